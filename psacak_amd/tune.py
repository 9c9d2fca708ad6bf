"""Autotuner for the tile shapes of psacak_amd (run on the GPU machine).

The kernels' tile and workgroup shapes are build-time constants:

* ``PSK_RS_ROUNDS`` -- rounds of 256 pairs per radix-sort tile (tile = 256 x rounds);
* ``PSK_RS_CT``     -- radix-count tiles per workgroup;
* ``PSK_MS_ROUNDS`` -- rounds of 256 items per multi-split tile (level-0 induce).

The tuner searches them one coordinate at a time: for each candidate value it
builds the library with hipcc (``build.build_hip(defines=...)``) into a
scratch directory, times a real psacak8 call on a seeded DNA text in a fresh
process that loads that build (``PSACAK_AMD_LIB``), checks the SA there with
the device checker, and keeps the fastest.

A second phase tunes choices keyed by the input, with the built library and
its context switches (no rebuild): for each text size class (below and above
64 MiB) it times the level-0 final induce both ways -- bucket steps and chain
sorts (``Context.set_chain_induce``) -- on seeded random texts over a range of
alphabet sizes, and takes as the class's threshold the smallest alphabet from
which the chain sorts are faster at every larger one measured
(``PSK_CHAIN_SIGMA``, ``PSK_CHAIN_SIGMA_SMALL``).

The same phase times the prefix-doubling threshold (the share of distinct
symbols from which a recursion level is sorted by prefix doubling instead of
recursing, ``Context.set_doubling``) on Markov DNA and random texts over 12,
24 and 64 symbols, and keeps the candidate with the least total time
(``PSK_DOUBLING``).

The result is written to ``psacak_amd/tuning.json`` ("defines": tile shapes,
"runtime": thresholds), which ``build.py`` applies to every later build.

    python -m psacak_amd.tune [--n BYTES] [--reps K] [--out FILE] [--runtime-only]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TUNING = os.path.join(HERE, "tuning.json")

SPACE = {
    "PSK_RS_ROUNDS": [8, 4, 16],
    "PSK_RS_CT": [1, 2, 4],
    "PSK_MS_ROUNDS": [16, 8, 32],
}

# Timed in a child process: the library under test is chosen by PSACAK_AMD_LIB
# before the package is imported.
_CHILD = r"""
import json, sys, time
sys.path.insert(0, {root!r})
import numpy as np, torch
import psacak_amd as P
from psacak_amd import gen
n, reps = {n}, {reps}
t = gen.markov_dna(n, seed=20260101, repeat_frac=0.10)
dt = torch.from_numpy(t).to("cuda:0")
ds = torch.empty(n, dtype=torch.int32, device="cuda:0")
ctx = P.Context()
P.psacak8_dev(dt, ds, n, ctx=ctx)
torch.cuda.synchronize()
ms = []
for _ in range(reps):
    P.psacak8_dev(dt, ds, n, ctx=ctx)
    ms.append(float(ctx.last_timing_ms()[0]))
ok = P.check8_dev(dt, ds, n, ctx=ctx)
print(json.dumps({{"ms": min(ms), "all_ms": ms, "ok": bool(ok)}}))
"""


# The alphabet sweep of one size class: both final-induce paths per alphabet.
_CHILD_SIGMA = r"""
import json, sys
sys.path.insert(0, {root!r})
import numpy as np, torch
import psacak_amd as P
from psacak_amd import gen
n, reps, sigmas = {n}, {reps}, {sigmas!r}
ds = torch.empty(n, dtype=torch.int32, device="cuda:0")
rows = []
for sig in sigmas:
    t = gen.uniform8(n, alphabet=sig, seed=7100 + sig)
    dt = torch.from_numpy(t).to("cuda:0")
    row = {{"sigma": sig}}
    for name, thr in (("buckets", 0), ("chain", 1)):
        with P.Context() as ctx:
            ctx.set_chain_induce(thr)
            P.psacak8_dev(dt, ds, n, ctx=ctx)
            ms = []
            for _ in range(reps):
                P.psacak8_dev(dt, ds, n, ctx=ctx)
                ms.append(float(ctx.last_timing_ms()[0]))
            row[name] = min(ms)
            row[name + "_ok"] = bool(P.check8_dev(dt, ds, n, ctx=ctx))
    rows.append(row)
    print(json.dumps(row), file=sys.stderr, flush=True)
print(json.dumps(rows))
"""

SIGMAS = [4, 8, 12, 16, 24, 32, 48, 64, 80, 96, 128, 160, 192, 256]

# The doubling sweep: every candidate threshold on every text, alternating.
_CHILD_DOUBLING = r"""
import json, sys
sys.path.insert(0, {root!r})
import numpy as np, torch
import psacak_amd as P
from psacak_amd import gen
n, reps, cands = {n}, {reps}, {cands!r}
texts = {{"markov": gen.markov_dna(n, seed=20260101, repeat_frac=0.10)}}
for s in (12, 24, 64):
    texts["k%d" % s] = gen.uniform8(n, alphabet=s, seed=7100 + s)
ds = torch.empty(n, dtype=torch.int32, device="cuda:0")
rows = []
for name, t in texts.items():
    dt = torch.from_numpy(t).to("cuda:0")
    row = {{"text": name}}
    for thr in cands + cands:
        with P.Context() as ctx:
            ctx.set_doubling(thr)
            P.psacak8_dev(dt, ds, n, ctx=ctx)
            ms = []
            for _ in range(reps):
                P.psacak8_dev(dt, ds, n, ctx=ctx)
                ms.append(float(ctx.last_timing_ms()[0]))
            ok = bool(P.check8_dev(dt, ds, n, ctx=ctx))
        key = str(thr)
        row[key] = min(row.get(key, 1e30), min(ms))
        row[key + "_ok"] = row.get(key + "_ok", True) and ok
    rows.append(row)
    print(json.dumps(row), file=sys.stderr, flush=True)
print(json.dumps(rows))
"""

DOUBLING = [16, 32, 64, 128, 256]


def tune_doubling(reps: int, n: int = 1 << 28, log=print, lib: str | None = None) -> tuple:
    env = dict(os.environ)
    if lib:
        env["PSACAK_AMD_LIB"] = lib
    code = _CHILD_DOUBLING.format(root=ROOT, n=n, reps=reps, cands=DOUBLING)
    r = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE,
                       text=True, timeout=1800)
    if r.returncode != 0:
        raise RuntimeError(f"doubling sweep failed ({r.returncode})")
    rows = json.loads(r.stdout.strip().splitlines()[-1])
    if not all(x[str(c) + "_ok"] for x in rows for c in DOUBLING):
        raise RuntimeError(f"an SA of the doubling sweep failed the device check: {rows}")
    total = {c: sum(x[str(c)] for x in rows) for c in DOUBLING}
    for c in DOUBLING:
        log(f"[tune] doubling {c}/1024: " + ", ".join(f"{x['text']} {x[str(c)]:.2f}" for x in rows)
            + f" ms; total {total[c]:.2f} ms")
    best = min(DOUBLING, key=lambda c: total[c])
    log(f"[tune] PSK_DOUBLING = {best}")
    return best, {"n": n, "rows": rows}
SIZE_CLASSES = {"PSK_CHAIN_SIGMA_SMALL": 1 << 24, "PSK_CHAIN_SIGMA": 1 << 28}


def _threshold(rows) -> int:
    """Smallest alphabet from which the chain sorts beat the bucket steps at
    every larger alphabet measured (256 + 1 = never, i.e. 0)."""
    thr = 0
    for r in sorted(rows, key=lambda r: -r["sigma"]):
        if r["chain"] < r["buckets"]:
            thr = r["sigma"]
        else:
            break
    return thr


def tune_runtime(reps: int, log=print, lib: str | None = None) -> dict:
    out = {}
    sweeps = {}
    for name, n in SIZE_CLASSES.items():
        env = dict(os.environ)
        if lib:
            env["PSACAK_AMD_LIB"] = lib
        code = _CHILD_SIGMA.format(root=ROOT, n=n, reps=reps, sigmas=SIGMAS)
        # the child's per-alphabet lines go straight to stderr (progress)
        r = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE,
                           text=True, timeout=1200)
        if r.returncode != 0:
            raise RuntimeError(f"sweep failed ({r.returncode})")
        rows = json.loads(r.stdout.strip().splitlines()[-1])
        if not all(x["buckets_ok"] and x["chain_ok"] for x in rows):
            raise RuntimeError(f"an SA of the alphabet sweep failed the device check: {rows}")
        for x in rows:
            log(f"[tune] n={n} sigma={x['sigma']}: bucket steps {x['buckets']:.3f} ms, "
                f"chain sorts {x['chain']:.3f} ms")
        out[name] = _threshold(rows)
        sweeps[name] = {"n": n, "rows": rows}
        log(f"[tune] {name} = {out[name]}")
    out["PSK_DOUBLING"], sweeps["PSK_DOUBLING"] = tune_doubling(reps, log=log, lib=lib)
    return {"runtime": out, "sweeps": sweeps}


def _measure(lib: str, n: int, reps: int) -> dict:
    env = dict(os.environ, PSACAK_AMD_LIB=lib)
    code = _CHILD.format(root=ROOT, n=n, reps=reps)
    r = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"timing run failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def tune(n: int, reps: int, out: str, log=print) -> dict:
    from psacak_amd import build

    best = {k: v[0] for k, v in SPACE.items()}
    tried = {}
    with tempfile.TemporaryDirectory() as tmp:
        def run(cfg):
            key = tuple(sorted(cfg.items()))
            if key in tried:
                return tried[key]
            lib = os.path.join(tmp, f"lib{len(tried)}.so")
            t0 = time.time()
            build.build_hip(out_path=lib, defines=[f"{k}={v}" for k, v in cfg.items()])
            res = _measure(lib, n, reps)
            if not res["ok"]:
                raise RuntimeError(f"candidate {cfg} produced an invalid SA")
            tried[key] = res["ms"]
            log(f"[tune] {cfg}: {res['ms']:.3f} ms (build+run {time.time() - t0:.0f} s)")
            return res["ms"]

        cur = run(best)
        for name, values in SPACE.items():
            for v in values:
                if v == best[name]:
                    continue
                cfg = dict(best, **{name: v})
                ms = run(cfg)
                if ms < cur * 0.99:  # keep a change only for a clear gain
                    best, cur = cfg, ms
    rec = {
        "defines": best,
        "ms": cur,
        "workload": f"psacak8 on {n} bytes of seeded Markov DNA (bench generator), best of {reps}",
        "candidates": [{"defines": dict(k), "ms": v} for k, v in tried.items()],
    }
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    log(f"[tune] best {best}: {cur:.3f} ms -> {out}")
    return rec


def merge_runtime(rt: dict, out: str) -> dict:
    """Adds the input-keyed thresholds to the record at out (kept otherwise)."""
    rec = {}
    if os.path.exists(out):
        with open(out) as f:
            rec = json.load(f)
    elif os.path.exists(TUNING):
        with open(TUNING) as f:
            rec = json.load(f)
    rec["runtime"] = rt["runtime"]
    rec["runtime_sweeps"] = rt["sweeps"]
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    return rec


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=1 << 28, help="text bytes of the timed call")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=TUNING)
    ap.add_argument("--runtime-only", action="store_true",
                    help="only the input-keyed thresholds, with the current build")
    a = ap.parse_args()
    if not a.runtime_only:
        tune(a.n, a.reps, a.out)
    merge_runtime(tune_runtime(a.reps), a.out)


if __name__ == "__main__":
    main()
