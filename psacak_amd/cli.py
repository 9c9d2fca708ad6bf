"""Command-line counterpart of the crate's bench binary (src/bin/psacak.rs).

    python -m psacak_amd.cli [-t TIMES] [-c] INPUT

Reads INPUT, computes its suffix array TIMES times on the GPU (host buffers
in and out, as psacak_inplace does, bin/psacak.rs:40-46), and prints to stderr
the same lines the crate prints (bin/psacak.rs:38,46,48,50), which
bench/compare.sh parses:

    load N bytes from `INPUT`
     time: 0.123s 0.120s
      rss: 123.456MiB
    check: true

``-c`` runs an independent check: the O(n) Burkhardt-Kaerkkaeinen condition on
the device (permutation + adjacent-pair ranks), and a direct comparison of
sampled adjacent suffix pairs on the host.
"""
from __future__ import annotations

import argparse
import resource
import sys
import time

import numpy as np


def _sampled_pairs_ok(text: bytes, sa: np.ndarray, samples: int = 2048) -> bool:
    n = len(text)
    if n < 2:
        return True
    rng = np.random.default_rng(12345)
    for i in rng.integers(1, n, size=min(samples, n - 1)):
        a, b = int(sa[i - 1]), int(sa[i])
        if not text[a:] < text[b:]:
            return False
    return True


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="psacak", description="pSACAK suffix sorting test")
    ap.add_argument("-t", "--times", type=int, default=1, help="repeat multiple times")
    ap.add_argument("-c", "--check", action="store_true", help="check the suffix array")
    ap.add_argument("input", help="the data to sort")
    a = ap.parse_args(argv)

    import psacak_amd as P

    try:
        with open(a.input, "rb") as f:
            text = f.read()
    except OSError as e:
        print(f"error: {e!r}", file=sys.stderr)
        return 1
    print(f"load {len(text)} bytes from `{a.input}`", file=sys.stderr)
    suf = np.zeros(len(text), dtype=np.uint32)
    times = []
    ctx = P.Context()
    for _ in range(max(a.times, 1)):
        t0 = time.perf_counter()
        P.psacak_inplace(text, suf, ctx)
        times.append(time.perf_counter() - t0)
    print(" time: " + "".join(f"{t:.3f}s " for t in times), file=sys.stderr)
    rss_kib = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
    print(f"  rss: {rss_kib / 1024.0:.3f}MiB", file=sys.stderr)
    if a.check:
        import torch

        dt = torch.frombuffer(bytearray(text), dtype=torch.uint8).to("cuda:0") if text else None
        ok = True
        if text:
            ds = torch.from_numpy(suf.view(np.int32)).to("cuda:0")
            ok = P.check8_dev(dt, ds, len(text), ctx=ctx)
        ok = ok and _sampled_pairs_ok(text, suf)
        print(f"check: {str(ok).lower()}", file=sys.stderr)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
