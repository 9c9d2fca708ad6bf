"""Build the psacak_amd native libraries in-tree.

* ``libpsacak_amd.so`` -- the HIP library (C ABI in ``include/psacak.h``),
  compiled by hipcc for gfx950 only.
* ``libpsk_gen.so`` -- the seeded workload generators (host C).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsacak_amd.so")
GEN = os.path.join(HERE, "libpsk_gen.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    # code object v5 loads on the ROCm 7.0 runtime bundled with PyTorch too
    "-mcode-object-version=5",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 instructions stay off in every HIP build of this project
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


TUNING = os.path.join(HERE, "tuning.json")


def tuned_defines():
    """Tile constants and input-keyed thresholds chosen by the autotuner
    (psacak_amd/tune.py), if any."""
    if not os.path.exists(TUNING):
        return []
    import json

    with open(TUNING) as f:
        rec = json.load(f)
    items = list(rec.get("defines", {}).items()) + list(rec.get("runtime", {}).items())
    return [f"{k}={int(v)}" for k, v in items]


def build_hip(verbose: bool = False, out_path: str = LIB, defines=None) -> str:
    """Compile the library.  defines (e.g. "PSK_RS_ROUNDS=16") override tile
    constants; by default the autotuner's choice (tuning.json) applies."""
    if defines is None:
        defines = tuned_defines()
    tmp = out_path + ".tmp"
    out = _run([HIPCC, *HIP_FLAGS, *[f"-D{d}" for d in defines], "-o", tmp,
                os.path.join(CSRC, "psacak.hip")])
    os.replace(tmp, out_path)
    if verbose and out.strip():
        print(out)
    return out_path


def build_gen() -> str:
    tmp = GEN + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-pthread", "-o", tmp,
          os.path.join(CSRC, "gen.c"), "-lm"])
    os.replace(tmp, GEN)
    return GEN


def build_all(verbose: bool = False) -> None:
    build_gen()
    build_hip(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", LIB, GEN)
