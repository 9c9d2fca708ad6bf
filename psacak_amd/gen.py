"""Seeded synthetic workloads (ctypes over ``libpsk_gen.so``, built from
``csrc/gen.c``).  Each generator is a pure function of its arguments.

* :func:`iid_dna` -- i.i.d. uniform over ``ACGT`` (bytes 65, 67, 71, 84);
* :func:`markov_dna` -- order-3 Markov chain over ACGT with Dirichlet(0.5)
  transition rows, then repeats (1 KiB-100 KiB copies with 1% substitutions)
  until ``repeat_frac`` of the text is covered;
* :func:`zipf32` -- Zipf(s) symbols over ``sigma`` ranks mapped through a
  seeded permutation, then exact repeats of length 16-4096 covering
  ``repeat_frac``;
* :func:`uniform8` -- i.i.d. bytes in ``[0, alphabet)``;
* :func:`lengths` -- log-uniform text lengths; :func:`requests` -- Poisson
  request stream (gaps in ms and lengths).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_GEN_PATH = os.path.join(_HERE, "libpsk_gen.so")
_g = None


def _lib():
    global _g
    if _g is None:
        if not os.path.exists(_GEN_PATH):
            from . import build
            build.build_gen()
        g = ctypes.CDLL(_GEN_PATH)
        u64, vp, i32, dbl, u32 = (ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_uint32)
        g.psk_gen_iid_dna.argtypes = [u64, u64, vp, i32]
        g.psk_gen_uniform8.argtypes = [u64, u64, u32, vp, i32]
        g.psk_gen_markov_dna.argtypes = [u64, u64, dbl, vp, i32]
        g.psk_gen_zipf32.argtypes = [u64, u64, u32, dbl, dbl, vp, i32]
        g.psk_gen_lengths.argtypes = [u64, u64, u64, u64, vp]
        g.psk_gen_requests.argtypes = [u64, u64, dbl, u64, u64, dbl, u64, u64, vp, vp]
        for f in (g.psk_gen_iid_dna, g.psk_gen_uniform8, g.psk_gen_markov_dna,
                  g.psk_gen_zipf32, g.psk_gen_lengths, g.psk_gen_requests):
            f.restype = None
        _g = g
    return _g


def _threads(threads):
    if threads:
        return threads
    return min(16, os.cpu_count() or 1)


def iid_dna(n: int, seed: int = 1, threads: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _lib().psk_gen_iid_dna(seed, n, out.ctypes.data, _threads(threads))
    return out


def uniform8(n: int, alphabet: int = 256, seed: int = 1, threads: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _lib().psk_gen_uniform8(seed, n, alphabet, out.ctypes.data, _threads(threads))
    return out


def markov_dna(n: int, seed: int = 2, repeat_frac: float = 0.10, threads: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    _lib().psk_gen_markov_dna(seed, n, repeat_frac, out.ctypes.data, _threads(threads))
    return out


def zipf32(n: int, sigma: int = 1 << 20, s: float = 1.1, repeat_frac: float = 0.20,
           seed: int = 3, threads: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    _lib().psk_gen_zipf32(seed, n, sigma, s, repeat_frac, out.ctypes.data, _threads(threads))
    return out


def lengths(count: int, lo: int, hi: int, seed: int = 4) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    _lib().psk_gen_lengths(seed, count, lo, hi, out.ctypes.data)
    return out


def requests(count: int, mean_ms: float = 5.0, lo: int = 10 << 10, hi: int = 64 << 20,
             big_p: float = 0.01, big_lo: int = 1 << 30, big_hi: int = 2 << 30, seed: int = 5):
    gaps = np.empty(count, dtype=np.float64)
    lens = np.empty(count, dtype=np.uint64)
    _lib().psk_gen_requests(seed, count, mean_ms, lo, hi, big_p, big_lo, big_hi,
                            gaps.ctypes.data, lens.ctypes.data)
    return gaps, lens
