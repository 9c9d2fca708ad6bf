"""CPU oracle for the pSACAK path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker or as the timed
CPU baseline; the product package ``psacak_amd`` never imports it.

It wraps ``oracle/libpsk_oracle.so`` (built from ``psk_oracle.c`` and
``psk_par.c`` by ``oracle/Makefile``): a serial C restatement of
hucsmn/psacak 0.1.0 (``psacak8.rs``, ``psacak32.rs``, ``naming.rs``,
``common.rs``) and a multi-threaded one with the crate's pipelined induce
(``pipeline.rs``), the timed CPU baseline (file:line citations in the C
sources).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpsk_oracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build() -> str:
    """Compile the oracle library with the committed Makefile (always, not
    only when a source is newer than the library)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for name, args, res in [
            ("or_saca_naive8", [_u8p, ctypes.c_uint32, _u32p], None),
            ("or_saca_naive32", [_u32p, ctypes.c_uint32, _u32p], None),
            ("or_sacak8", [_u8p, ctypes.c_uint32, _u32p], ctypes.c_int),
            ("or_sacak32", [_u32p, ctypes.c_uint32, _u32p], ctypes.c_int),
            ("or_translate32", [_u32p, ctypes.c_uint32, _u32p], ctypes.c_int),
            ("or_check_sa8", [_u8p, ctypes.c_uint32, _u32p], ctypes.c_int),
            ("or_check_sa32", [_u32p, ctypes.c_uint32, _u32p], ctypes.c_int),
            ("or_types8", [_u8p, ctypes.c_uint32, _u8p], None),
            ("or_types32", [_u32p, ctypes.c_uint32, _u8p], None),
            ("or_buckets8", [_u8p, ctypes.c_uint32, _u32p, _u32p], None),
            ("or_name8", [_u8p, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p], ctypes.c_uint32),
            ("or_name32", [_u32p, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p], ctypes.c_uint32),
            ("or_stage1_8", [_u8p, ctypes.c_uint32, _u32p], ctypes.c_uint32),
            ("or_stage1_32", [_u32p, ctypes.c_uint32, _u32p], ctypes.c_uint32),
            ("or_level_stats8", [_u8p, ctypes.c_uint32, _u64p, ctypes.c_int], ctypes.c_int),
            ("or_par_sacak8", [_u8p, ctypes.c_uint32, _u32p], ctypes.c_int),
            ("or_par_sacak32", [_u32p, ctypes.c_uint32, _u32p], ctypes.c_int),
            ("or_par_threads", [], ctypes.c_int),
            ("or_par_set_threads", [ctypes.c_int], None),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _u8(text) -> np.ndarray:
    if isinstance(text, (bytes, bytearray)):
        return np.frombuffer(bytes(text), dtype=np.uint8)
    return np.ascontiguousarray(text, dtype=np.uint8)


def _u32(text) -> np.ndarray:
    return np.ascontiguousarray(text, dtype=np.uint32)


def saca_naive8(text) -> np.ndarray:
    t = _u8(text)
    sa = np.empty(len(t), dtype=np.uint32)
    lib().or_saca_naive8(_p(t, _u8p), len(t), _p(sa, _u32p))
    return sa


def saca_naive32(text) -> np.ndarray:
    t = _u32(text)
    sa = np.empty(len(t), dtype=np.uint32)
    lib().or_saca_naive32(_p(t, _u32p), len(t), _p(sa, _u32p))
    return sa


def sacak8(text) -> np.ndarray:
    t = _u8(text)
    sa = np.empty(len(t), dtype=np.uint32)
    rc = lib().or_sacak8(_p(t, _u8p), len(t), _p(sa, _u32p))
    if rc != 0:
        raise ValueError(f"or_sacak8 failed: {rc}")
    return sa


def par_sacak8(text) -> np.ndarray:
    """psacak8 with the pipelined parallel induce on OpenMP threads
    (OMP_NUM_THREADS); same SA as :func:`sacak8`."""
    t = _u8(text)
    sa = np.empty(len(t), dtype=np.uint32)
    rc = lib().or_par_sacak8(_p(t, _u8p), len(t), _p(sa, _u32p))
    if rc != 0:
        raise ValueError(f"or_par_sacak8 failed: {rc}")
    return sa


def par_sacak32_raw(text) -> np.ndarray:
    """:func:`sacak32_raw` with the multi-threaded psacak32."""
    t = _u32(text)
    tr = translate32(t)
    if tr is None:
        return saca_naive32(t)
    sa = np.empty(len(t), dtype=np.uint32)
    rc = lib().or_par_sacak32(_p(tr, _u32p), len(tr), _p(sa, _u32p))
    if rc != 0:
        raise ValueError(f"or_par_sacak32 failed: {rc}")
    return sa


def par_threads() -> int:
    return int(lib().or_par_threads())


def set_par_threads(p: int) -> None:
    lib().or_par_set_threads(int(p))


def sacak32(text) -> np.ndarray:
    """psacak32 on a bucket-pointer text (psacak32.rs:22 precondition)."""
    t = _u32(text)
    sa = np.empty(len(t), dtype=np.uint32)
    rc = lib().or_sacak32(_p(t, _u32p), len(t), _p(sa, _u32p))
    if rc != 0:
        raise ValueError(f"or_sacak32 failed: {rc}")
    return sa


def translate32(text):
    """translate_sample (psacak32.rs:912-943); None if all symbols distinct."""
    t = _u32(text)
    out = np.empty(len(t), dtype=np.uint32)
    ok = lib().or_translate32(_p(t, _u32p), len(t), _p(out, _u32p))
    return out if ok else None


def sacak32_raw(text) -> np.ndarray:
    """SA of an arbitrary u32 text: translate then psacak32, or the value sort
    when every symbol is distinct (psacak32.rs:873-881)."""
    t = _u32(text)
    tr = translate32(t)
    if tr is None:
        return saca_naive32(t)
    return sacak32(tr)


def check_sa8(text, sa) -> int:
    t = _u8(text)
    s = _u32(sa)
    if len(s) != len(t):
        return 1
    return lib().or_check_sa8(_p(t, _u8p), len(t), _p(s, _u32p))


def check_sa32(text, sa) -> int:
    t = _u32(text)
    s = _u32(sa)
    if len(s) != len(t):
        return 1
    return lib().or_check_sa32(_p(t, _u32p), len(t), _p(s, _u32p))


def types8(text) -> np.ndarray:
    t = _u8(text)
    out = np.zeros(len(t), dtype=np.uint8)
    lib().or_types8(_p(t, _u8p), len(t), _p(out, _u8p))
    return out


def types32(text) -> np.ndarray:
    t = _u32(text)
    out = np.zeros(len(t), dtype=np.uint8)
    lib().or_types32(_p(t, _u32p), len(t), _p(out, _u8p))
    return out


def buckets8(text):
    t = _u8(text)
    bounds = np.zeros(257, dtype=np.uint32)
    lms = np.zeros(256, dtype=np.uint32)
    lib().or_buckets8(_p(t, _u8p), len(t), _p(bounds, _u32p), _p(lms, _u32p))
    return bounds, lms


def stage1_8(text) -> np.ndarray:
    t = _u8(text)
    out = np.zeros(max(len(t) // 2, 1), dtype=np.uint32)
    m1 = lib().or_stage1_8(_p(t, _u8p), len(t), _p(out, _u32p))
    return out[:m1].copy()


def stage1_32(text) -> np.ndarray:
    t = _u32(text)
    out = np.zeros(max(len(t) // 2, 1), dtype=np.uint32)
    m1 = lib().or_stage1_32(_p(t, _u32p), len(t), _p(out, _u32p))
    return out[:m1].copy()


def name8(text, lms_sorted):
    t = _u8(text)
    l = _u32(lms_sorted)
    red = np.zeros(max(len(l), 1), dtype=np.uint32)
    k = lib().or_name8(_p(t, _u8p), len(t), _p(l, _u32p), len(l), _p(red, _u32p))
    return int(k), red[: len(l)].copy()


def name32(text, lms_sorted):
    t = _u32(text)
    l = _u32(lms_sorted)
    red = np.zeros(max(len(l), 1), dtype=np.uint32)
    k = lib().or_name32(_p(t, _u32p), len(t), _p(l, _u32p), len(l), _p(red, _u32p))
    return int(k), red[: len(l)].copy()


def level_stats8(text, max_levels: int = 64):
    """Per-level (m, m1, mL, k, sum of LMS-substring lengths) of the recursion."""
    t = _u8(text)
    st = np.zeros(5 * max_levels, dtype=np.uint64)
    lv = lib().or_level_stats8(_p(t, _u8p), len(t), _p(st, _u64p), max_levels)
    return st[: 5 * lv].reshape(lv, 5)
