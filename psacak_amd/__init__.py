"""psacak_amd -- pSACAK suffix-array construction on AMD MI355X (gfx950).

A drop-in for the public API of the hucsmn/psacak crate (lib.rs):

* :func:`psacak` (``lib.rs:64-69``) -- SA of a byte text, as a new u32 array;
* :func:`psacak_inplace` (``lib.rs:77-82``) -- SA into a caller's u32 buffer;
* :func:`psacak32` (``psacak32.rs:24``) -- SA of a u32 text in bucket-pointer
  form, and :func:`psacak32_raw` for an arbitrary u32 text (translated as in
  ``psacak32.rs:912-943``).

The SA excludes the virtual sentinel (``lib.rs:30-43``).  Every call runs the
HIP kernels of ``libpsacak_amd.so`` through its C ABI (``include/psacak.h``);
there is no CPU fallback: if the library is missing the import fails.  The
crate panics on bad arguments; here they raise :class:`PskError` subclasses
before any GPU work.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "PskError", "Context", "psacak", "psacak_inplace", "psacak32", "psacak32_raw",
    "psacak8_dev", "psacak32_dev", "psacak32_raw_dev", "check8_dev", "check32_dev", "batch8",
    "batch8_dev",
    "workspace_bytes8", "workspace_bytes32", "lib_path", "version", "kernel_profile",
    "kernel_stats", "Server", "psacak8_dev_noalloc",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
# PSACAK_AMD_LIB selects another build of the library (the autotuner times
# candidate builds this way); by default the in-tree build
_LIB_PATH = os.environ.get("PSACAK_AMD_LIB") or os.path.join(_HERE, "libpsacak_amd.so")

PSK_OK, PSK_EINVAL, PSK_ETOOLONG, PSK_EPRECOND = 0, -1, -2, -3
PSK_ENOMEM, PSK_EBUDGET, PSK_EDEVICE, PSK_EBUSY, PSK_ENOTFOUND = -4, -5, -6, -7, -8


class PskError(RuntimeError):
    """A non-zero status from the C ABI; ``.code`` is the PSK_* value."""

    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{_strerror(code)} ({code})" + (f": {msg}" if msg else ""))


def _load():
    # PyTorch-ROCm ships its own HIP runtime (libamdhip64, SONAME
    # libamdhip64.so.7) and loads it by file name.  If this library's
    # dependency on the system runtime were resolved first, torch would load
    # a second runtime into the process and find no GPU.  Importing torch
    # first makes both share torch's runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(_LIB_PATH):
        raise ImportError(
            f"psacak_amd: native library {_LIB_PATH} is missing; build it with "
            "`python -m psacak_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(_LIB_PATH)
    vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    sig = {
        "psk_strerror": ([i32], ctypes.c_char_p),
        "psk_version": ([], ctypes.c_char_p),
        "psk_ctx_create": ([ctypes.POINTER(vp), u64], i32),
        "psk_ctx_destroy": ([vp], None),
        "psk_last_error": ([vp], ctypes.c_char_p),
        "psk_ctx_trim": ([vp], i32),
        "psk_ctx_workspace": ([vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)], i32),
        "psk_ctx_set_direct_rounds": ([vp, i32], i32),
        "psk_ctx_set_dict_naming": ([vp, i32], i32),
        "psk_ctx_set_doubling": ([vp, i32], i32),
        "psk_ctx_set_chain_induce": ([vp, i32], i32),
        "psk_workspace_bytes8": ([u64], u64),
        "psk_workspace_bytes32": ([u64], u64),
        "psk_psacak_inplace": ([vp, vp, u64, vp, u64], i32),
        "psk_psacak": ([vp, vp, u64, vp], i32),
        "psk_psacak32": ([vp, vp, u64, vp, u64], i32),
        "psk_psacak32_raw": ([vp, vp, u64, vp, u64], i32),
        "psk_psacak8_dev": ([vp, vp, u64, vp, vp], i32),
        "psk_psacak32_dev": ([vp, vp, u64, vp, vp], i32),
        "psk_psacak32_raw_dev": ([vp, vp, u64, vp, vp], i32),
        "psk_last_levels": ([vp, vp, i32], i32),
        "psk_last_timing": ([vp, vp, i32], i32),
        "psk_batch8": ([vp, vp, vp, vp, u64], i32),
        "psk_batch8_dev": ([vp, vp, vp, u64, vp, vp], i32),
        "psk_check8_dev": ([vp, vp, u64, vp, ctypes.POINTER(i32), vp], i32),
        "psk_profile": ([vp, i32], i32),
        "psk_profile_only": ([vp, ctypes.c_char_p], i32),
        "psk_server_create": ([ctypes.POINTER(vp), u64, i32], i32),
        "psk_server_footprint": ([u64, i32], u64),
        "psk_server_submit": ([vp, vp, u64, ctypes.POINTER(u64)], i32),
        "psk_server_submit_into": ([vp, vp, u64, vp, ctypes.POINTER(u64)], i32),
        "psk_server_poll": ([vp, u64, ctypes.POINTER(i32), vp, ctypes.POINTER(ctypes.c_double)], i32),
        "psk_server_wait": ([vp, u64, vp, ctypes.POINTER(ctypes.c_double), ctypes.c_double], i32),
        "psk_server_peak_bytes": ([vp], u64),
        "psk_server_pending": ([vp], u64),
        "psk_server_stats": ([vp, vp, i32], i32),
        "psk_server_destroy": ([vp], None),
        "psk_profile_stats": ([vp, ctypes.c_char_p, u64, vp, vp, vp, i32], i32),
        "psk_psacak8_dev_noalloc": ([vp, u64, vp, vp], i32),
        "psk_check32_dev": ([vp, vp, u64, vp, ctypes.POINTER(i32), vp], i32),
        "psk_ctx_set_capture": ([vp, i32], i32),
        "psk_last_capture": ([vp, vp, u64, ctypes.POINTER(u64)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


class _Lib:
    """Loads libpsacak_amd.so on first use (so ``psacak_amd.build`` can run
    before the library exists); a missing library raises ImportError."""

    _handle = None

    def __getattr__(self, name):
        if _Lib._handle is None:
            _Lib._handle = _load()
        return getattr(_Lib._handle, name)


_lib = _Lib()


def lib_path() -> str:
    return _LIB_PATH


def version() -> str:
    return _lib.psk_version().decode()


def _strerror(code: int) -> str:
    return _lib.psk_strerror(code).decode()


def workspace_bytes8(n: int) -> int:
    return int(_lib.psk_workspace_bytes8(n))


def workspace_bytes32(n: int) -> int:
    return int(_lib.psk_workspace_bytes32(n))


class Context:
    """A device context: one HIP stream plus cached device workspace.

    ``hbm_budget`` (bytes, 0 = none) bounds the workspace it may reserve.
    """

    def __init__(self, hbm_budget: int = 0):
        h = ctypes.c_void_p()
        rc = _lib.psk_ctx_create(ctypes.byref(h), hbm_budget)
        if rc != PSK_OK:
            raise PskError(rc)
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _lib.psk_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_error(self) -> str:
        return _lib.psk_last_error(self._h).decode()

    def trim(self):
        self._check(_lib.psk_ctx_trim(self._h))

    def workspace(self):
        """(reserved, peak) device workspace of this context, in bytes."""
        r, p = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._check(_lib.psk_ctx_workspace(self._h, ctypes.byref(r), ctypes.byref(p)))
        return int(r.value), int(p.value)

    def set_direct_rounds(self, rounds: int):
        """Rounds of direct LMS-substring sorting before induced sorting takes
        over; 0 = always induced sorting.  Results are identical either way."""
        self._check(_lib.psk_ctx_set_direct_rounds(self._h, int(rounds)))

    def set_doubling(self, distinct_per_1024: int):
        """Recursion levels whose text has at least distinct_per_1024/1024
        distinct symbols are sorted by prefix doubling (0 = always recurse).
        Results are identical either way."""
        self._check(_lib.psk_ctx_set_doubling(self._h, int(distinct_per_1024)))

    def set_chain_induce(self, min_distinct_bytes: int):
        """Byte texts with at least this many distinct bytes take their final
        induce as chain sorts instead of bucket steps (0 = always bucket
        steps).  Results are identical either way."""
        self._check(_lib.psk_ctx_set_chain_induce(self._h, int(min_distinct_bytes)))

    def set_dict_naming(self, on: bool):
        """Round 1 of direct naming at level 0 through a dictionary of the
        distinct LMS-substring keys (True, default) or by sorting every key
        (False).  Results are identical either way."""
        self._check(_lib.psk_ctx_set_dict_naming(self._h, 1 if on else 0))

    def last_levels(self):
        """Per-level rows (m, m1, mL, k, sum of LMS-substring lengths, symbol bytes)."""
        rows = np.zeros(6 * 64, dtype=np.uint64)
        n = _lib.psk_last_levels(self._h, rows.ctypes.data, 64)
        return rows[: 6 * max(n, 0)].reshape(-1, 6)

    def set_capture(self, level: int):
        """Capture the translated reduced text of recursion level ``level`` in
        the following calls (-1: off); read it with :meth:`last_capture`."""
        self._check(_lib.psk_ctx_set_capture(self._h, int(level)))

    def last_capture(self):
        """The captured text of the last call (uint32 array), or None if the
        last call did not name the captured level."""
        cnt = ctypes.c_uint64(0)
        rc = _lib.psk_last_capture(self._h, None, 0, ctypes.byref(cnt))
        if rc == PSK_ENOTFOUND:
            return None
        self._check(rc)
        out = np.empty(cnt.value, dtype=np.uint32)
        self._check(_lib.psk_last_capture(self._h, out.ctypes.data, len(out), ctypes.byref(cnt)))
        return out

    def last_timing_ms(self):
        ms = np.zeros(4, dtype=np.float64)
        _lib.psk_last_timing(self._h, ms.ctypes.data, 4)
        return ms

    def _check(self, rc: int):
        if rc != PSK_OK:
            raise PskError(rc, self.last_error())


def _ctx_handle(ctx):
    return ctx.handle if ctx is not None else None


def _raise(rc: int, ctx=None):
    if rc != PSK_OK:
        msg = ctx.last_error() if ctx is not None else ""
        raise PskError(rc, msg)


def _as_u8(text) -> np.ndarray:
    if isinstance(text, (bytes, bytearray, memoryview)):
        return np.frombuffer(bytes(text), dtype=np.uint8)
    a = np.asarray(text)
    if a.dtype != np.uint8:
        raise TypeError("psacak expects a byte text (uint8)")
    return np.ascontiguousarray(a)


def psacak(text, ctx: Context | None = None) -> np.ndarray:
    """SA of a byte text (``lib.rs:64-69``); the sentinel is not included."""
    t = _as_u8(text)
    sa = np.empty(len(t), dtype=np.uint32)
    rc = _lib.psk_psacak(_ctx_handle(ctx), t.ctypes.data, len(t), sa.ctypes.data)
    _raise(rc, ctx)
    return sa


def psacak_inplace(text, suf: np.ndarray, ctx: Context | None = None) -> None:
    """SA of ``text`` into ``suf[:len(text)]`` (``lib.rs:77-82``).

    ``suf`` must be a writable, C-contiguous uint32 array at least as long as
    the text; entries past ``len(text)`` are not touched."""
    t = _as_u8(text)
    if not isinstance(suf, np.ndarray) or suf.dtype != np.uint32 or not suf.flags.c_contiguous:
        raise TypeError("suf must be a C-contiguous numpy uint32 array")
    rc = _lib.psk_psacak_inplace(_ctx_handle(ctx), t.ctypes.data, len(t), suf.ctypes.data,
                                 len(suf))
    _raise(rc, ctx)


def psacak32(text, ctx: Context | None = None) -> np.ndarray:
    """SA of a u32 text in bucket-pointer form (``psacak32.rs:22-24``)."""
    t = np.ascontiguousarray(text, dtype=np.uint32)
    sa = np.empty(len(t), dtype=np.uint32)
    rc = _lib.psk_psacak32(_ctx_handle(ctx), t.ctypes.data, len(t), sa.ctypes.data, len(sa))
    _raise(rc, ctx)
    return sa


def psacak32_raw(text, ctx: Context | None = None) -> np.ndarray:
    """SA of an arbitrary u32 text (translated first, ``psacak32.rs:912-943``)."""
    t = np.ascontiguousarray(text, dtype=np.uint32)
    sa = np.empty(len(t), dtype=np.uint32)
    rc = _lib.psk_psacak32_raw(_ctx_handle(ctx), t.ctypes.data, len(t), sa.ctypes.data, len(sa))
    _raise(rc, ctx)
    return sa


def _dev_ptr(x):
    """Device pointer of a torch tensor (or an int address)."""
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream  # torch.cuda.Stream


def psacak8_dev(d_text, d_sa, n: int | None = None, ctx: Context | None = None, stream=None):
    """SA of a device byte text into a device u32 buffer (torch tensors or
    raw device addresses); returns when the SA is complete."""
    if n is None:
        n = d_text.numel()
    rc = _lib.psk_psacak8_dev(_ctx_handle(ctx), _dev_ptr(d_text), n, _dev_ptr(d_sa),
                              _stream_ptr(stream))
    _raise(rc, ctx)


def psacak32_dev(d_text, d_sa, n: int | None = None, ctx: Context | None = None, stream=None):
    if n is None:
        n = d_text.numel()
    rc = _lib.psk_psacak32_dev(_ctx_handle(ctx), _dev_ptr(d_text), n, _dev_ptr(d_sa),
                               _stream_ptr(stream))
    _raise(rc, ctx)


def psacak32_raw_dev(d_text, d_sa, n: int | None = None, ctx: Context | None = None,
                     stream=None):
    if n is None:
        n = d_text.numel()
    rc = _lib.psk_psacak32_raw_dev(_ctx_handle(ctx), _dev_ptr(d_text), n, _dev_ptr(d_sa),
                                   _stream_ptr(stream))
    _raise(rc, ctx)


# psk_psacak8_dev_noalloc limits (include/psacak.h)
NOALLOC_SCRATCH_BYTES = 256 << 20
NOALLOC_MAX_N = 2 << 20


def psacak8_dev_noalloc(d_text, d_sa, n: int | None = None, stream=None):
    """psacak8_dev without a context and without any device allocation: the
    workspace comes from the library's fixed per-GPU scratch array (texts of
    at most NOALLOC_MAX_N bytes)."""
    if n is None:
        n = d_text.numel()
    rc = _lib.psk_psacak8_dev_noalloc(_dev_ptr(d_text), n, _dev_ptr(d_sa), _stream_ptr(stream))
    _raise(rc)


def check8_dev(d_text, d_sa, n: int | None = None, ctx: Context | None = None,
               stream=None) -> bool:
    """Device-side SA check (permutation + adjacent-pair rank condition)."""
    if n is None:
        n = d_text.numel()
    ok = ctypes.c_int(0)
    rc = _lib.psk_check8_dev(_ctx_handle(ctx), _dev_ptr(d_text), n, _dev_ptr(d_sa),
                             ctypes.byref(ok), _stream_ptr(stream))
    _raise(rc, ctx)
    return bool(ok.value)


def check32_dev(d_text, d_sa, n: int | None = None, ctx: Context | None = None,
                stream=None) -> bool:
    """Device-side SA check of an integer (u32) text."""
    if n is None:
        n = d_text.numel()
    ok = ctypes.c_int(0)
    rc = _lib.psk_check32_dev(_ctx_handle(ctx), _dev_ptr(d_text), n, _dev_ptr(d_sa),
                              ctypes.byref(ok), _stream_ptr(stream))
    _raise(rc, ctx)
    return bool(ok.value)


def kernel_profile(ctx: Context, on: bool, only=None) -> None:
    """Start (clearing old totals) or stop HIP-event profiling of the tagged
    kernels launched by ``ctx``; ``only`` (names) limits it to those tags."""
    names = ",".join(only) if only else ""
    ctx._check(_lib.psk_profile_only(ctx.handle, names.encode()))
    ctx._check(_lib.psk_profile(ctx.handle, 1 if on else 0))


def kernel_stats(ctx: Context):
    """[{name, launches, ms, bytes}] per tagged kernel since kernel_profile(ctx, True)."""
    cap = 256
    names = ctypes.create_string_buffer(1 << 16)
    launches = np.zeros(cap, dtype=np.uint64)
    ms = np.zeros(cap, dtype=np.float64)
    by = np.zeros(cap, dtype=np.float64)
    n = _lib.psk_profile_stats(ctx.handle, names, len(names), launches.ctypes.data,
                               ms.ctypes.data, by.ctypes.data, cap)
    if n < 0:
        raise PskError(n, ctx.last_error())
    nm = names.value.decode().split("\n") if n else []
    return [{"name": nm[i], "launches": int(launches[i]), "ms": float(ms[i]),
             "bytes": float(by[i])} for i in range(n)]


def batch8(texts, ctx: Context | None = None):
    """SAs of many independent byte texts (host buffers)."""
    arrs = [_as_u8(t) for t in texts]
    outs = [np.empty(len(a), dtype=np.uint32) for a in arrs]
    k = len(arrs)
    tp = (ctypes.c_void_p * k)(*[a.ctypes.data for a in arrs])
    lp = (ctypes.c_uint64 * k)(*[len(a) for a in arrs])
    sp = (ctypes.c_void_p * k)(*[o.ctypes.data for o in outs])
    rc = _lib.psk_batch8(_ctx_handle(ctx), tp, lp, sp, k)
    _raise(rc, ctx)
    return outs


def batch8_dev(d_texts, offsets, d_sas, ctx: Context | None = None, stream=None):
    """SAs of concatenated device texts: text i is d_texts[offsets[i]:offsets[i+1]]."""
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    rc = _lib.psk_batch8_dev(_ctx_handle(ctx), _dev_ptr(d_texts), off.ctypes.data,
                             len(off) - 1, _dev_ptr(d_sas), _stream_ptr(stream))
    _raise(rc, ctx)


class Server:
    """Continuous-batching SA server (``psk_server_*``): requests are admitted
    in arrival order against an HBM budget, small ones are coalesced into
    batches, and ``workers`` worker threads (one stream each) overlap their
    jobs on the device.  Results equal single :func:`psacak` calls."""

    def __init__(self, hbm_budget: int = 0, workers: int = 2):
        h = ctypes.c_void_p()
        rc = _lib.psk_server_create(ctypes.byref(h), hbm_budget, workers)
        if rc != PSK_OK:
            raise PskError(rc)
        self._h = h
        self._n = {}

    @staticmethod
    def footprint(nbytes: int, batched: bool = False) -> int:
        return int(_lib.psk_server_footprint(nbytes, 1 if batched else 0))

    def submit(self, text) -> int:
        """Queue a copy of text; the SA is copied out by poll/wait."""
        t = _as_u8(text)
        rid = ctypes.c_uint64()
        rc = _lib.psk_server_submit(self._h, t.ctypes.data, len(t), ctypes.byref(rid))
        _raise(rc)
        self._n[rid.value] = (len(t), None, None)
        return rid.value

    def submit_into(self, text, out=None) -> int:
        """Queue text without copying it; the SA is written straight into out
        (a new uint32 array if None), which poll/wait return.  The server keeps
        references to both until the request is done."""
        t = _as_u8(text)
        if out is None:
            out = np.empty(len(t), dtype=np.uint32)
        if out.dtype != np.uint32 or not out.flags.c_contiguous or len(out) < len(t):
            raise ValueError("out must be a contiguous uint32 array of at least len(text)")
        rid = ctypes.c_uint64()
        rc = _lib.psk_server_submit_into(self._h, t.ctypes.data, len(t), out.ctypes.data,
                                         ctypes.byref(rid))
        _raise(rc)
        self._n[rid.value] = (len(t), t, out)
        return rid.value

    def _out(self, rid):
        n, _, out = self._n.get(rid, (0, None, None))
        return (out if out is not None else np.empty(n, dtype=np.uint32)), out is not None

    def poll(self, rid: int):
        """(done, sa or None, latency_ms or None); a finished request is released."""
        done = ctypes.c_int(0)
        lat = ctypes.c_double(0)
        sa, borrowed = self._out(rid)
        rc = _lib.psk_server_poll(self._h, rid, ctypes.byref(done),
                                  None if borrowed else sa.ctypes.data, ctypes.byref(lat))
        if not done.value and rc == PSK_OK:
            return False, None, None
        self._n.pop(rid, None)
        _raise(rc)
        return True, sa[: len(sa)], lat.value

    def wait(self, rid: int, timeout_ms: float = -1.0):
        """Block until the request finishes; returns (sa, latency_ms)."""
        lat = ctypes.c_double(0)
        sa, borrowed = self._out(rid)
        rc = _lib.psk_server_wait(self._h, rid, None if borrowed else sa.ctypes.data,
                                  ctypes.byref(lat), timeout_ms)
        if rc != PSK_EBUSY:
            self._n.pop(rid, None)
        _raise(rc)
        return sa, lat.value

    def peak_bytes(self) -> int:
        return int(_lib.psk_server_peak_bytes(self._h))

    def pending(self) -> int:
        return int(_lib.psk_server_pending(self._h))

    def stats(self) -> dict:
        """Admission counters (psk_server_stats)."""
        out = np.zeros(7, dtype=np.uint64)
        n = _lib.psk_server_stats(self._h, out.ctypes.data, 7)
        _raise(n if n < 0 else PSK_OK)
        keys = ["jobs", "admitted", "blocked", "trims", "max_inflight", "order_violations",
                "peak_bytes"]
        return {k: int(v) for k, v in zip(keys, out[:n])}

    def close(self):
        if getattr(self, "_h", None):
            _lib.psk_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
