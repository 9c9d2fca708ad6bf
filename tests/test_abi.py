"""The C ABI without a GPU: the library loads, exports every symbol that
include/psacak.h declares, and rejects bad arguments with status codes before
touching the device (the crate panics on the same inputs: lib.rs:65,78-79,
psacak32.rs:25)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import psacak_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "psacak.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(psk_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", P.lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (psk_[a-z0-9_]+)", out))
    decl = declared_symbols()
    assert len(decl) >= 15
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    L = ctypes.CDLL(P.lib_path())
    for s in decl:
        getattr(L, s)


def test_version_and_strerror():
    assert "gfx950" in P.version()
    for code in range(-8, 1):
        assert P._strerror(code)


def test_argument_checks_without_gpu():
    L = P._lib
    buf = np.zeros(16, dtype=np.uint32)
    txt = np.zeros(16, dtype=np.uint8)
    # sa_len < n (lib.rs:78)
    assert L.psk_psacak_inplace(None, txt.ctypes.data, 16, buf.ctypes.data, 15) == P.PSK_EINVAL
    # n >= u32::MAX (lib.rs:65,79)
    assert L.psk_psacak_inplace(None, txt.ctypes.data, (1 << 32) - 1, buf.ctypes.data,
                                1 << 40) == P.PSK_ETOOLONG
    # null pointers
    assert L.psk_psacak_inplace(None, None, 4, buf.ctypes.data, 4) == P.PSK_EINVAL
    assert L.psk_psacak(None, txt.ctypes.data, 4, None) == P.PSK_EINVAL
    # psacak32 length limit (psacak32.rs:25)
    assert L.psk_psacak32(None, buf.ctypes.data, 1 << 31, buf.ctypes.data, 1 << 31) == \
        P.PSK_ETOOLONG
    assert L.psk_psacak32_raw(None, buf.ctypes.data, 8, buf.ctypes.data, 7) == P.PSK_EINVAL
    # empty texts are fine and need no device
    assert L.psk_psacak_inplace(None, None, 0, None, 0) == P.PSK_OK
    assert P.psacak(b"").tolist() == []
    assert P.psacak32(np.zeros(0, dtype=np.uint32)).tolist() == []
    with pytest.raises(P.PskError) as e:
        P.psacak_inplace(b"abc", np.zeros(2, dtype=np.uint32))
    assert e.value.code == P.PSK_EINVAL
    with pytest.raises(TypeError):
        P.psacak_inplace(b"abc", np.zeros(3, dtype=np.int64))
    # batch: one bad member rejects the whole call before any work
    with pytest.raises(P.PskError):
        P._raise(L.psk_batch8(None, None, None, None, 3))


def test_workspace_footprint_is_monotone():
    prev = 0
    for n in [0, 10, 1 << 10, 1 << 20, 1 << 26, 1 << 30]:
        b = P.workspace_bytes8(n)
        assert b >= prev
        prev = b
    assert P.workspace_bytes8(1 << 30) < 200 * (1 << 30)


def test_context_create_without_device():
    with P.Context(hbm_budget=1 << 30) as c:
        assert c.last_error() == ""
        assert c.last_levels().shape[1] == 6


def test_direct_rounds_setting_without_gpu():
    with P.Context() as c:
        for r in (0, 1, 6, 64):
            c.set_direct_rounds(r)
        for r in (-1, 65):
            with pytest.raises(P.PskError) as e:
                c.set_direct_rounds(r)
            assert e.value.code == P.PSK_EINVAL


def test_doubling_setting_without_gpu():
    with P.Context() as c:
        for v in (0, 1, 900, 1024):
            c.set_doubling(v)
        for v in (-1, 1025):
            with pytest.raises(P.PskError) as e:
                c.set_doubling(v)
            assert e.value.code == P.PSK_EINVAL


def test_chain_induce_setting_without_gpu():
    with P.Context() as c:
        for v in (0, 1, 64, 256):
            c.set_chain_induce(v)
        for v in (-1, 257):
            with pytest.raises(P.PskError) as e:
                c.set_chain_induce(v)
            assert e.value.code == P.PSK_EINVAL
    assert P._lib.psk_ctx_set_chain_induce(None, 1) == P.PSK_EINVAL


def test_dict_naming_setting_without_gpu():
    with P.Context() as c:
        c.set_dict_naming(False)
        c.set_dict_naming(True)
    assert P._lib.psk_ctx_set_dict_naming(None, 1) == P.PSK_EINVAL


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_an_error_not_an_abort():
    with pytest.raises(P.PskError) as e:
        P.psacak(b"mississippi")
    assert e.value.code == P.PSK_EDEVICE


def test_server_admission_without_gpu():
    with P.Server(hbm_budget=1 << 30, workers=2) as s:
        # alone larger than the budget: rejected at submit, before any device work
        big = P.Server.footprint(1 << 30)
        assert big > 1 << 30
        with pytest.raises(P.PskError) as e:
            s._n  # noqa: B018
            rc = P._lib.psk_server_submit(s._h, np.zeros(8, np.uint8).ctypes.data, 1 << 30,
                                          ctypes.byref(ctypes.c_uint64()))
            P._raise(rc)
        assert e.value.code == P.PSK_EBUDGET
        assert s.pending() == 0
    assert P.Server.footprint(1000, True) > P.Server.footprint(1000, False)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_server_reports_device_error_without_gpu():
    with P.Server(workers=1) as s:
        rid = s.submit(b"mississippi")
        with pytest.raises(P.PskError) as e:
            s.wait(rid, timeout_ms=30000)
        assert e.value.code == P.PSK_EDEVICE


def test_tuned_defines_read_from_tuning_file(tmp_path, monkeypatch):
    import json

    from psacak_amd import build

    monkeypatch.setattr(build, "TUNING", str(tmp_path / "none.json"))
    assert build.tuned_defines() == []
    f = tmp_path / "tuning.json"
    f.write_text(json.dumps({"defines": {"PSK_RS_ROUNDS": 8, "PSK_MS_ROUNDS": 16}}))
    monkeypatch.setattr(build, "TUNING", str(f))
    assert sorted(build.tuned_defines()) == ["PSK_MS_ROUNDS=16", "PSK_RS_ROUNDS=8"]


def test_server_concurrent_waits_on_one_id():
    # two threads wait on the same request: exactly one receives its result
    # (the SA, or the device error where there is no GPU), the other finds
    # the request released (PSK_ENOTFOUND); neither reads freed state
    import threading

    with P.Server(workers=1) as s:
        rid = s.submit(b"mississippi" * 50)
        codes = []
        lock = threading.Lock()

        def waiter():
            try:
                s.wait(rid, timeout_ms=60000)
                rc = P.PSK_OK
            except P.PskError as e:
                rc = e.code
            with lock:
                codes.append(rc)

        ts = [threading.Thread(target=waiter) for _ in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert codes.count(P.PSK_ENOTFOUND) == 1, codes
        assert set(codes) - {P.PSK_ENOTFOUND} <= {P.PSK_OK, P.PSK_EDEVICE}, codes
        # poll after the release: unknown id
        done = ctypes.c_int(0)
        assert P._lib.psk_server_poll(s._h, rid, ctypes.byref(done), None, None) == P.PSK_ENOTFOUND
        assert s.pending() == 0
        st = s.stats()
        assert st["admitted"] == 1 and st["order_violations"] == 0


def test_default_context_trim_and_workspace():
    # ctx NULL names the calling thread's default context, as in the entry points
    r, p = ctypes.c_uint64(7), ctypes.c_uint64(7)
    assert P._lib.psk_ctx_workspace(None, ctypes.byref(r), ctypes.byref(p)) == P.PSK_OK
    assert P._lib.psk_ctx_trim(None) == P.PSK_OK
    assert P._lib.psk_ctx_workspace(None, ctypes.byref(r), None) == P.PSK_OK
    assert r.value == 0


def test_capture_settings_without_gpu():
    with P.Context() as c:
        c.set_capture(0)
        c.set_capture(-1)
        assert c.last_capture() is None
        for v in (-2, 64):
            with pytest.raises(P.PskError) as e:
                c.set_capture(v)
            assert e.value.code == P.PSK_EINVAL
    assert P._lib.psk_check32_dev(None, None, 4, None, None, None) == P.PSK_EINVAL
