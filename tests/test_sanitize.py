"""Sanitizer builds (CPU): the library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (device code unchanged), driven through the C ABI
without a GPU -- argument checks, contexts, settings, workspace bounds and the
server's threads -- and the CPU oracle under the same sanitizers on random and
adversarial texts.  Every failure must be a status code, with no memory error,
undefined behaviour or leak."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if not shutil.which("gcc") or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("toolchain absent")
    out = tmp_path_factory.mktemp("san")
    r = subprocess.run(["make", "-s", "-C", HERE, f"OUT={out}", "-j2"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return out


def _run(exe):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("PSACAK_AMD_FAIL_ALLOC", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "LeakSanitizer" not in r.stderr
    return r.stdout


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="host-only run (no GPU present)")
def test_library_host_code_under_sanitizers(built):
    assert "all expectations hold" in _run(built / "abi_san")


def test_oracle_under_sanitizers(built):
    assert "all agree" in _run(built / "oracle_san")
