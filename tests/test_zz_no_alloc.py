"""psk_psacak8_dev_noalloc: the zero-allocation variant of psk_psacak8_dev
(no counterpart in the reference).  Called through the C ABI and checked
against psk_psacak8_dev on the same inputs (bit-exact), on consecutive calls
whose second text is smaller, on texts at the size limit, and for the
rejection of longer texts."""
import ctypes

import numpy as np
import pytest

import oracle
import psacak_amd as P
from psacak_amd import gen

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _dev_sa(t, ctx):
    import torch
    dt = torch.from_numpy(np.array(t, dtype=np.uint8)).to("cuda:0")
    ds = torch.empty(len(t), dtype=torch.int32, device="cuda:0")
    P.psacak8_dev(dt, ds, ctx=ctx)
    torch.cuda.synchronize()
    return ds.cpu().numpy().view(np.uint32)


def _noalloc_sa(t, stream=None):
    import torch
    dt = torch.from_numpy(np.array(t, dtype=np.uint8)).to("cuda:0")
    ds = torch.full((len(t),), -1, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    rc = P._lib.psk_psacak8_dev_noalloc(dt.data_ptr(), len(t), ds.data_ptr(),
                                        None if stream is None else stream.cuda_stream)
    assert rc == P.PSK_OK, rc
    torch.cuda.synchronize()
    return ds.cpu().numpy().view(np.uint32)


def test_noalloc_equals_psacak8_dev_two_calls_second_smaller():
    note("noalloc two calls")
    big = gen.markov_dna(P.NOALLOC_MAX_N, seed=41)
    small = gen.iid_dna(300000, seed=42)
    with P.Context() as c:
        ref_big, ref_small = _dev_sa(big, c), _dev_sa(small, c)
    got_big = _noalloc_sa(big)
    got_small = _noalloc_sa(small)
    assert np.array_equal(got_big, ref_big)
    assert np.array_equal(got_small, ref_small)
    assert np.array_equal(got_small, oracle.sacak8(small))


def test_noalloc_texts_at_the_limit_and_on_a_stream():
    import torch
    n = P.NOALLOC_MAX_N
    fib_a, fib_b = b"a", b"ab"
    while len(fib_b) < n:
        fib_a, fib_b = fib_b, fib_b + fib_a
    texts = {
        "ab_period": np.frombuffer(b"ab" * (n // 2), dtype=np.uint8),
        "fibonacci": np.frombuffer(fib_b[:n], dtype=np.uint8),
        "uniform256": gen.uniform8(n, alphabet=256, seed=3),
        "uniform2": gen.uniform8(n, alphabet=2, seed=4),
        "all_same": np.full(n, 65, dtype=np.uint8),
        "tiny": np.frombuffer(b"mississippi", dtype=np.uint8),
    }
    s = torch.cuda.Stream()
    with P.Context() as c:
        for name, t in texts.items():
            note(f"noalloc limit {name}")
            assert np.array_equal(_noalloc_sa(t, stream=s), _dev_sa(t, c)), name


def test_noalloc_makes_no_device_allocation():
    import torch
    note("noalloc memory")
    t = gen.markov_dna(P.NOALLOC_MAX_N, seed=43)
    _noalloc_sa(t)  # module and scratch loaded
    dt = torch.from_numpy(t).to("cuda:0")
    ds = torch.empty(len(t), dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        assert P._lib.psk_psacak8_dev_noalloc(dt.data_ptr(), len(t), ds.data_ptr(), None) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    # the context path would reserve >= 64 MiB chunks for this text
    assert free0 - free1 < (8 << 20), (free0, free1)


def test_noalloc_rejects_longer_texts():
    note("noalloc limit check")
    import torch
    dt = torch.zeros(8, dtype=torch.uint8, device="cuda:0")
    ds = torch.zeros(8, dtype=torch.int32, device="cuda:0")
    # the length check comes before any device access
    rc = P._lib.psk_psacak8_dev_noalloc(dt.data_ptr(), P.NOALLOC_MAX_N + 1, ds.data_ptr(), None)
    assert rc == P.PSK_ETOOLONG
    assert P._lib.psk_psacak8_dev_noalloc(None, 5, ds.data_ptr(), None) == P.PSK_EINVAL
    assert P._lib.psk_psacak8_dev_noalloc(dt.data_ptr(), 0, ds.data_ptr(), None) == P.PSK_OK
