"""The ordering rule behind direct LMS-substring naming, checked against the
oracle's induced naming (oracle/psk_oracle.c, naming.rs:20-259).

The GPU names LMS substrings by sorting digit strings: symbol ranks + 1 up to
and including the terminal LMS symbol, then a terminator above every symbol
(or, for the substring that runs to the text's end, the sentinel digit 0),
then zeros.  Here a plain Python sort of the same digit strings must give the
names (after the crate's translation to bucket-pointer form) that induced
sorting gives, on random and periodic texts of bytes and of u32 symbols.
"""
import numpy as np
import pytest

import oracle


def lms_positions(ty):
    return [i for i in range(1, len(ty)) if ty[i] & 2]


def direct_translated_names(t, rank, sigma):
    ty = oracle.types8(t) if t.dtype == np.uint8 else oracle.types32(t)
    lms = lms_positions(ty)
    m1 = len(lms)
    n = len(t)
    keys = []
    for r, p in enumerate(lms):
        last = r + 1 == m1
        ln = (n - p) if last else (lms[r + 1] - p + 1)
        keys.append(tuple(rank[int(t[p + i])] + 1 for i in range(ln)) + (0 if last else sigma + 1,))
    width = max((len(k) for k in keys), default=0)
    keys = [k + (0,) * (width - len(k)) for k in keys]
    names = np.zeros(m1, dtype=np.uint32)
    name, prev = -1, None
    for r in sorted(range(m1), key=lambda r: keys[r]):
        if keys[r] != prev:
            name, prev = name + 1, keys[r]
        names[r] = name
    k = name + 1
    if m1 == 0:
        return 0, names
    # the crate's translation (naming.rs:263-274)
    tails = np.cumsum(np.bincount(names, minlength=k))
    nty = oracle.types32(names)
    out = np.array([(tails[c] - 1) if (nty[i] & 1) else (tails[c - 1] if c else 0)
                    for i, c in enumerate(names)], dtype=np.uint32)
    return k, out


def induced(t):
    if t.dtype == np.uint8:
        return oracle.name8(t, oracle.stage1_8(t))
    return oracle.name32(t, oracle.stage1_32(t))


@pytest.mark.parametrize("seed", range(4))
def test_byte_texts(seed):
    rng = np.random.default_rng(seed)
    for it in range(250):
        n = int(rng.integers(4, 260))
        sigma = int(rng.choice([2, 3, 4, 8, 256]))
        t = rng.integers(0, sigma, n).astype(np.uint8)
        if it % 9 == 0:
            unit = rng.integers(97, 100, int(rng.integers(1, 6))).astype(np.uint8)
            t = np.tile(unit, n // len(unit) + 1)[:n]
        present = sorted(set(t.tolist()))
        rank = {c: i for i, c in enumerate(present)}
        k1, r1 = direct_translated_names(t, rank, len(present))
        k2, r2 = induced(t)
        assert k1 == k2 and np.array_equal(r1, r2), t.tolist()


@pytest.mark.parametrize("seed", range(3))
def test_bucket_pointer_texts(seed):
    # u32 levels rank symbols by their bucket-pointer values themselves
    rng = np.random.default_rng(100 + seed)
    for _ in range(200):
        n = int(rng.integers(4, 200))
        u = rng.integers(0, int(rng.choice([2, 3, 10, 50])), n).astype(np.uint32)
        tr = oracle.translate32(u)
        if tr is None:
            continue
        k1, r1 = direct_translated_names(tr, {v: v for v in set(tr.tolist())}, n)
        k2, r2 = induced(tr)
        assert k1 == k2 and np.array_equal(r1, r2), u.tolist()


def test_a_dropped_terminator_is_caught():
    # with the terminator below the symbols the order is not the induced one
    t = np.frombuffer(b"cabacbcbcacbab", dtype=np.uint8).copy()
    rank = {c: i for i, c in enumerate(sorted(set(t.tolist())))}
    wrong = direct_translated_names(t, rank, -2)  # terminator digit -1 < every symbol
    assert not np.array_equal(wrong[1], induced(t)[1])
