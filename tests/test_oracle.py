"""Pins the CPU oracle (oracle/psk_oracle.c) before it is trusted as the
checker of the GPU path.

Each check is independent of the oracle's own formulas: the crate's literal
known-answer vectors, Python brute force (slice comparison), the crate's
brute-force type definitions (common.rs:370-396), closed forms, and planted
mistakes the SA checker must catch.
"""
import json
import os
import random

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def brute_sa(t):
    t = list(t)
    return sorted(range(len(t)), key=lambda i: t[i:])


# ---- known answers ---------------------------------------------------------

def test_doc_tests_mississippi():
    d = GOLD["doc_tests"]
    text = d["text"].encode()
    assert oracle.sacak8(text).tolist() == d["psacak"]
    suf = np.zeros(len(text) + 1, dtype=np.uint32)
    suf[0] = len(text)
    suf[1:] = oracle.sacak8(text)
    assert suf.tolist() == d["with_sentinel"]


@pytest.mark.parametrize("i", range(11))
def test_table_texts(i):
    case = GOLD["table"][i]
    assert oracle.sacak8(np.array(case["text"], dtype=np.uint8)).tolist() == case["sa"]
    assert oracle.sacak32_raw(np.array(case["text"], dtype=np.uint32)).tolist() == case["sa"]
    assert oracle.saca_naive8(np.array(case["text"], dtype=np.uint8)).tolist() == case["sa"]


def test_naive_against_python_brute_force():
    rng = random.Random(7)
    for _ in range(300):
        n = rng.randint(0, 60)
        t = [rng.randrange(rng.choice([1, 2, 4, 256])) for _ in range(n)]
        assert oracle.saca_naive8(np.array(t, dtype=np.uint8)).tolist() == brute_sa(t)
        u = [rng.randrange(1 << 32) if rng.random() < 0.3 else rng.randrange(5) for _ in range(n)]
        assert oracle.saca_naive32(np.array(u, dtype=np.uint32)).tolist() == brute_sa(u)


# ---- full SACA-K restatements ----------------------------------------------

@settings(max_examples=300, deadline=None)
@given(st.lists(st.integers(0, 255), max_size=300), st.sampled_from([1, 2, 3, 4, 8, 256]))
def test_sacak8_matches_brute_force(raw, sigma):
    t = [x % sigma for x in raw]
    assert oracle.sacak8(np.array(t, dtype=np.uint8)).tolist() == brute_sa(t)


@settings(max_examples=300, deadline=None)
@given(st.lists(st.integers(0, 40), max_size=200))
def test_sacak32_matches_brute_force(t):
    assert oracle.sacak32_raw(np.array(t, dtype=np.uint32)).tolist() == brute_sa(t)


@pytest.mark.parametrize("text", [
    b"a" * 1000, b"ab" * 700, bytes(range(256)) * 3, bytes(reversed(range(256))) * 2,
    b"abracadabra" * 97, b"\x00" * 513 + b"\x01", b"\xff\x00" * 333,
])
def test_sacak8_adversarial(text):
    assert oracle.sacak8(text).tolist() == brute_sa(text)


def test_sacak8_larger_random_against_checker_and_numpy():
    rng = np.random.default_rng(11)
    for sigma in (2, 4, 20, 256):
        t = rng.integers(0, sigma, 20000).astype(np.uint8)
        sa = oracle.sacak8(t)
        assert oracle.check_sa8(t, sa) == 0
    # fibonacci word: deep recursion
    a, b = b"a", b"ab"
    while len(b) < 30000:
        a, b = b, b + a
    sa = oracle.sacak8(b)
    assert sa.tolist() == brute_sa(b)


# ---- checker: planted mistakes ---------------------------------------------

def test_checker_catches_planted_mistakes():
    rng = np.random.default_rng(3)
    t = rng.integers(0, 4, 5000).astype(np.uint8)
    sa = oracle.saca_naive8(t)
    assert oracle.check_sa8(t, sa) == 0
    bad = sa.copy(); bad[[10, 11]] = bad[[11, 10]]
    assert oracle.check_sa8(t, bad) == 2
    bad = sa.copy(); bad[5] = bad[6]
    assert oracle.check_sa8(t, bad) == 1
    bad = sa.copy(); bad[0] = len(t)
    assert oracle.check_sa8(t, bad) == 1
    bad = sa.copy(); bad[100:200] = bad[100:200][::-1]
    assert oracle.check_sa8(t, bad) == 2
    u = t.astype(np.uint32)
    assert oracle.check_sa32(u, sa) == 0
    bad = sa.copy(); bad[[0, len(sa) - 1]] = bad[[len(sa) - 1, 0]]
    assert oracle.check_sa32(u, bad) == 2


# ---- typing: common.rs:370-396 brute-force definitions ------------------------

def is_stype(t, i):
    for j in range(i + 1, len(t)):
        if t[i] > t[j]:
            return False
        if t[i] < t[j]:
            return True
    return False


def is_left_most(t, i):
    if i == 0:
        return False
    c0, c1 = t[i - 1], t[i]
    if c0 == c1:
        return False
    return c0 > c1 if is_stype(t, i) else c0 < c1


@settings(max_examples=200, deadline=None)
@given(st.lists(st.integers(0, 5), max_size=80))
def test_types_against_definitions(t):
    ty8 = oracle.types8(np.array(t, dtype=np.uint8))
    ty32 = oracle.types32(np.array(t, dtype=np.uint32))
    for i in range(len(t)):
        s = is_stype(t, i)
        lms = s and is_left_most(t, i)
        assert bool(ty8[i] & 1) == s and bool(ty8[i] & 2) == lms
        assert ty32[i] == ty8[i]


def test_buckets_closed_form():
    rng = np.random.default_rng(5)
    t = rng.integers(0, 7, 3000).astype(np.uint8)
    bounds, lms = oracle.buckets8(t)
    cnt = np.bincount(t, minlength=256)
    assert bounds.tolist() == [0] + np.cumsum(cnt).tolist()
    ty = oracle.types8(t)
    assert lms.tolist() == np.bincount(t[(ty & 2) > 0], minlength=256).tolist()


# ---- naming / translation ---------------------------------------------------

def lms_substring(t, p, lms_positions):
    nxt = [q for q in lms_positions if q > p]
    if not nxt:
        return None  # runs into the sentinel: unique
    return list(t[p:nxt[0] + 1])


@settings(max_examples=150, deadline=None)
@given(st.lists(st.integers(0, 3), min_size=4, max_size=120))
def test_naming_is_canonical(t):
    t8 = np.array(t, dtype=np.uint8)
    ty = oracle.types8(t8)
    lms = [i for i in range(len(t)) if ty[i] & 2]
    if not lms:
        return
    # LMS positions sorted by their full suffixes (brute force)
    srt = sorted(lms, key=lambda i: t[i:])
    k, red = oracle.name8(t8, np.array(srt, dtype=np.uint32))
    # The translated text maps a name to its bucket: L-type -> head,
    # S-type -> tail-1.  Along the suffix order, different substrings must
    # move to a strictly later bucket; equal ones stay in the bucket (L before
    # S, so values never decrease) and equal types give equal values.
    by_pos = dict(zip(lms, red.tolist()))
    rty = dict(zip(lms, (oracle.types32(red) & 1).tolist()))
    subs = {p: lms_substring(t, p, lms) for p in lms}
    for a, b in zip(srt, srt[1:]):
        eq = subs[a] is not None and subs[a] == subs[b]
        if eq:
            assert by_pos[a] <= by_pos[b]
            if rty[a] == rty[b]:
                assert by_pos[a] == by_pos[b]
        else:
            assert by_pos[a] < by_pos[b]
    distinct = len({tuple(s) if s is not None else ("$", p) for p, s in subs.items()})
    assert k == distinct


@settings(max_examples=200, deadline=None)
@given(st.lists(st.integers(0, 30), min_size=1, max_size=100))
def test_translate_closed_form(t):
    a = np.array(t, dtype=np.uint32)
    tr = oracle.translate32(a)
    if len(set(t)) >= len(t):
        assert tr is None
        return
    ty = oracle.types32(a)
    srt = sorted(t)
    import bisect
    for i, c in enumerate(t):
        head = bisect.bisect_left(srt, c)
        tail = bisect.bisect_right(srt, c)
        assert tr[i] == (tail - 1 if ty[i] & 1 else head)


def test_level_stats_shape():
    rng = np.random.default_rng(2)
    t = (rng.integers(0, 4, 1 << 16)).astype(np.uint8)
    st_ = oracle.level_stats8(t)
    assert st_[0, 0] == len(t)
    for a, b in zip(st_, st_[1:]):
        assert b[0] == a[1]  # next level's size = this level's LMS count
    assert st_[-1, 3] == st_[-1, 1]  # the last level names everything uniquely


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_parallel_restatement_equals_serial(threads):
    """or_par_sacak8 (the pipelined parallel induce, timed CPU baseline) gives
    the serial oracle's SA, below and above the 2-block parallel threshold
    (psacak8.rs:15), on texts whose induces defer many writes."""
    rng = np.random.default_rng(11)
    old = oracle.par_threads()
    oracle.set_par_threads(threads)
    try:
        for n in [0, 1, 2, 3, 4, 9, 1000, (1 << 18) - 1, (1 << 18) + 5, 700_001]:
            texts = [
                rng.integers(0, 4, n).astype(np.uint8),
                rng.integers(0, 256, n).astype(np.uint8),
                np.full(n, 7, np.uint8),
                np.tile(np.frombuffer(b"abcab", np.uint8), n // 5 + 1)[:n].copy(),
            ]
            if n > 1000:
                # long repeats drive a deep recursion through the u32 levels
                base = rng.integers(0, 4, 5000).astype(np.uint8)
                texts.append(np.tile(base, n // 5000 + 1)[:n].copy())
            for t in texts:
                assert np.array_equal(oracle.par_sacak8(t), oracle.sacak8(t)), (n, t[:8])
        for n in [5, 3000, 400_003]:
            z = np.minimum(rng.zipf(1.3, n), 1 << 16).astype(np.uint32)
            assert np.array_equal(oracle.par_sacak32_raw(z), oracle.sacak32_raw(z)), n
    finally:
        oracle.set_par_threads(old)
