"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the
crate's known answers.  The SA is unique, so equality with the oracle is
equality with the crate.  Integer work: the bar is bit-exact."""
import json
import os
import sys

import numpy as np
import pytest

import oracle
import psacak_amd as P
from psacak_amd import gen

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat.json")))


def note(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = P.Context()
    yield c
    c.close()


def test_doc_tests_mississippi(ctx):
    note("mississippi")
    d = GOLD["doc_tests"]
    text = d["text"].encode()
    assert P.psacak(text, ctx).tolist() == d["psacak"]
    suf = np.zeros(len(text), dtype=np.uint32)
    P.psacak_inplace(text, suf, ctx)
    assert suf.tolist() == d["psacak_inplace"]
    suf = np.zeros(len(text) + 1, dtype=np.uint32)
    suf[0] = len(text)
    P.psacak_inplace(text, suf[1:], ctx)
    assert suf.tolist() == d["with_sentinel"]


def test_inplace_leaves_tail_untouched(ctx):
    note("inplace tail")
    suf = np.full(20, 7777, dtype=np.uint32)
    P.psacak_inplace(b"banana", suf, ctx)
    assert suf[:6].tolist() == [5, 3, 1, 0, 4, 2]
    assert (suf[6:] == 7777).all()


@pytest.mark.parametrize("i", range(11))
def test_table_texts(ctx, i):
    note(f"table {i}")
    case = GOLD["table"][i]
    assert P.psacak(np.array(case["text"], dtype=np.uint8), ctx).tolist() == case["sa"]
    assert P.psacak32_raw(np.array(case["text"], dtype=np.uint32), ctx).tolist() == case["sa"]
    tr = oracle.translate32(np.array(case["text"], dtype=np.uint32))
    if tr is not None:
        assert P.psacak32(tr, ctx).tolist() == case["sa"]


def test_tiny_and_empty(ctx):
    note("tiny")
    for t in [b"", b"a", b"ba", b"ab", b"aa", b"cba", b"aaa", b"aba", b"abca", b"zzzz"]:
        assert P.psacak(t, ctx).tolist() == oracle.saca_naive8(t).tolist(), t


def test_random_small_many(ctx):
    rng = np.random.default_rng(1234)
    for it in range(400):
        n = int(rng.integers(4, 600))
        sigma = int(rng.choice([1, 2, 3, 4, 5, 16, 256]))
        t = rng.integers(0, sigma, n).astype(np.uint8)
        if it % 5 == 0:
            t = t + np.uint8(60)
        note(f"small it={it} n={n} sigma={sigma}")
        got = P.psacak(t, ctx)
        exp = oracle.sacak8(t)
        assert np.array_equal(got, exp), (it, n, sigma, t.tolist())


ADVERSARIAL = {
    "all_same_4k": b"a" * 4096,
    "all_same_70k": b"q" * 70000,
    "period2": b"ab" * 30000,
    "period3": b"abc" * 20000,
    "descending": bytes(range(255, -1, -1)) * 64,
    "ascending": bytes(range(256)) * 64,
    "runs": b"".join(bytes([65 + (i % 4)]) * (1 + (i * 7919) % 300) for i in range(2000)),
    "zero_runs": b"\x00" * 10000 + b"\x01" + b"\x00" * 9999,
    "ff00": b"\xff\x00" * 20000,
    "ff_runs": (b"\xff\xff\xfe" * 9000 + b"\xff" * 700 + b"\x00\xff\xfd\xff") * 3,
}


@pytest.mark.parametrize("name", sorted(ADVERSARIAL))
def test_adversarial(ctx, name):
    note(f"adversarial {name}")
    t = ADVERSARIAL[name]
    assert np.array_equal(P.psacak(t, ctx), oracle.sacak8(t))


def test_fibonacci_word(ctx):
    note("fibonacci")
    a, b = b"a", b"ab"
    while len(b) < 200000:
        a, b = b, b + a
    assert np.array_equal(P.psacak(b, ctx), oracle.sacak8(b))


@pytest.mark.parametrize("n", [1 << 10, 1 << 16, 1 << 20, 1 << 22])
def test_iid_dna_sizes(ctx, n):
    note(f"iid dna n={n}")
    t = gen.iid_dna(n, seed=n)
    assert np.array_equal(P.psacak(t, ctx), oracle.sacak8(t))


@pytest.mark.parametrize("n", [1 << 16, 1 << 22])
def test_markov_dna(ctx, n):
    note(f"markov dna n={n}")
    t = gen.markov_dna(n, seed=7)
    assert np.array_equal(P.psacak(t, ctx), oracle.sacak8(t))


@pytest.mark.parametrize("alphabet", [2, 4, 128, 256])
def test_uniform_alphabets(ctx, alphabet):
    note(f"uniform alphabet={alphabet}")
    t = gen.uniform8(1 << 20, alphabet=alphabet, seed=alphabet)
    assert np.array_equal(P.psacak(t, ctx), oracle.sacak8(t))


@pytest.mark.parametrize("rounds,dict_naming",
                         [(0, True), (1, True), (2, True), (6, True), (1, False), (6, False),
                          (32, False)])
def test_naming_strategies_agree(rounds, dict_naming):
    # direct LMS-substring sorting with at most `rounds` rounds, induced
    # sorting where ties remain (0 = induced only), at every level; round 1
    # at level 0 by dictionary or by sorting
    note(f"naming rounds={rounds} dict={dict_naming}")
    rng = np.random.default_rng(31)
    texts = [gen.markov_dna(1 << 20, seed=9), ADVERSARIAL["descending"], ADVERSARIAL["runs"],
             gen.uniform8(200000, alphabet=3, seed=2), b"abcab" * 20000 + b"x" + b"abcab" * 9000]
    a, b = b"a", b"ab"
    while len(b) < 50000:
        a, b = b, b + a
    texts.append(b)
    with P.Context() as c:
        c.set_direct_rounds(rounds)
        c.set_dict_naming(dict_naming)
        for i, t in enumerate(texts):
            note(f"naming rounds={rounds} text {i}")
            assert np.array_equal(P.psacak(t, c), oracle.sacak8(t)), i
        for it in range(40):
            n = int(rng.integers(2, 400))
            u = rng.integers(0, int(rng.choice([2, 3, 10, 50])), n).astype(np.uint32)
            tr = oracle.translate32(u)
            if tr is not None:
                assert np.array_equal(P.psacak32(tr, c), oracle.sacak32(tr)), it
        z = gen.zipf32(1 << 18, sigma=1 << 12, s=1.1, repeat_frac=0.3, seed=3)
        assert np.array_equal(P.psacak32_raw(z, c), oracle.sacak32_raw(z))


@pytest.mark.parametrize("case", ["long_runs", "many_keys", "mid_keys", "skewed"])
def test_dictionary_naming_cases(case):
    # level-0 naming through the dictionary: long LMS substrings left tied
    # after round 1 (homopolymer runs), more distinct keys than the
    # dictionary holds (falls back to sorting), more than its LDS histogram
    # holds, and a few very frequent keys
    rng = np.random.default_rng(5)
    if case == "long_runs":
        parts = [bytes([int(rng.choice(list(b"ACGT")))]) * int(rng.integers(1, 90))
                 for _ in range(40000)]
        t = np.frombuffer(b"".join(parts), dtype=np.uint8)
    elif case == "many_keys":
        t = gen.uniform8(1 << 23, alphabet=256, seed=77)
    elif case == "mid_keys":
        # ~7*10^4 distinct keys: LDS histogram plus global counts
        t = gen.uniform8(1 << 22, alphabet=8, seed=77)
    else:
        t = np.frombuffer((b"CA" * 300000) + gen.iid_dna(1 << 20, seed=4).tobytes(), dtype=np.uint8)
    note(f"dictionary case={case} n={len(t)}")
    with P.Context() as c:
        got = P.psacak(t, c)
        c.set_dict_naming(False)
        assert np.array_equal(P.psacak(t, c), got)
    assert np.array_equal(got, oracle.sacak8(t))


@pytest.mark.parametrize("sigma", [2, 3, 7, 15, 31, 63, 127, 255])
def test_packed_digit_widths_and_offsets(ctx, sigma):
    # level-0 keys come out of a stream of bit_width(sigma + 1) bits per
    # byte (2..9 here); device texts at odd byte offsets take the packing
    # path for unaligned texts; round 1 by dictionary and by sorting
    import torch
    rng = np.random.default_rng(sigma)
    n = 150000
    base = rng.integers(0, sigma, n + 8).astype(np.uint8)
    base[n // 2:n // 2 + 3000] = base[1000:4000]
    dt_all = torch.from_numpy(base).to("cuda:0")
    with P.Context() as c:
        for dict_naming in (True, False):
            c.set_dict_naming(dict_naming)
            for off in (0, 1, 3):
                note(f"packed sigma={sigma} off={off} dict={dict_naming}")
                t = base[off:off + n]
                ds = torch.empty(n, dtype=torch.int32, device="cuda:0")
                P.psacak8_dev(dt_all[off:off + n], ds, ctx=c)
                torch.cuda.synchronize()
                assert np.array_equal(ds.cpu().numpy().view(np.uint32), oracle.sacak8(t)), off


@pytest.mark.parametrize("threshold", [0, 1, 512, 900, 1024])
def test_doubling_levels(threshold):
    # recursion levels sorted by prefix doubling when distinct enough (1 =
    # every level below level 0, 0 = never): periodic and repetitive texts
    # keep suffixes tied for many rounds
    note(f"doubling threshold={threshold}")
    a, b = b"a", b"ab"
    while len(b) < 200000:
        a, b = b, b + a
    blk = gen.iid_dna(1 << 19, seed=5).tobytes()
    texts = [gen.markov_dna(1 << 21, seed=3), gen.iid_dna(1 << 20, seed=4), b,
             b"abcab" * 30000 + b"x" + b"abcab" * 9000, ADVERSARIAL["runs"],
             gen.uniform8(300000, alphabet=3, seed=6),
             # every suffix of a copied block ties with its twin: doubling
             # gives up after a few rounds and the level recurses
             blk + blk, blk + b"A" + blk + blk[:1000]]
    with P.Context() as c:
        c.set_doubling(threshold)
        for i, t in enumerate(texts):
            note(f"doubling threshold={threshold} text {i}")
            assert np.array_equal(P.psacak(t, c), oracle.sacak8(t)), i
        z = gen.zipf32(1 << 19, sigma=1 << 14, s=1.1, repeat_frac=0.3, seed=8)
        assert np.array_equal(P.psacak32_raw(z, c), oracle.sacak32_raw(z))


@pytest.mark.parametrize("sigma", [64, 200, 256])
def test_doubling_singleton_symbols(sigma):
    # random texts over large alphabets: the names of their level-1 text are
    # nearly all distinct, so prefix doubling places almost every position
    # directly (a symbol that occurs once) and sorts only the few repeated
    # ones; a copied block keeps some repeated names
    rng = np.random.default_rng(sigma)
    with P.Context() as c:
        c.set_doubling(1)
        for n in (1 << 16, 1 << 20, (1 << 20) + 12345):
            t = rng.integers(0, sigma, n).astype(np.uint8)
            t[n // 2:n // 2 + 5000] = t[100:5100]
            note(f"doubling singletons sigma={sigma} n={n}")
            assert np.array_equal(P.psacak(t.tobytes(), c), oracle.sacak8(t)), n
        z = gen.zipf32(1 << 20, sigma=1 << 18, s=0.6, repeat_frac=0.01, seed=sigma)
        note(f"doubling singletons zipf32 sigma=2^18 (seed {sigma})")
        assert np.array_equal(P.psacak32_raw(z, c), oracle.sacak32_raw(z))


def test_psacak32_random_bucket_pointer_texts(ctx):
    rng = np.random.default_rng(99)
    for it in range(150):
        n = int(rng.integers(2, 400))
        u = rng.integers(0, int(rng.choice([2, 3, 10, 50, 1000])), n).astype(np.uint32)
        note(f"psacak32 it={it} n={n}")
        tr = oracle.translate32(u)
        if tr is None:
            continue
        assert np.array_equal(P.psacak32(tr, ctx), oracle.sacak32(tr)), (it, u.tolist())
        assert np.array_equal(P.psacak32_raw(u, ctx), oracle.sacak32(tr))


def test_psacak32_dev_unaligned_texts(ctx):
    # u32 device texts 4-byte but not 16-byte aligned (caller buffers at an
    # element offset) take the scalar typing and label paths
    import torch
    z = gen.zipf32(200003, sigma=1 << 12, s=1.1, repeat_frac=0.2, seed=12)
    tr = oracle.translate32(z)
    assert tr is not None
    n = len(z)
    for name, text, fn, exp in (("raw", z, P.psacak32_raw_dev, oracle.sacak32_raw(z)),
                                ("bucket-pointer", tr, P.psacak32_dev, oracle.sacak32(tr))):
        base = np.zeros(n + 4, dtype=np.uint32)
        for off in (0, 1, 2, 3):
            note(f"psacak32 {name} device text at element offset {off}")
            base[off:off + n] = text
            dt_all = torch.from_numpy(base.view(np.int32)).to("cuda:0")
            ds = torch.empty(n, dtype=torch.int32, device="cuda:0")
            fn(dt_all[off:off + n], ds, ctx=ctx)
            torch.cuda.synchronize()
            assert np.array_equal(ds.cpu().numpy().view(np.uint32), exp), (name, off)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_psacak32_long_equal_runs(ctx, seed):
    # long runs of one symbol keep the items of a type run tied for many
    # rounds of a level's final chain sorts (the ties halve per round), so
    # the sorts outlast the rounds whose keys carry run offsets and then
    # write every label again; L runs (a run followed by a smaller symbol)
    # and S runs (followed by a larger one) both occur
    rng = np.random.default_rng(seed)
    parts, total = [], 0
    while total < 300000:
        ln = int(rng.integers(1, 3000)) if rng.random() < 0.3 else int(rng.integers(1, 8))
        parts.append(np.full(ln, int(rng.integers(0, 50)), dtype=np.uint32))
        total += ln
    u = np.concatenate(parts)[:300000]
    note(f"psacak32 long runs seed={seed} n={len(u)}")
    assert np.array_equal(P.psacak32_raw(u, ctx), oracle.sacak32_raw(u))


def test_psacak32_raw_zipf(ctx):
    note("psacak32_raw zipf 1M")
    z = gen.zipf32(1 << 20, sigma=1 << 16, s=1.1, repeat_frac=0.2, seed=5)
    assert np.array_equal(P.psacak32_raw(z, ctx), oracle.sacak32_raw(z))
    note("psacak32_raw all distinct")
    perm = np.random.default_rng(1).permutation(5000).astype(np.uint32)
    assert np.array_equal(P.psacak32_raw(perm, ctx), np.argsort(perm, kind="stable").astype(np.uint32))


def test_psacak32_precondition(ctx):
    note("psacak32 precondition")
    with pytest.raises(P.PskError) as e:
        P.psacak32(np.array([5, 1, 0, 2], dtype=np.uint32), ctx)
    assert e.value.code == P.PSK_EPRECOND


def test_device_api_and_checker(ctx):
    import torch
    note("device api")
    t = gen.markov_dna(1 << 20, seed=3)
    dt = torch.from_numpy(t).to("cuda:0")
    ds = torch.empty(len(t), dtype=torch.int32, device="cuda:0")
    P.psacak8_dev(dt, ds, ctx=ctx)
    torch.cuda.synchronize()
    got = ds.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, oracle.sacak8(t))
    assert P.check8_dev(dt, ds, ctx=ctx)
    bad = ds.clone()
    bad[[100, 101]] = bad[[101, 100]]
    assert not P.check8_dev(dt, bad, ctx=ctx)
    bad = ds.clone()
    bad[7] = bad[8]
    assert not P.check8_dev(dt, bad, ctx=ctx)
    lv = ctx.last_levels()
    assert lv.shape[0] >= 2 and lv[0, 0] == len(t)


def test_batch(ctx):
    note("batch8 mixed sizes")
    rng = np.random.default_rng(8)
    L = gen.lengths(60, 4, 20000, seed=8)
    texts = [gen.iid_dna(int(n), seed=i) for i, n in enumerate(L)]
    texts += [b"", b"a", b"ab", b"ba", b"aaa", b"mississippi", b"", bytes(range(255, 0, -1))]
    texts += [rng.integers(0, 3, int(rng.integers(1, 50))).astype(np.uint8) for _ in range(40)]
    outs = P.batch8(texts, ctx)
    for i, (t, o) in enumerate(zip(texts, outs)):
        assert np.array_equal(o, oracle.sacak8(t)), i


def test_batch_many_small_and_dev(ctx):
    import torch
    note("batch8_dev 3000 texts")
    L = gen.lengths(3000, 1, 3000, seed=21)
    texts = [gen.markov_dna(int(n), seed=100 + i, repeat_frac=0.0) for i, n in enumerate(L)]
    off = np.zeros(len(texts) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(t) for t in texts])
    cat = np.concatenate(texts)
    dt = torch.from_numpy(cat).to("cuda:0")
    ds = torch.zeros(len(cat), dtype=torch.int32, device="cuda:0")
    P.batch8_dev(dt, off, ds, ctx=ctx)
    torch.cuda.synchronize()
    sa = ds.cpu().numpy().view(np.uint32)
    for i, t in enumerate(texts):
        assert np.array_equal(sa[off[i]:off[i + 1]], oracle.sacak8(t)), i


@pytest.mark.parametrize("dict_naming,rounds", [(True, 32), (False, 32), (True, 1), (False, 0)])
def test_batch_text_ends(dict_naming, rounds):
    # each text's last LMS substring runs into its own sentinel: identical
    # tails in many texts (equal symbols, distinct sentinels), tails longer
    # than one round's digits, thousands of texts (wide text indices), tiny
    # and empty texts; every naming strategy gives each text its own SA
    note(f"batch text ends dict={dict_naming} rounds={rounds}")
    rng = np.random.default_rng(41)
    base = [b"GATTACA", b"CAGCAG", b"ACGTACGA", b"TTTTTTTTTTTTTTTTTTTTTTTTTTTTTA", b"CA", b"A", b""]
    texts = []
    for i in range(4000):
        k = int(rng.integers(0, 6))
        if k == 0:
            texts.append(base[int(rng.integers(0, len(base)))])
        elif k == 1:
            texts.append(gen.iid_dna(int(rng.integers(1, 40)), seed=i).tobytes()
                         + base[int(rng.integers(0, 4))])
        elif k == 2:
            texts.append(gen.markov_dna(int(rng.integers(50, 3000)), seed=i).tobytes()
                         + b"A" * int(rng.integers(1, 40)))
        else:
            texts.append(gen.iid_dna(int(rng.integers(1, 600)), seed=7 * i).tobytes())
    with P.Context() as c:
        c.set_dict_naming(dict_naming)
        c.set_direct_rounds(rounds)
        outs = P.batch8(texts, c)
    for i, (t, o) in enumerate(zip(texts, outs)):
        assert np.array_equal(o, oracle.sacak8(np.frombuffer(t, dtype=np.uint8))), i


def test_host_staged_transfers(ctx):
    # host buffers of several staging chunks (64 MiB) with odd sizes: the
    # host entry points return exactly what the device entry points compute
    import torch
    note("host staged transfers u8")
    t = gen.markov_dna((24 << 20) + 5, seed=12)
    got = P.psacak(t, ctx)
    dt = torch.from_numpy(t).to("cuda:0")
    ds = torch.empty(len(t), dtype=torch.int32, device="cuda:0")
    P.psacak8_dev(dt, ds, ctx=ctx)
    torch.cuda.synchronize()
    assert np.array_equal(got, ds.cpu().numpy().view(np.uint32))
    assert P.check8_dev(dt, ds, len(t), ctx=ctx)
    note("host staged transfers u32")
    z = gen.zipf32((20 << 20) + 3, sigma=1 << 16, s=1.1, repeat_frac=0.1, seed=5)
    g32 = P.psacak32_raw(z, ctx)
    dz = torch.from_numpy(z.view(np.int32)).to("cuda:0")
    dzs = torch.empty(len(z), dtype=torch.int32, device="cuda:0")
    P.psacak32_raw_dev(dz, dzs, len(z), ctx=ctx)
    torch.cuda.synchronize()
    assert np.array_equal(g32, dzs.cpu().numpy().view(np.uint32))
    note("host staged transfers batch")
    L = gen.lengths(60, 1 << 10, 3 << 20, seed=19)
    texts = [gen.iid_dna(int(n), seed=300 + i) for i, n in enumerate(L)]
    outs = P.batch8(texts, ctx)
    off = np.zeros(len(texts) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(x) for x in texts])
    dt = torch.from_numpy(np.concatenate(texts)).to("cuda:0")
    ds = torch.empty(int(off[-1]), dtype=torch.int32, device="cuda:0")
    P.batch8_dev(dt, off, ds, ctx=ctx)
    torch.cuda.synchronize()
    sa = ds.cpu().numpy().view(np.uint32)
    for i, o in enumerate(outs):
        assert np.array_equal(o, sa[off[i]:off[i + 1]]), i


def test_batch_matches_single_calls_on_repeats(ctx):
    note("batch8 identical texts")
    t = gen.markov_dna(50000, seed=5)
    texts = [t, t[:49999], t, t[1:], t[:1000] * 50]
    outs = P.batch8(texts, ctx)
    for x, o in zip(texts, outs):
        assert np.array_equal(o, P.psacak(x, ctx))
        assert np.array_equal(o, oracle.sacak8(x))


def test_server_results_equal_alone_and_in_the_mix(ctx):
    import threading
    import time
    note("server mix")
    rng = np.random.default_rng(17)
    L = gen.lengths(120, 8, 400000, seed=33)
    texts = [gen.markov_dna(int(n), seed=200 + i) for i, n in enumerate(L)]
    texts += [b"banana", b"", b"a", gen.iid_dna(3 << 20, seed=1)]
    alone = [P.psacak(t, ctx) for t in texts]
    with P.Server(hbm_budget=64 << 30, workers=3) as srv:
        ids = {}

        def producer():
            # copying and borrowing submissions, interleaved
            for i, t in enumerate(texts):
                ids[i] = srv.submit(t) if i % 2 else srv.submit_into(t)
                time.sleep(float(rng.exponential(0.0005)))

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        lat = []
        for i, t in enumerate(texts):
            sa, ms = srv.wait(ids[i], timeout_ms=120000)
            lat.append(ms)
            assert np.array_equal(sa, alone[i]), i
        assert srv.peak_bytes() > 0
        assert srv.pending() == 0
    for i in range(0, len(texts), 17):
        assert np.array_equal(alone[i], oracle.sacak8(texts[i]))


def test_cli_matches_crate_output_format(tmp_path, capsys):
    from psacak_amd import cli
    note("cli")
    f = tmp_path / "in.txt"
    f.write_bytes(b"mississippi" * 1000)
    assert cli.main(["-t", "2", "-c", str(f)]) == 0
    err = capsys.readouterr().err.splitlines()
    assert err[0] == f"load 11000 bytes from `{f}`"
    assert err[1].startswith(" time: ") and err[1].count("s ") == 2
    assert err[2].startswith("  rss: ") and err[2].endswith("MiB")
    assert err[3] == "check: true"


@pytest.mark.parametrize("case", ["markov", "iid", "long_run"])
def test_large_queues_against_oracle(ctx, case):
    """Texts large enough for level-0 queues of more than 4M entries: filled
    without host synchronisation (device-side seed counts) when no run of
    equal bytes exceeds 256, with exact readbacks when one does (long_run)."""
    n = 40 << 20  # bucket parts of 5M+ entries
    note(f"large queues {case} n={n}")
    if case == "markov":
        t = gen.markov_dna(n, seed=21)
    else:
        t = gen.iid_dna(n, seed=22)
    if case == "long_run":
        t = t.copy()
        t[n // 2: n // 2 + 1000] = ord("G")
    assert np.array_equal(P.psacak(t, ctx), oracle.sacak8(t))


@pytest.mark.parametrize("threshold", [1, 0])
def test_chain_induce_level0(threshold):
    """Level 0's final induce as chain sorts (threshold 1: every byte text)
    or as bucket steps (0): the same SA as the oracle on small and large
    alphabets, runs, periodic and repetitive texts, and the level statistics
    of either path equal."""
    note(f"chain induce threshold={threshold}")
    a, b = b"a", b"ab"
    while len(b) < 100000:
        a, b = b, b + a
    blk = gen.uniform8(1 << 17, alphabet=200, seed=9).tobytes()
    texts = [gen.markov_dna(1 << 20, seed=11), gen.iid_dna(1 << 19, seed=12), b,
             gen.uniform8(1 << 20, alphabet=3, seed=13), gen.uniform8(1 << 20, alphabet=40, seed=14),
             gen.uniform8(1 << 20, alphabet=256, seed=15), ADVERSARIAL["runs"],
             bytes(range(256)) * 3000, bytes(range(255, -1, -1)) * 3000,
             blk + blk + blk[:5000], b"\xff" * 70000 + b"\x00" + b"\xfe" * 30000,
             b"mississippi", b"ab", b"aab", b"ba" * 7]
    with P.Context() as c:
        c.set_chain_induce(threshold)
        for i, t in enumerate(texts):
            note(f"chain induce threshold={threshold} text {i} n={len(t)}")
            assert np.array_equal(P.psacak(t, c), oracle.sacak8(t)), i
    rng = np.random.default_rng(31)
    with P.Context() as c:
        c.set_chain_induce(threshold)
        for it in range(60):
            n = int(rng.integers(4, 3000))
            t = rng.integers(0, int(rng.choice([2, 5, 60, 256])), n).astype(np.uint8).tobytes()
            note(f"chain induce threshold={threshold} random it={it} n={n}")
            assert np.array_equal(P.psacak(t, c), oracle.sacak8(t)), (it, t)


def test_chain_induce_large_alphabet_device_check(ctx):
    """A 64 MiB text over 256 symbols (the chain path by default) passes the
    device checker and the oracle's checker, within the workspace estimate."""
    import torch
    n = 64 << 20
    note(f"chain induce 256 symbols n={n}")
    t = gen.uniform8(n, alphabet=256, seed=16)
    dev = torch.device("cuda:0")
    d_t = torch.from_numpy(t).to(dev)
    d_sa = torch.empty(n, dtype=torch.int32, device=dev)
    with P.Context() as c:
        P.psacak8_dev(d_t, d_sa, ctx=c)
        torch.cuda.synchronize()
        reserved, peak = c.workspace()
        assert 0 < reserved <= peak <= P.workspace_bytes8(n), (reserved, peak)
    assert P.check8_dev(d_t, d_sa, ctx=ctx)
    # and the oracle's O(n) checker on the host
    assert oracle.check_sa8(t, d_sa.cpu().numpy().view(np.uint32)) == 0


@pytest.mark.parametrize("kind,n", [("iid", 64 << 20), ("markov", 256 << 20)])
def test_workspace_within_estimate(kind, n):
    """The workspace a call reserves stays within psk_workspace_bytes8, which
    the server's HBM admission relies on; a second call of another size on
    the same context reuses the chunks (no growth past the estimate)."""
    note(f"workspace {kind} {n}")
    import torch
    dev = torch.device("cuda:0")
    t = gen.iid_dna(n, seed=5) if kind == "iid" else gen.markov_dna(n, seed=5)
    with P.Context() as c:
        for m in (n, n // 2):
            d_t = torch.from_numpy(t[:m].copy()).to(dev)
            d_sa = torch.empty(m, dtype=torch.int32, device=dev)
            P.psacak8_dev(d_t, d_sa, ctx=c)
            torch.cuda.synchronize()
            assert P.check8_dev(d_t, d_sa)  # the default context, not c
        reserved, peak = c.workspace()
        assert 0 < reserved <= peak <= P.workspace_bytes8(n), (reserved, peak)


def test_workspace32_within_estimate():
    """psacak32_raw (translation included) stays within psk_workspace_bytes32."""
    note("workspace32")
    import torch
    dev = torch.device("cuda:0")
    n = 1 << 24
    z = gen.zipf32(n, seed=3)
    with P.Context() as c:
        d_t = torch.from_numpy(z.view(np.int32)).to(dev)
        d_sa = torch.empty(n, dtype=torch.int32, device=dev)
        P.psacak32_raw_dev(d_t, d_sa, ctx=c)
        torch.cuda.synchronize()
        sa = d_sa.cpu().numpy().view(np.uint32)
        assert np.array_equal(np.sort(sa), np.arange(n, dtype=np.uint32))
        reserved, peak = c.workspace()
        assert 0 < reserved <= peak <= P.workspace_bytes32(n), (reserved, peak)


def test_budget_is_enforced():
    note("budget")
    with P.Context(hbm_budget=1 << 20) as c:
        with pytest.raises(P.PskError) as e:
            P.psacak(gen.iid_dna(1 << 22), c)
        assert e.value.code in (P.PSK_EBUDGET, P.PSK_ENOMEM)


def test_failed_call_leaves_workspace_reusable():
    # a budget that runs out partway through a device call: the call fails
    # with PSK_EBUDGET and leaves the context's workspace stack where it found
    # it, so a smaller call on the same context succeeds inside the chunks
    # already reserved (lib.rs:77-82: a call leaves no state behind)
    import torch
    note("budget failure then reuse")
    dev = torch.device("cuda:0")
    big = gen.markov_dna(32 << 20, seed=21)
    small = gen.markov_dna(1 << 20, seed=22)
    budget = 256 << 20
    with P.Context(hbm_budget=budget) as c:
        d_big = torch.from_numpy(big).to(dev)
        s_big = torch.empty(len(big), dtype=torch.int32, device=dev)
        with pytest.raises(P.PskError) as e:
            P.psacak8_dev(d_big, s_big, ctx=c)
        assert e.value.code == P.PSK_EBUDGET
        r1, _ = c.workspace()
        assert 0 < r1 <= budget
        d_small = torch.from_numpy(small).to(dev)
        s_small = torch.empty(len(small), dtype=torch.int32, device=dev)
        exp = oracle.sacak8(small)
        for _ in range(2):
            P.psacak8_dev(d_small, s_small, ctx=c)
            torch.cuda.synchronize()
            assert np.array_equal(s_small.cpu().numpy().view(np.uint32), exp)
            r2, _ = c.workspace()
            assert r2 == r1, (r1, r2)
        # the same through the host entry point and psacak32_raw
        with pytest.raises(P.PskError) as e:
            P.psacak(big, c)
        assert e.value.code == P.PSK_EBUDGET
        assert np.array_equal(P.psacak(small, c), exp)
        z = gen.zipf32(1 << 16, sigma=1 << 12, seed=4)
        assert np.array_equal(P.psacak32_raw(z, c), oracle.sacak32_raw(z))
        assert c.workspace()[0] <= budget


def test_c3_zipf_sigma_2_20_against_oracle(ctx):
    # the psacak32 configuration's own alphabet: Zipf(1.1) over 2^20 symbols
    # with 20% exact repeats of 16-4096 symbols (psacak32.rs:912-943
    # translation, then psacak32)
    note("C3 zipf32 n=2^22 sigma=2^20")
    z = gen.zipf32(1 << 22, sigma=1 << 20, s=1.1, repeat_frac=0.2, seed=20260103)
    got = P.psacak32_raw(z, ctx)
    assert np.array_equal(got, oracle.sacak32_raw(z))


def test_check32_dev_full_c3_text(ctx):
    # the full 2^28-symbol C3 text, verified on the device (permutation plus
    # the adjacent-pair rank condition), and planted mistakes caught
    import torch
    note("C3 zipf32 n=2^28 device check")
    n = 1 << 28
    z = gen.zipf32(n, sigma=1 << 20, s=1.1, repeat_frac=0.2, seed=20260103)
    dz = torch.from_numpy(z.view(np.int32)).to("cuda:0")
    del z
    ds = torch.empty(n, dtype=torch.int32, device="cuda:0")
    P.psacak32_raw_dev(dz, ds, n, ctx=ctx)
    torch.cuda.synchronize()
    assert P.check32_dev(dz, ds, n, ctx=ctx)
    a, b = ds[12345].item(), ds[12346].item()
    ds[12345], ds[12346] = b, a
    assert not P.check32_dev(dz, ds, n, ctx=ctx)
    ds[12345], ds[12346] = a, b
    ds[7] = ds[8].item()
    assert not P.check32_dev(dz, ds, n, ctx=ctx)
    del dz, ds
    torch.cuda.empty_cache()


def test_check32_dev_small_against_oracle_checker(ctx):
    import torch
    note("check32_dev small")
    rng = np.random.default_rng(3)
    for it in range(20):
        n = int(rng.integers(2, 3000))
        u = rng.integers(0, int(rng.choice([2, 5, 100, 1 << 20])), n).astype(np.uint32)
        sa = oracle.sacak32_raw(u)
        du = torch.from_numpy(u.view(np.int32)).to("cuda:0")
        ds = torch.from_numpy(sa.view(np.int32).copy()).to("cuda:0")
        assert P.check32_dev(du, ds, n, ctx=ctx)
        bad = sa.copy()
        i = int(rng.integers(0, n - 1))
        bad[i], bad[i + 1] = bad[i + 1], bad[i]
        ds = torch.from_numpy(bad.view(np.int32)).to("cuda:0")
        assert P.check32_dev(du, ds, n, ctx=ctx) == (oracle.check_sa32(u, bad) == 0)


def _oracle_reduced(text_level, width):
    """The oracle's translated reduced text of one level (stage 1, names,
    translation): the next level's text, and its number of names."""
    if width == 1:
        lms = oracle.stage1_8(text_level)
        k, red = oracle.name8(text_level, lms)
    else:
        lms = oracle.stage1_32(text_level)
        k, red = oracle.name32(text_level, lms)
    return k, red


@pytest.mark.parametrize("kind", ["markov", "iid", "uniform3", "runs"])
def test_level_intermediates_match_oracle(kind):
    # canonical intermediates of levels 0 and 1 (bit-exact): the per-level
    # counters (m, m1, mL, k, sum of LMS-substring lengths), which also feed
    # the bench's roofline model, and the translated reduced texts
    # (naming.rs:20-274) the recursion receives
    note(f"level intermediates {kind}")
    if kind == "markov":
        t = gen.markov_dna(1 << 22, seed=71)
    elif kind == "iid":
        t = gen.iid_dna(1 << 22, seed=72)
    elif kind == "uniform3":
        t = gen.uniform8(1 << 21, alphabet=3, seed=73)
    else:
        t = np.frombuffer(ADVERSARIAL["runs"] * 8, dtype=np.uint8)
    ref_rows = oracle.level_stats8(t)
    k0, red0 = _oracle_reduced(t, 1)
    with P.Context() as c:
        for level in (0, 1):
            c.set_capture(level)
            sa = P.psacak(t, c)
            rows = c.last_levels()
            cap = c.last_capture()
            assert np.array_equal(rows[0, :5], ref_rows[0]), (rows[0], ref_rows[0])
            assert rows[0, 3] == k0
            if level == 0:
                assert cap is not None and np.array_equal(cap, red0)
            elif kind in ("markov", "iid") or (rows.shape[0] > 1 and rows[1, 1] != 0):
                # level 1 was named (not sorted by prefix doubling)
                assert np.array_equal(rows[1, :5], ref_rows[1]), (rows[1], ref_rows[1])
                k1, red1 = _oracle_reduced(red0, 4)
                assert rows[1, 3] == k1
                assert cap is not None and np.array_equal(cap, red1)
            elif rows.shape[0] > 1:
                assert rows[1, 0] == ref_rows[1, 0]
                assert cap is None
        assert np.array_equal(sa, oracle.sacak8(t))


def test_level_intermediates_psacak32_zipf():
    note("level intermediates zipf32")
    z = gen.zipf32(1 << 20, sigma=1 << 16, s=1.1, repeat_frac=0.2, seed=74)
    tr = oracle.translate32(z)
    k0, red0 = _oracle_reduced(tr, 4)
    with P.Context() as c:
        c.set_capture(0)
        sa = P.psacak32_raw(z, c)
        cap = c.last_capture()
        rows = c.last_levels()
    assert rows[0, 3] == k0
    assert cap is not None and np.array_equal(cap, red0)
    assert np.array_equal(sa, oracle.sacak32(tr))


def test_server_under_a_binding_budget(ctx):
    # a budget that holds the large request's footprint but not much else:
    # admission must wait for memory (blocked checks), stay in arrival order,
    # keep the reserved device memory inside the budget, and every result
    # must equal the same text's SA computed alone
    note("server binding budget")
    rng = np.random.default_rng(23)
    small = [gen.markov_dna(int(n), seed=400 + i)
             for i, n in enumerate(gen.lengths(60, 10 << 10, 2 << 20, seed=44))]
    big = gen.iid_dna(64 << 20, seed=45)
    texts = small[:30] + [big] + small[30:]
    alone = [P.psacak(t, ctx) for t in texts]
    budget = int(P.Server.footprint(len(big)) * 1.15)
    with P.Server(hbm_budget=budget, workers=3) as srv:
        ids = []
        for i, t in enumerate(texts):
            ids.append(srv.submit(t) if i % 2 else srv.submit_into(t))
            if rng.random() < 0.3:
                import time
                time.sleep(0.001)
        for i, rid in enumerate(ids):
            sa, _ = srv.wait(rid, timeout_ms=120000)
            assert np.array_equal(sa, alone[i]), i
        st = srv.stats()
        assert srv.pending() == 0
    assert st["admitted"] == len(texts)
    assert st["order_violations"] == 0
    assert st["blocked"] > 0, st        # the budget did bind
    assert st["max_inflight"] <= budget, st
    assert st["peak_bytes"] <= budget, st
    for i in (0, 30, 45):
        assert np.array_equal(alone[i], oracle.sacak8(texts[i]))


def test_text_above_2_31_bytes(ctx):
    # the crate's domain runs to n < 2^32 - 1 (lib.rs:65,79): a text of
    # 2^31 + 2^20 bytes (level 1 then has about 6.4*10^8 symbols), verified on
    # the device (permutation + adjacent-pair rank condition) and by direct
    # comparison of sampled adjacent suffixes
    import torch
    n = (1 << 31) + (1 << 20)
    note(f"text of {n} bytes")
    t = gen.iid_dna(n, seed=31)
    dt = torch.from_numpy(t).to("cuda:0")
    ds = torch.empty(n, dtype=torch.int32, device="cuda:0")
    P.psacak8_dev(dt, ds, ctx=ctx)
    torch.cuda.synchronize()
    assert P.check8_dev(dt, ds, ctx=ctx)
    lv = ctx.last_levels()
    assert lv[0, 0] == n and lv.shape[0] >= 2
    rng = np.random.default_rng(5)
    idx = rng.integers(1, n, 256)
    sa_lo = ds[torch.from_numpy(idx - 1).to("cuda:0")].cpu().numpy().view(np.uint32)
    sa_hi = ds[torch.from_numpy(idx).to("cuda:0")].cpu().numpy().view(np.uint32)
    tb = memoryview(t)
    for a, b in zip(sa_lo, sa_hi):
        a, b = int(a), int(b)
        x, y = bytes(tb[a:a + 4096]), bytes(tb[b:b + 4096])
        assert x != y, (a, b)
        k = next(j for j in range(min(len(x), len(y))) if x[j] != y[j])
        assert x[k] < y[k], (a, b)
    del dt, ds
    torch.cuda.empty_cache()


def test_injected_allocation_failures_leave_contexts_usable():
    # every k-th workspace allocation fails (PSACAK_AMD_FAIL_ALLOC=k): each
    # entry point returns a status code (no abort, no wrong SA), and once the
    # injection stops the same context computes correct SAs again
    import torch
    note("fault injection")
    t = gen.markov_dna(6 << 20, seed=51)
    exp = oracle.sacak8(t)
    z = gen.zipf32(1 << 18, sigma=1 << 14, seed=52)
    expz = oracle.sacak32_raw(z)
    dt = torch.from_numpy(t).to("cuda:0")
    ds = torch.empty(len(t), dtype=torch.int32, device="cuda:0")
    failures = 0
    try:
        for k in (1, 2, 3, 5, 8):
            with P.Context() as c:
                os.environ["PSACAK_AMD_FAIL_ALLOC"] = str(k)
                for call in ("dev", "host", "raw32", "batch", "check"):
                    try:
                        if call == "dev":
                            P.psacak8_dev(dt, ds, ctx=c)
                        elif call == "host":
                            P.psacak(t, c)
                        elif call == "raw32":
                            P.psacak32_raw(z, c)
                        elif call == "batch":
                            P.batch8([t[:100000], t[100000:300000]], c)
                        else:
                            P.check8_dev(dt, ds, ctx=c)
                    except P.PskError as e:
                        assert e.code in (P.PSK_ENOMEM, P.PSK_EBUDGET), (k, call, e)
                        failures += 1
                os.environ["PSACAK_AMD_FAIL_ALLOC"] = "0"
                P.psacak8_dev(dt, ds, ctx=c)
                torch.cuda.synchronize()
                assert np.array_equal(ds.cpu().numpy().view(np.uint32), exp), k
                assert np.array_equal(P.psacak(t, c), exp), k
                assert np.array_equal(P.psacak32_raw(z, c), expz), k
    finally:
        os.environ.pop("PSACAK_AMD_FAIL_ALLOC", None)
    assert failures > 0
