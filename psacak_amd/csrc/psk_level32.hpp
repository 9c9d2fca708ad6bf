// psk_level32.hpp -- one recursion level on an integer (u32) text in
// bucket-pointer form (psacak32.rs:24-56): an L-type symbol is the head of
// its bucket, an S-type symbol the bucket's tail - 1.
//
// Induced sorting with a large alphabet cannot proceed bucket by bucket (the
// buckets number ~m/3), so each scan is computed by "chain sorting", which
// yields exactly the order the sequential scan produces:
//
//  * L-scan.  An L-type suffix p is placed in its bucket's L part in the order
//    of its successor's rank (the induce reads SA in order), so within a
//    bucket, L suffixes sort by (T[p], rank(p+1)); p+1 is either L-type
//    again or the chain's terminal: an LMS suffix (whose rank is its slot at
//    the bucket tail) or the sentinel (rank below everything).  Sorting the
//    L suffixes by these chain keys is prefix doubling over the chains:
//    labels start as bucket-level ranks (first slot of the bucket's L part),
//    terminals are unique, and each round sorts the tied items by
//    (label(p), label(p+h)), splits ties and doubles h.
//  * S-scan.  Same with S suffixes, whose chains end at an L-type suffix whose
//    final slot the L-scan has fixed.
//
// Labels are final slot + 1 (so the sentinel can be 0) under the
// "first slot of the group" convention, which keeps every group inside the
// slot range it will finally occupy; a round turns a group's label into
// label + (index of the subgroup's first item - index of the group's first
// item) in the sorted order.
#pragma once

#include "psk_lmsname.hpp"

namespace psk {

#ifndef PSK_CHAIN_SIGMA
#define PSK_CHAIN_SIGMA 96
#endif
#ifndef PSK_DOUBLING
#define PSK_DOUBLING 16
#endif
#ifndef PSK_CHAIN_SIGMA_SMALL
#define PSK_CHAIN_SIGMA_SMALL 96
#endif

struct LevelStat {
    u64 m, m1, mL, k, sumlen, csize;
};

// The first LMS position of a level goes to the call's deferred statistics
// (err[1 + depth]); the level's sum of LMS-substring lengths follows from it
// when the call ends, without a readback of its own.
inline void record_first_lms(Dev &d, const u32 *lms_text, u32 m1, int depth) {
    if (!m1 || depth >= 62) return;
    u32 *slot = d.err + 1 + depth;
    for_each(d, 1, [=] __device__(u64) { *slot = lms_text[0]; });
}

struct Runtime {
    Dev d;
    std::vector<LevelStat> levels;
    int max_rounds = 0;
    // rounds of direct LMS-substring sorting before induced sorting takes
    // over (long repeated substrings); 0 = always induced sorting
    int direct_rounds = 32;
    // a recursion level whose text has at least this fraction of distinct
    // symbols (in 1/1024) is sorted by prefix doubling instead of recursing;
    // 0 = always recurse (default from the autotuner's record)
    u32 doubling_distinct = PSK_DOUBLING;
    bool doubling_gave_up = false;  // in this call (reset by each entry point)
    int depth = 0;                  // recursion level being solved (traces)
    // capture_depth >= 0: the translated reduced text of that level (the
    // next level's text) is copied to the host as the level computes it
    int capture_depth = -1;
    std::vector<u32> captured;
    bool captured_ok = false;
    // round 1 of direct naming at level 0 through a dictionary of the
    // distinct keys (false: by sorting every key)
    bool dict_naming = true;
    // a byte text with at least this many distinct bytes takes its level-0
    // final induce as chain sorts (final_induce32) instead of one multi-split
    // step per bucket and scan; 0 = always bucket steps.  Texts of fewer than
    // CHAIN_SMALL_N bytes use chain_sigma_small (their bucket steps are
    // shorter, so launch costs weigh more).  The defaults come from the
    // autotuner's record (psacak_amd/tuning.json) through the build.
    u32 chain_sigma = PSK_CHAIN_SIGMA;
    u32 chain_sigma_small = PSK_CHAIN_SIGMA_SMALL;
    static constexpr u64 CHAIN_SMALL_N = 1ull << 26;
    // level 0 places its sorted LMS suffixes from 16-byte (position, window)
    // records gathered by reduced index (one random read each, 16 bytes per
    // LMS suffix of extra workspace), or, when false, by gathering the
    // position and then the window (two random reads, no extra workspace)
    bool lms_records = true;
};

// Workspace bounds: level 0 holds type/LMS bit vectors (n/4), the digit
// scratch (n), LMS lists and rank table (<= 4 x 2n), the placement records
// (16 bytes per LMS suffix, <= 8n) unless the context is lean, plus the
// recursion on <= n/2 symbols; each u32 level needs <= ~36 bytes per symbol.
// Bounded by the worst case m1 = n/2 at every level (measured peaks: 31-37
// bytes per input byte lean, 36-42 with records).
inline u64 levels_bytes(u64 n) {
    u64 total = 0, m = n;
    while (m > 3) {
        total += m * 36 + (64ull << 20);
        m /= 2;
    }
    return total + (64ull << 20);
}

inline u64 workspace_bytes32(u64 n) {
    // the translation of an arbitrary text into bucket-pointer form at the
    // first level adds about 4 bytes per symbol (measured: 81.7 bytes per
    // symbol for psacak32_raw on 2^28 Zipf symbols)
    return levels_bytes(n) + 4 * n;
}

inline u64 lean_bytes8(u64 n) {
    // bit vectors, digit cache, byte windows (8n), LMS lists and names
    return n / 4 + n + 17 * n + levels_bytes(n / 2) + (64ull << 20);
}

inline u64 workspace_bytes8(u64 n) { return lean_bytes8(n) + 8 * n; }

// positions p < n whose bit in wordfn(w) is set, ascending, into out.
template <class WF>
inline void positions_of(Dev &d, u64 n, WF wordfn, u32 *out) {
    bit_list(d, n, wordfn, out, nullptr);
}

// Sort items[0..A) (null: the items 0..A-1) by chain keys (see above); RK
// holds the labels of every position 0..m (m = sentinel) and receives the
// final labels of the items.
// If last_slot_labels, the items' initial group labels name the group's last
// slot + 1 instead of its first (S parts are known by their tail); the first
// round converts them, so later rounds use first-slot labels.
// The high half of a key names the item's group.  Round 1 uses the label
// itself, or, given rank (a dense rank of each position's symbol, k values),
// the rank; every later round uses the dense id of the item's group among the
// groups still being refined, with the groups' labels in a table.  Keys then
// carry bit_width(groups) + bit_width(m) bits instead of 2 bit_width(m), and
// the group is read with the item instead of gathered from RK.
// tails (with rank): number of symbols of rank <= c, so the bucket of rank c
// is [tails[c-1], tails[c]) and the initial labels are its head + 1 (L) or
// its tail (S, last-slot labels).
// giveup: return false (RK partly relabelled, SA partly written) when the
// ties look like long repeats: after round 2 still more than a sixteenth of
// the items tied and fewer than a quarter of round 1's ties resolved, or after
// round 4 more than an eighth tied.
// SA (optional): receives SA[label - 1] = p for every item as its label
// becomes final, so the items' SA entries need no pass of their own.
// runbits (optional, the S-type bits; only for a level's final L and S
// sorts, items = every position of one type): labels that no later round and
// no later sort reads are not written (ChainSink), using the items' run
// offsets carried in their keys while the key bits leave room.
// final (with rank, runbits and SA; 1 = the final L sort, 2 = the final S
// sort): items come as [the items whose successor's slot is already final,
// in that slot's order] ++ [the others, in any order].  Round 1 then sorts by
// (rank(p), rank(p+1), type of p+1) only: a stable sort keeps the first part
// in its order within equal keys, and those items leave the round as
// singletons (the L sort: successor an LMS suffix or the sentinel; the S
// sort: successor an L suffix).  The labels and ties that come out equal
// those of the full keys (rank(p), label(p+1)).  kpre (optional): the
// round-1 keys of items[0..npre), run offset included at RUN_KEY_SHIFT
// (the caller builds them where it has their symbols at hand).
inline bool chain_sort(Runtime &rt, u64 m, u32 *RK, const u32 *items, u32 A,
                       bool last_slot_labels, const u32 *rank = nullptr,
                       const u32 *tails = nullptr, u32 k = 0, bool giveup = false,
                       u32 *SA = nullptr, const u32 *runbits = nullptr, int final_mode = 0,
                       const u64 *kpre = nullptr, u32 npre = 0) {
    Dev &d = rt.d;
    if (A == 0) return true;
    const u32 A0 = A;
    u32 Aprev = A;
    ArenaScope mk(*d.a);
    u32 *act = d.a->get<u32>(A);
    u32 *vt = d.a->get<u32>(A);
    u32 *gid = d.a->get<u32>(A), *gid2 = d.a->get<u32>(A);
    const u64 gcap = std::max<u64>(rank ? k : 0, (u64)A / 2 + 1);
    u32 *glab = d.a->get<u32>(gcap), *glab2 = d.a->get<u32>(gcap);
    u64 *keys = d.a->get<u64>(A);
    u64 *kt = d.a->get<u64>(A);
    u32 *cnt = d.a->get<u32>(4);  // after a round: items and groups left, large-group tiles and items
    const int B = bit_width(m);
    int Bh = B;
    const bool byrank = rank != nullptr && tails != nullptr && k > 0;
    if (byrank) {
        u32 *gl = glab;
        const bool lsl = last_slot_labels;
        for_each(d, k, [=] __device__(u64 c) {
            gl[c] = lsl ? tails[c] : (c ? tails[c - 1] : 0u) + 1;
        });
        Bh = bit_width(k - 1);
    }
    const int fm = byrank && runbits ? final_mode : 0;
    // round-1 key of the final sorts: rank(p) | rank(p+1) + 1 (0: the
    // sentinel) | p+1 S-type | alone
    const int B1 = fm ? bit_width(k) + 2 : B;
    u64 h = 1;
    int rounds = 0;
    u32 G = 0;             // tie groups entering the round (after round 1)
    u32 NT = 0, LI = 0;    // ... the tiles and items of the large ones
    bool carried = false;  // the group ids hold run offsets
    bool skipped = false;  // a round left labels unwritten
    while (A > 0) {
        rounds++;
        const bool first = rounds == 1;
        const u64 hh = h, mm = m;
        const u32 *ac = act;
        const u32 *gi = gid;
        u32 *err = d.err;
        u32 *acw = act;
        // run offsets in this round's keys: the sorted bits end below them
        // (at most 7 radix passes) and the group ids leave room for them
        const bool cr = runbits && (first || carried) && 2 * h <= RUN_CAP && Bh + B <= 56 &&
                        A < (1u << 28);
        const bool prevc = carried, sit = last_slot_labels;
        const u32 *rb = runbits;
        const bool f1 = first && fm;
        const int fmm = fm, BB1 = B1;
        for_each(d, A, [=] __device__(u64 i) {
            // round 1 reads the caller's items and makes the working copy
            u32 p;
            if (first) {
                p = items ? items[i] : (u32)i;
                acw[i] = p;
            } else {
                p = ac[i];
            }
            u64 key;
            const u32 gv = first ? 0u : gi[i];
            if (f1 && i < npre) {
                key = kpre[i];
                if (!cr) key &= (1ull << RUN_KEY_SHIFT) - 1;
                keys[i] = key;
                return;
            }
            if (f1) {
                const u64 q = (u64)p + 1;
                u32 c2 = 0, ob = 0;
                if (q < mm) {
                    c2 = rank[q] + 1;
                    ob = getbit(rb, q);
                }
                const u32 so = fmm == 1 ? (q < mm ? ob : 1u) : (ob ^ 1u);
                key = ((u64)rank[p] << BB1) | ((u64)c2 << 2) | ((u64)ob << 1) | so;
            } else {
                u64 q = (u64)p + hh;
                u32 lo = 0;
                if (q <= mm) lo = RK[q];
                else atomicOr(err, 16u);
                u32 hi = first ? (byrank ? rank[p] : RK[p]) : (prevc ? gv & RUN_GID_MASK : gv);
                key = ((u64)hi << B) | (u64)lo;
            }
            if (cr)
                key |= (u64)(first ? run_offset(rb, p, sit) : gv >> RUN_GID_SHIFT)
                       << RUN_KEY_SHIFT;
            keys[i] = key;
        });
        // round 1 sorts everything; later rounds keep every tie group in its
        // place and sort it by the label bits alone
        bool sw = first ? radix_sort_kv(d, A, keys, act, kt, vt, Bh + (f1 ? B1 : B))
                        : group_sort_kv(d, A, G, gid, prevc ? RUN_GID_MASK : ~0u, keys, act, kt,
                                        vt, B, NT, LI);
        const u64 *K = sw ? kt : keys;
        const u32 *V = sw ? vt : act;
        u32 *nxt = sw ? act : vt;  // the buffer not holding V
        const u64 AA = A;
        int pr = -1;
        {
            const bool lastconv = last_slot_labels && first;
            ArenaScope m3(*d.a);
            const u32 nt = ceil_div(AA, REL_TILE);
            u32 *tg = d.a->get<u32>(nt), *ts = d.a->get<u32>(nt);
            u32 *tk = d.a->get<u32>(nt), *te = d.a->get<u32>(nt);
            u32 *tq = d.a->get<u32>(nt);
            RelIn in{K, AA, f1 ? B1 : B, false, cr ? (1ull << RUN_KEY_SHIFT) - 1 : ~0ull, f1};
            k_rel_reduce<<<nt, BLOCK, 0, d.s>>>(in, tg, ts, tk, te, tq);
            PSK_LAUNCH_CHECK();
            k_rel_tiles<<<lastconv ? 5 : 4, BLOCK, 0, d.s>>>(tg, ts, tk, tq, te, nt, cnt);
            PSK_LAUNCH_CHECK();
            const u32 *gl = (first && !byrank) ? nullptr : glab;
            // 28 B per item; the SA entries of the items that finish are
            // added once their count is known
            pr = d.pbegin("chain_relabel", 28.0 * AA);
            const ChainSink sink{RK, lastconv, SA, mm, err, cr ? (u32)(2 * h) : 0u,
                                 last_slot_labels, !cr && skipped};
            k_rel_apply<<<nt, BLOCK, 0, d.s>>>(in, V, sink, nxt,
                                               tg, ts, tk, te,
                                               tq, lastconv, mm, gl, gid2, glab2, err, cnt);
            PSK_LAUNCH_CHECK();
            d.pend(pr);
            m3.reset();
        }
        u32 ag[4];
        d.read(ag, cnt, sizeof(ag));
        if (SA) d.padd(pr, 4.0 * (double)(AA - ag[0]));
        A = ag[0];
        if (nxt != act) std::swap(act, vt);
        std::swap(gid, gid2);
        std::swap(glab, glab2);
        if (d.trace) {
            char what[64];
            snprintf(what, sizeof(what), "level %d %s %s", rt.depth, giveup ? "doubling" : SA ? "final" : "stage1",
                     last_slot_labels ? "S" : "L");
            trace_round(d, what, rounds, AA, Bh + B, gid, ag[0], cr ? RUN_GID_MASK : ~0u, ag[1]);
        }
        Bh = bit_width(ag[1] ? ag[1] - 1 : 0);
        G = ag[1];
        NT = ag[2];
        LI = ag[3];
        carried = cr;
        skipped = skipped || cr;
        h *= 2;
        if (h > m + 1 && A > 0) fail(PSK_EDEVICE, "chain sort did not converge");
        if (giveup && A > 0 &&
            ((rounds == 2 && A > A0 / 16 && (u64)A * 4 > (u64)Aprev * 3) ||
             (rounds == 4 && A > A0 / 8))) {
            mk.reset();
            return false;
        }
        Aprev = A;
    }
    rt.max_rounds = std::max(rt.max_rounds, rounds);
    mk.reset();
    return true;
}

// Initial labels for the scans of a bucket-pointer text s[0..m): an L
// position's L-scan label is its bucket head + 1 and an S position's S-scan
// label the last slot of its bucket's S part + 1 (tail - 1 is the S-type
// symbol itself; chain_sort is told to convert them): s[x] + 1 either way.
// labels_L writes s[x] + 1 at every position (the LMS labels of the L-scan
// then go over the LMS positions) and RK[m] = 0: a plain stream, four
// positions per thread with 16-byte accesses.
inline void labels_L(Dev &d, const u32 *s, u64 m, const u32 *, u32 *RK) {
    if (((((uintptr_t)s) | ((uintptr_t)RK)) & 15u) != 0) {
        // (a caller's text off 16-byte alignment) one position per thread
        for_each(d, m + 1, [=] __device__(u64 x) { RK[x] = x < m ? s[x] + 1 : 0u; });
        return;
    }
    const u64 nq = m / 4;
    for_each(d, nq + 1, [=] __device__(u64 q) {
        if (q < nq) {
            const uint4 v = *(const uint4 *)(s + 4 * q);
            *(uint4 *)(RK + 4 * q) = make_uint4(v.x + 1, v.y + 1, v.z + 1, v.w + 1);
        } else {
            for (u64 x = 4 * nq; x < m; x++) RK[x] = s[x] + 1;  // at most 3
            RK[m] = 0;
        }
    });
}

// The S-scan labels, after labels_L and the L-scan: the L sort writes L
// positions only and the LMS labels went over the LMS positions, so only
// those get s[x] + 1 back (four positions per thread, from the type bits).
inline void labels_S(Dev &d, const u32 *s, u64 m, const u32 *sbits, u32 *RK) {
    const u64 nq = (m + 3) / 4;
    for_each(d, nq, [=] __device__(u64 q) {
        const u64 x = 4 * q;
        const u32 w = (u32)(x >> 5), b = (u32)(x & 31);
        const u32 S = sbits[w] >> b;
        const u32 prev = b ? sbits[w] >> (b - 1) : (w ? sbits[w - 1] >> 31 : 1u);
        u32 lms = S & ~((S << 1) | (prev & 1u)) & 15u;  // S-type after an L-type position
        if (m - x < 4) lms &= (1u << (m - x)) - 1u;
        while (lms) {
            const u32 j = (u32)__ffs(lms) - 1;
            lms &= lms - 1;
            RK[x + j] = s[x + j] + 1;
        }
    });
}

// Labels of LMS suffixes placed at their bucket tails in the order of list
// (grouped by symbol value, ascending, e.g. sorted by it): the last of a group
// gets slot s[x] = tail-1, the one before it tail-2, ...  vals: the symbol of
// each list item when the caller has it (sorted keys), else nullptr (one
// gather per item).  grp: scratch of m+1 u32.
// need (optional, one key per list item): only items whose key carries a
// run offset >= 1 at RUN_KEY_SHIFT get their label written (the final L sort
// with short round-1 keys reads an LMS label only through a chain of at
// least two L positions before it).
inline void labels_lms_at_tails(Dev &d, const u32 *s, const u32 *list, const u64 *vals, u64 cnt,
                                u32 *RK, u32 *grp, const u64 *need = nullptr) {
    if (cnt == 0) return;
    u32 *err = d.err;
    ArenaScope mk(*d.a);
    u32 *vb = vals ? nullptr : d.a->get<u32>(cnt);
    // group ends: a wave compares neighbours by shuffle
    for_each_waves(d, cnt, [=] __device__(u64 j, bool live) {
        u32 v = live ? (vals ? (u32)vals[j] : s[list[j]]) : 0u;
        u32 vn = __shfl_down(v, 1, WAVE);
        if (lane_id() == WAVE - 1 && j + 1 < cnt) vn = vals ? (u32)vals[j + 1] : s[list[j + 1]];
        if (live) {
            if (vb) vb[j] = v;
            if (j + 1 == cnt || vn != v) grp[v] = (u32)j;
        }
    });
    for_each(d, cnt, [=] __device__(u64 j) {
        if (need && (need[j] >> RUN_KEY_SHIFT) == 0) return;
        u32 v = vals ? (u32)vals[j] : vb[j];
        u32 back = grp[v] - (u32)j;
        if (back <= v) RK[list[j]] = v - back + 1;
        else atomicOr(err, 256u);
    });
    mk.reset();
}

void level32(Runtime &rt, const u32 *s, u64 m, const u32 *sbits, u32 *SA, int depth,
             const u32 *rank = nullptr, const u32 *tails = nullptr, u32 k = 0);

// LMS positions in the order of their stage-1 labels (distinct slots
// 1..m): each LMS position x marks its slot with x + 1 in a zeroed array of
// m words, then one compaction keeps the marks in slot order.  lms_text is in
// text order, so the label reads are monotone; only the m1 mark writes are
// scattered.
inline void extract_lms_sorted(Dev &d, u64 m, const u32 *RK, const u32 *lms_text, u32 m1,
                               u32 *mark, u32 *lms_sorted) {
    u32 *err = d.err;
    d.zero(mark, m * sizeof(u32));
    for_each(d, m1, [=] __device__(u64 r) {
        u32 x = lms_text[r];
        u32 l = RK[x];
        if (l >= 1 && l <= m) mark[l - 1] = x + 1;
        else atomicOr(err, 64u);
    });
    scan_exclusive(
        d, m, [=] __device__(u64 i) { return mark[i] ? 1u : 0u; },
        [=] __device__(u64 i, u32 off, u32 v) {
            if (v) lms_sorted[off] = mark[i] - 1;
        },
        nullptr);
}

// The reduced problem of a level from its names (reduced, tails, k): types,
// bucket-pointer translation and the next level, or the inverse when every
// name is distinct.  SA1 receives the sorted reduced indices (indices into
// the level's text-order LMS list): the level maps them to positions itself,
// through records that cost one random read per entry.
inline void recurse_reduced(Runtime &rt, const u32 *reduced, const u32 *tails, u32 k, u32 m1,
                            u32 *SA1, int depth) {
    Dev &d = rt.d;
    ArenaScope mk(*d.a);
    auto capture = [&](const u32 *text1) {
        if (rt.capture_depth != depth) return;
        rt.captured.resize(m1);
        PSK_HIP(hipMemcpyAsync(rt.captured.data(), text1, (u64)m1 * sizeof(u32),
                               hipMemcpyDeviceToHost, d.s));
        PSK_HIP(hipStreamSynchronize(d.s));
        rt.captured_ok = true;
    };
    if (k < m1) {
        u32 *sb = d.a->get<u32>(((u64)m1 + 31) / 32 + 1);
        compute_types(d, reduced, m1, sb);
        u32 *s1 = d.a->get<u32>(m1);
        translate_names(d, reduced, s1, m1, sb, tails);
        capture(s1);
        // the names stay as the dense symbol ranks of the next level
        level32(rt, s1, m1, sb, SA1, depth + 1, reduced, tails, k);
    } else {
        // every name distinct: the bucket-pointer form is the names themselves
        capture(reduced);
        inverse_names(d, reduced, m1, SA1);
    }
    mk.reset();
}

// Induced path: names from the sorted LMS substrings, then the reduced problem.
template <class C>
inline void solve_reduced(Runtime &rt, const C *T, const u32 *lms_sorted, const u32 *lms_text,
                          const u32 *lmsbits, u64 nwords, u32 *NP, u32 m1, u32 *SA1, int depth,
                          LevelStat &st, const u32 *EB = nullptr) {
    Dev &d = rt.d;
    if (m1 == 0) return;
    ArenaScope mk(*d.a);
    u32 *reduced = d.a->get<u32>(m1);
    u32 *tails = d.a->get<u32>(m1);
    u8 *diff = d.a->get<u8>(m1);
    u32 *dk = d.a->get<u32>(1);
    u32 k = name_lms(d, T, lms_sorted, lms_text, lmsbits, nwords, NP, m1, diff, reduced, tails, dk,
                     EB);
    st.k = k;
    rt.levels[depth] = st;
    recurse_reduced(rt, reduced, tails, k, m1, SA1, depth);
    mk.reset();
}

// Direct path: names by sorting the LMS substrings, then the reduced problem.
// Returns false if the substrings need more rounds (nothing is written).
template <class RF>
inline bool solve_reduced_direct(Runtime &rt, RF rk, u64 sigma, u64 m, const u32 *lms_text,
                                 u32 m1, u32 *SA1, int depth, LevelStat &st,
                                 LmsEnds ends = LmsEnds{}) {
    Dev &d = rt.d;
    if (m1 == 0) return true;
    if (rt.direct_rounds == 0) return false;
    ArenaScope mk(*d.a);
    u32 *reduced = d.a->get<u32>(m1);
    u32 *tails = d.a->get<u32>(m1);
    u32 k = 0;
    // a byte text's LMS substrings take few distinct values: try the
    // dictionary first there
    bool named;
    if constexpr (std::is_same<RF, ByteRanks>::value) {
        // keys come out of the packed digit stream of the text, freed before
        // the recursion
        ArenaScope mp(*d.a);
        const PackedRanks pk = pack_ranks(d, rk, m, sigma);
        named = (rt.dict_naming && name_lms_dict(d, pk, sigma, m, lms_text, m1, reduced, tails,
                                                 &k, rt.direct_rounds, ends)) ||
                name_lms_direct(d, pk, sigma, m, lms_text, m1, reduced, tails, &k,
                                rt.direct_rounds, ends);
        mp.reset();
    } else {
        named = name_lms_direct(d, rk, sigma, m, lms_text, m1, reduced, tails, &k,
                                rt.direct_rounds, ends);
    }
    if (!named) {
        mk.reset();
        return false;
    }
    if (d.trace) {
        char what[48];
        snprintf(what, sizeof(what), "level %d naming", depth);
        d.phase(what);
    }
    st.k = k;
    rt.levels[depth] = st;
    recurse_reduced(rt, reduced, tails, k, m1, SA1, depth);
    mk.reset();
    return true;
}

// One psacak32 level: s[0..m) in bucket-pointer form with its S-type bits;
// SA receives the level's SA (positions of s).  rank, tails (optional):
// dense rank of every symbol of s (k distinct values) and the cumulative
// symbol counts per rank.
// Long exact repeats in a level's text (symbols rank[0..m)): the share of
// 32-symbol windows that occur more than once, estimated on content-defined
// anchors (positions whose 8-symbol hash is 0 mod 16, so the copies of a
// repeat are anchored alike).  Anchor hashes are sorted and neighbours
// compared.  Returns duplicates / anchors (0 when there are none).
constexpr int REP_WIN = 32;
__device__ __forceinline__ u32 rep_hash8(const u32 *rank, u64 i) {
    u32 h = 0x243F6A88u;
    #pragma unroll
    for (int j = 0; j < 8; j++) h = (h ^ rank[i + j]) * 0x9E3779B1u;
    return h ^ (h >> 15);
}
inline double repeat_share(Dev &d, const u32 *rank, u64 m) {
    if (m < 4 * (u64)REP_WIN) return 0.0;
    ArenaScope mk(*d.a);
    const u64 np = m - REP_WIN + 1;
    u32 *ctr = d.a->get<u32>(2);
    d.zero(ctr, 2 * sizeof(u32));
    const u64 cap = np / 8 + 1024;  // anchors are ~1/16 of the positions
    u64 *hk = d.a->get<u64>(cap), *hk2 = d.a->get<u64>(cap);
    u32 *hv = d.a->get<u32>(cap), *hv2 = d.a->get<u32>(cap);
    // anchor bits (one ballot per 64 positions), then the anchors as a list
    // (at most cap of them) and their window hashes in a dense pass
    // (32-bit multiplicative hashing, native multiplies; no shared counter)
    const u64 nwp = (np + 31) / 32;
    u32 *af = d.a->get<u32>(nwp + 2);
    for_each_waves(d, np, [=] __device__(u64 i, bool live) {
        const bool a = live && (rep_hash8(rank, i) >> 28) == 0;
        const u64 b = ballot(a);
        if ((lane_id() & 31) == 0 && (i >> 5) < nwp) af[i >> 5] = (u32)(b >> (lane_id() & 32));
    });
    u32 *alist = d.a->get<u32>(cap);
    const u64 cc = cap;
    bit_list_f(
        d, np, [=] __device__(u64 w) { return af[w]; },
        [=] __device__(u64 j, u32 p) {
            if (j < cc) alist[j] = p;
        },
        ctr);
    u32 na = d.read_u32(ctr);
    if (na > cap) na = (u32)cap;  // (a skewed text; the estimate uses the first cap)
    for_each(d, na, [=] __device__(u64 j) {
        const u64 i = alist[j];
        u32 h1 = 0x85A308D3u, h2 = 0x13198A2Eu;
        #pragma unroll 4
        for (int q = 0; q < REP_WIN; q++) {
            const u32 r = rank[i + q];
            h1 = (h1 ^ r) * 0x85EBCA6Bu;
            h2 = (h2 + r) * 0xC2B2AE35u;
            h2 ^= h2 >> 13;
        }
        hk[j] = ((u64)h1 << 32) | h2;
        hv[j] = 0;
    });
    double share = 0.0;
    if (na > 1) {
        bool sw = radix_sort_kv(d, na, hk, hv, hk2, hv2, 64);
        const u64 *K = sw ? hk2 : hk;
        const u64 nn = na;
        u32 *dup = ctr + 1;
        scan_exclusive(
            d, na,
            [=] __device__(u64 i) {
                return ((i > 0 && K[i - 1] == K[i]) || (i + 1 < nn && K[i + 1] == K[i])) ? 1u : 0u;
            },
            [=] __device__(u64, u32, u32) {}, dup);
        share = (double)d.read_u32(dup) / (double)na;
    }
    mk.reset();
    return share;
}

// SA of a level by prefix doubling over all its suffixes: every position is an
// item, its initial label is the first slot of its symbol's bucket (+1), the
// sentinel (label 0) is the only terminal, and chain_sort's rounds compare
// (label(p), label(p+h)) with h doubling.  A position whose symbol occurs
// once is final from the start (its bucket is its slot): only the others go
// through the rounds (almost none in the deep levels of texts over large
// alphabets, whose names are nearly all distinct).  A suffix that still ties after a
// round has its first 2h symbols in common with another and no sentinel
// among them, so p + h stays <= m.  When almost every symbol of the level is
// distinct (the names of a deep level), few suffixes tie for long and this
// costs less than naming, recursing and inducing.  Long repeats keep many
// suffixes tied round after round (prefix doubling is then O(m log m)): when
// the ties stop shrinking (chain_sort's give-up rule) this returns false,
// with SA untouched, and the level recurses instead.
inline bool level32_doubling(Runtime &rt, u64 m, u32 *SA, const u32 *rank, const u32 *tails,
                             u32 k) {
    Dev &d = rt.d;
    ArenaScope mk(*d.a);
    u32 *RK = d.a->get<u32>(m + 1);
    const u64 mm = m, nw = (m + 31) / 32;
    // initial labels; a position whose symbol occurs once is final already
    // (its SA entry is its bucket's slot), the others are flagged as items
    u32 *flag = d.a->get<u32>(nw + 2);
    u32 *cnt = d.a->get<u32>(1);
    for_each_waves(d, m + 1, [=] __device__(u64 x, bool live) {
        bool item = false;
        if (live) {
            if (x == mm) {
                RK[x] = 0;
            } else {
                const u32 c = rank[x];
                const u32 h = c ? tails[c - 1] : 0u;
                RK[x] = h + 1;
                if (tails[c] - h == 1) SA[h] = (u32)x;
                else item = true;
            }
        }
        const u64 b = ballot(item);
        if ((lane_id() & 31) == 0 && (x >> 5) < nw) flag[x >> 5] = (u32)(b >> (lane_id() & 32));
    });
    u32 *items = d.a->get<u32>(m);
    bit_list(d, m, [=] __device__(u64 w) { return flag[w]; }, items, cnt);
    const u32 A = d.read_u32(cnt);
    if (!chain_sort(rt, m, RK, items, A, false, rank, tails, k, true, SA)) {
        mk.reset();
        return false;
    }
    mk.reset();
    return true;
}

// Final induce of a level from its sorted LMS suffixes (psacak32.rs:103-121,
// then the L-scan and the S-scan as chain sorts): s[0..m) in bucket-pointer
// form with its S-type bits, rank/tails/k its dense symbol ranks (rank may be
// null: no short round-1 keys), lms_text its m1 LMS positions in text order,
// SA1 their sorted reduced indices, Litems/Sitems its L and S positions in
// ascending order.  RK and cnt: scratch of m + 1 words.  Writes all of SA.
inline void final_induce32(Runtime &rt, const u32 *s, u64 m, const u32 *sbits, const u32 *rank,
                           const u32 *tails, u32 k, const u32 *lms_text, u32 m1, const u32 *SA1,
                           const u32 *Litems, u32 nL, const u32 *Sitems, u32 nS, u32 *RK,
                           u32 *cnt, u32 *SA) {
    Dev &d = rt.d;
    ArenaScope mk(*d.a);
    labels_L(d, s, m, sbits, RK);
    // The final L sort's round 1 takes short keys (chain_sort, final) when
    // the ranks leave room for a run offset beside them: its items are the
    // predecessors of the sorted LMS suffixes in that order, then the L items
    // whose successor is L or the sentinel.  cnt is free once the LMS labels
    // are placed and holds that item order.
    const bool shortL = rank && m1 && (u64)k < (1ull << 27);
    u32 *ordL = cnt;
    u64 *kpre = shortL ? d.a->get<u64>(m1) : nullptr;
    const int B1 = bit_width(k) + 2;
    // sorted LMS suffixes are grouped by first symbol already
    if (m1) {
        // SA1 holds reduced indices: one record per LMS position in text order
        // (sequential reads), then one record gather per sorted entry gives
        // the list and its symbols.  Records are (symbol, position), or with
        // the short L keys (position, symbol, rank of the predecessor, rank + 1
        // | the predecessor's L-run offset << 27)
        ArenaScope m2(*d.a);
        u64 *vals = d.a->get<u64>(m1);
        u32 *list = d.a->get<u32>(m1 + 1);
        u32 *err = d.err;
        const u32 mm1 = m1;
        if (shortL) {
            uint4 *R = d.a->get<uint4>(m1);
            const u32 *sb = sbits;
            for_each(d, m1, [=] __device__(u64 p) {
                const u32 x = lms_text[p];  // an LMS position: x >= 1, x-1 L-type
                const u32 off = run_offset(sb, x - 1, false);
                R[p] = make_uint4(x, s[x], rank[x - 1], (rank[x] + 1) | (off << 27));
            });
            const int b1 = B1;
            for_each(d, m1, [=] __device__(u64 j) {
                const u32 p = SA1[j];
                uint4 r = make_uint4(0, 0, 0, 0);
                if (p < mm1) r = R[p];
                else atomicOr(err, 512u);
                list[j] = r.x;
                vals[j] = r.y;
                kpre[j] = ((u64)r.z << b1) | ((u64)(r.w & ((1u << 27) - 1)) << 2) | 3ull |
                          ((u64)(r.w >> 27) << RUN_KEY_SHIFT);
            });
        } else {
            u64 *R = d.a->get<u64>(m1);
            for_each(d, m1, [=] __device__(u64 p) {
                const u32 x = lms_text[p];
                R[p] = ((u64)s[x] << 32) | x;
            });
            for_each(d, m1, [=] __device__(u64 j) {
                const u32 p = SA1[j];
                u64 r = 0;
                if (p < mm1) r = R[p];
                else atomicOr(err, 512u);
                list[j] = (u32)r;
                vals[j] = r >> 32;
            });
        }
        labels_lms_at_tails(d, s, list, vals, m1, RK, cnt, kpre);
        if (shortL) for_each(d, m1, [=] __device__(u64 j) { ordL[j] = list[j] - 1; });
        m2.reset();
    }
    // every position is an L or an S item: the two sorts write the whole SA
    if (shortL) {
        // the L positions whose successor is L-type or the sentinel, from the
        // type bits (positions from m on read as L-type)
        const u64 nw = (m + 31) / 32;
        bit_list(
            d, m,
            [=] __device__(u64 w) {
                const u32 S = sbits[w], Sn = w + 1 < nw ? sbits[w + 1] : 0u;
                return ~S & ~((S >> 1) | (Sn << 31));
            },
            ordL + m1, nullptr);
        chain_sort(rt, m, RK, ordL, nL, false, rank, tails, k, false, SA, sbits, 1, kpre, m1);
    } else {
        chain_sort(rt, m, RK, Litems, nL, false, rank, tails, k, false, SA, sbits);
    }
    labels_S(d, s, m, sbits, RK);
    chain_sort(rt, m, RK, Sitems, nS, true, rank, tails, k, false, SA, sbits);
    mk.reset();
}

inline void level32(Runtime &rt, const u32 *s, u64 m, const u32 *sbits, u32 *SA, int depth,
                    const u32 *rank, const u32 *tails, u32 k) {
    Dev &d = rt.d;
    if (m == 0) return;
    rt.depth = depth;
    if (m <= 3) {
        k_naive_small<u32><<<1, 1, 0, d.s>>>(s, (u32)m, SA);
        PSK_LAUNCH_CHECK();
        return;
    }
    if (rank && tails && k && rt.doubling_distinct && !rt.doubling_gave_up &&
        (u64)k * 1024 >= (u64)rt.doubling_distinct * m) {
        if (repeat_share(d, rank, m) >= 1.0 / 16 ||
            !level32_doubling(rt, m, SA, rank, tails, k)) {
            // the repeats that kept this level's suffixes tied recur below
            rt.doubling_gave_up = true;
        } else {
            if (d.trace) {
                char what[48];
                snprintf(what, sizeof(what), "level %d doubling", depth);
                d.phase(what);
            }
            if ((int)rt.levels.size() <= depth) rt.levels.resize(depth + 1);
            rt.levels[depth] = LevelStat{m, 0, 0, k, 0, 4};
            rt.levels.resize(depth + 1);
            return;
        }
    }
    ArenaScope mk(*d.a);
    const u64 nwords = (m + 31) / 32;
    u32 *lmsbits = d.a->get<u32>(nwords + 1);
    u32 *ctr = d.a->get<u32>(2);
    d.zero(ctr, 2 * sizeof(u32));
    k_lms32<<<std::min<u32>(ceil_div(nwords, BLOCK), 2048), BLOCK, 0, d.s>>>(sbits, lmsbits, nwords,
                                                                           ctr);
    PSK_LAUNCH_CHECK();
    // L item list (ascending positions); then the LMS and L counts in one read
    u32 *Litems = d.a->get<u32>(m);
    u32 *dcnt = ctr + 1;
    bit_list(d, m, [=] __device__(u64 w) { return ~sbits[w]; }, Litems, dcnt);
    u32 cts[2];
    d.read(cts, ctr, sizeof(cts));
    const u32 m1 = cts[0], nL = cts[1];
    const u32 nS = (u32)m - nL;
    u32 *lms_text = d.a->get<u32>(m1 + 1);
    bits_to_positions(d, lmsbits, nwords, lms_text);
    u32 *Sitems = Litems + nL;
    positions_of(d, m, [=] __device__(u64 w) { return sbits[w]; }, Sitems);

    LevelStat st{m, m1, nL, 0, 0, 4};
    if ((int)rt.levels.size() <= depth) rt.levels.resize(depth + 1);
    rt.levels[depth] = st;

    u32 *RK = d.a->get<u32>(m + 1);
    u32 *cnt = d.a->get<u32>(m + 1);

    // ---- LMS-substring names and the reduced problem ------------------------
    record_first_lms(d, lms_text, m1, depth);
    rt.levels[depth] = st;
    u32 *SA1 = d.a->get<u32>(m1 + 1);
    // by sorting the substrings: symbols ranked by the names of the level
    // above, else by their bucket-pointer values (order- and equality-
    // preserving: types inside an LMS substring follow from its symbols)
    if (d.trace) {
        char what[48];
        snprintf(what, sizeof(what), "level %d types, LMS, item lists", depth);
        d.phase(what);
    }
    const bool named =
        rank ? solve_reduced_direct(rt, WordRanks{rank}, k, m, lms_text, m1, SA1, depth, st)
             : solve_reduced_direct(rt, WordRanks{s}, m, m, lms_text, m1, SA1, depth, st);
    if (!named) {
        // induced stage 1 (long repeated LMS substrings): LMS characters at
        // their bucket tails, in arbitrary order (psacak32.rs:60-99)
        labels_L(d, s, m, sbits, RK);
        {
            // LMS characters grouped by symbol (stable: text order inside a group)
            ArenaScope m2(*d.a);
            u64 *k0 = d.a->get<u64>(m1), *k1 = d.a->get<u64>(m1);
            u32 *v0 = d.a->get<u32>(m1), *v1 = d.a->get<u32>(m1);
            for_each(d, m1, [=] __device__(u64 r) {
                k0[r] = s[lms_text[r]];
                v0[r] = lms_text[r];
            });
            bool sw = radix_sort_kv(d, m1, k0, v0, k1, v1, bit_width(m));
            labels_lms_at_tails(d, s, sw ? v1 : v0, sw ? k1 : k0, m1, RK, cnt);
            m2.reset();
        }
        chain_sort(rt, m, RK, Litems, nL, false, rank, tails, k);
        labels_S(d, s, m, sbits, RK);
        chain_sort(rt, m, RK, Sitems, nS, true, rank, tails, k);
        u32 *lms_sorted = d.a->get<u32>(m1 + 1);
        extract_lms_sorted(d, m, RK, lms_text, m1, SA, lms_sorted);  // SA is scratch here
        u32 *NP = d.a->get<u32>(m / 2 + 1);
        solve_reduced(rt, s, lms_sorted, lms_text, lmsbits, nwords, NP, m1, SA1, depth, st);
    }

    // ---- final induce from sorted LMS suffixes (psacak32.rs:103-121) -------
    rt.depth = depth;
    if (d.trace) {
        char what[48];
        snprintf(what, sizeof(what), "level %d recursion below", depth);
        d.phase(what);
    }
    final_induce32(rt, s, m, sbits, rank, tails, k, lms_text, m1, SA1, Litems, nL, Sitems, nS, RK,
                   cnt, SA);
    if (d.trace) {
        char what[48];
        snprintf(what, sizeof(what), "level %d final induce", depth);
        d.phase(what);
    }
    mk.reset();
}

}  // namespace psk
