// psk_induce8.hpp -- induced sorting for byte texts (level 0), bucket by
// bucket, every step data-parallel over the whole GPU.
//
// Reference semantics: nonpar_induce_sort psacak8.rs:127-190 (the serial
// specification; par_induce_sort :201-344 is its CPU-pipelined form).
//
// L-scan (forward).  The read order is the SA order.  Bucket c's L part
// [head_c, head_c + nL_c) is a FIFO queue: it starts with the entries induced
// from buckets < c (in read order), and every entry x whose predecessor has
// the same byte (T[x-1] == c) appends x-1.  The queue therefore grows in
// generations: generation g+1 is the stable compaction of generation g by
// "T[x-1] == c", mapped x -> x-1, written right after it.  Once the queue is
// full, the bucket's read sequence (L part, then its LMS part at the tail) is
// final, and every x with T[x-1] > c sends x-1 to the next free slot of
// bucket T[x-1]'s L part: a stable multi-split of the sequence by T[x-1].
//
// S-scan (backward) mirrors it: bucket c's S part is filled from the tail
// downwards (generations of "T[x-1] == c" in read order), then the read
// sequence S part (tail -> down) followed by the L part (top -> down) sends
// every x with T[x-1] < c to the S part of bucket T[x-1].
//
// Every read position is final before it is read, so no empty marks are
// needed and the S parts are rebuilt from scratch (the LMS entries placed for
// the L-scan are overwritten, as in the reference).
#pragma once

#include "psk_types.hpp"

namespace psk {

struct Buckets8 {
    u32 cnt[256], nL[256], nLMS[256];
    u64 head[257];
};

// queue position of read index j: base + dir * j
struct QPos {
    u64 base;
    i64 dir;
    __device__ __forceinline__ u64 operator()(u64 j) const { return (u64)((i64)base + dir * (i64)j); }
};

// Predecessor of suffix x: none for x = 0 and, in a batch of concatenated
// texts, none at a text's first position (start bit vector SB).  DB is the
// "decided" mask (bit p: T[p] != T[p+1], or p ends a text).
struct Pred {
    const u8 *T;
    const u32 *SB;
    const u32 *DB;
    __device__ __forceinline__ bool has(u32 x) const { return x > 0 && !(SB && getbit(SB, x)); }
    __device__ __forceinline__ u32 byte(u32 x) const { return has(x) ? (u32)T[x - 1] : 0u; }
    // number of positions before x holding T[x] in an unbroken run: the gap
    // from x-1 down to the nearest decided position (a text start is preceded
    // by a text end, which is decided)
    __device__ __forceinline__ u32 run_before(u32 x) const {
        if (x == 0) return 0;
        u32 q = x - 1;
        u32 w = q >> 5;
        u32 m = DB[w] & (0xFFFFFFFFu >> (31 - (q & 31)));
        while (!m) {
            if (w == 0) return x;  // run reaches position 0
            m = DB[--w];
        }
        u32 top = w * 32 + (31 - __clz(m));
        return q - top;
    }
};

// Every queue entry x gets its predecessor byte cached in dig[j] (j = read
// index), gathered once: the chain step needs "T[x-1] == c" and the
// multi-split that follows needs T[x-1] as its digit.  A suffix without a
// predecessor caches 0; consumers test Pred::has themselves.

// Fill one bucket's queue: positions qp(0..limit); generation 0 (the entries
// induced from other buckets) is [0, *d_start).  Every further entry is a
// chain: seed x_k followed by x_k-1, x_k-2, ... while the predecessor byte is
// c, so seed k contributes r_k entries, its run of c's before it.  The queue
// is breadth-first: all depth-1 entries in seed order, then depth 2, ...;
// the entry (k, depth t) sits at a + sum_{t'<t} #{r >= t'} + #{k' < k :
// r_k' >= t}, i.e. the entries emitted seed by seed, stably sorted by depth.
// One pass measures the runs, one scan gives emission offsets, and one stable
// multi-split by depth places them (two passes if a run exceeds 256).
// Caches every entry's predecessor byte in dig[0..limit).  d_tmp: 2 u32.
inline void chain_queue8(Dev &d, Pred pr, u32 *SA, u8 *dig, QPos qp, u32 c,
                         const u32 *d_start, u64 limit, u32 *d_tmp) {
    const u64 a = d.read_u32(d_start);
    u32 *err = d.err;
    if (a > limit) fail(PSK_EDEVICE, "queue seed count exceeds its bucket part");
    Arena::Mark mk = d.a->mark();
    u32 *run = d.a->get<u32>(a + 1);
    u8 *stop = d.a->get<u8>(a + 1);
    u32 *maxr = d_tmp;
    d.zero(maxr, sizeof(u32));
    const u64 nchain = limit - a;
    {
        int prec = d.pbegin("chain8_runs", 10.0 * a);
        const u64 aa = a;
        // one thread per seed; the grid is a multiple of the wave so every
        // lane takes part in the wave-wide maximum (one atomic per wave)
        u64 blocks = (aa + BLOCK - 1) / BLOCK;
        k_for_each_waves<<<(u32)std::max<u64>(blocks, 1), BLOCK, 0, d.s>>>(
            [=] __device__(u64 k, bool live) {
                u32 r = 0;
                if (live) {
                    u32 x = SA[qp(k)];
                    u32 cp = pr.byte(x);
                    dig[k] = (u8)cp;
                    u32 sb = cp;
                    if (pr.has(x) && cp == c) {
                        r = pr.run_before(x);  // T[x] == c: the run of c's before x
                        sb = pr.byte(x - r);   // predecessor byte of the chain's last entry
                    }
                    run[k] = r;
                    stop[k] = (u8)sb;
                }
                // same-address atomics serialise in L2: only raise the maximum
                // when this wave beats the value already stored
                u32 wm = wave_max(r);
                if (lane_id() == 0 && wm > __atomic_load_n(maxr, __ATOMIC_RELAXED))
                    atomicMax(maxr, wm);
            },
            aa);
        PSK_LAUNCH_CHECK();
        d.pend(prec);
    }
    if (nchain == 0) { d.a->reset(mk); return; }
    // emission offsets
    u32 *eo = d.a->get<u32>(a + 1);
    u32 *tot = d_tmp + 1;
    scan_exclusive(
        d, a, [=] __device__(u64 k) { return run[k]; },
        [=] __device__(u64 k, u32 ex, u32) { eo[k] = ex; }, tot);
    u32 hv[2];
    d.read(hv, d_tmp, sizeof(hv));
    const u32 R = hv[0];
    if ((u64)hv[1] != nchain) fail(PSK_EDEVICE, "chain entries do not fill their bucket part");
    // emit (depth-1, entry, its predecessor byte) seed by seed
    u32 *ek = d.a->get<u32>(nchain), *ev = d.a->get<u32>(nchain);
    u8 *eb = d.a->get<u8>(nchain);
    for_each(d, a, [=] __device__(u64 k) {
        u32 r = run[k];
        if (!r) return;
        u32 x = SA[qp(k)], o = eo[k];
        for (u32 t = 1; t <= r; t++) {
            ek[o + t - 1] = t - 1;
            ev[o + t - 1] = x - t;
            eb[o + t - 1] = t < r ? (u8)c : stop[k];
        }
    });
    const u64 lim = limit, base = a;
    if (R <= 256) {
        multisplit(
            d, nchain,
            [=] __device__(u64 j, u32 *key) {
                *key = ek[j];
                return true;
            },
            [=] __device__(u64 j, u32, u32 gpos, u32) {
                u64 q = base + gpos;
                if (q < lim) {
                    SA[qp(q)] = ev[j];
                    dig[q] = eb[j];
                } else {
                    atomicOr(err, 2u);
                }
            },
            nullptr);
    } else {
        // long runs: stable LSD sort of (depth, emission index)
        u64 *k0 = d.a->get<u64>(nchain), *k1 = d.a->get<u64>(nchain);
        u32 *v0 = d.a->get<u32>(nchain), *v1 = d.a->get<u32>(nchain);
        for_each(d, nchain, [=] __device__(u64 j) {
            k0[j] = ek[j];
            v0[j] = (u32)j;
        });
        bool sw = radix_sort_kv(d, nchain, k0, v0, k1, v1, bit_width(R - 1));
        const u32 *V = sw ? v1 : v0;
        for_each(d, nchain, [=] __device__(u64 i) {
            u32 j = V[i];
            SA[qp(base + i)] = ev[j];
            dig[base + i] = eb[j];
        });
    }
    d.a->reset(mk);
}

// Per-bucket device tables: head, tail, fill (induce counters).
struct Induce8Tables {
    u32 *head, *tail, *fill, *counts, *tmp;
};

// Full induce (L-scan then S-scan) over SA for a byte text, from the LMS
// entries already placed at the bucket tails.  dig: scratch of >= n bytes.
// PB: n bytes indexed by SA slot.  On entry, PB holds T[x-1] at the slot of
// every placed LMS entry x (written by the placement, which reads T there
// anyway); the L-scan reads those instead of gathering T, and stores at every
// L slot the digit its entry will have in the S-scan (the predecessor byte if
// it is smaller than the bucket's, else 255), so the S-scan reads its L parts
// sequentially.  L slots are final after the L-scan, so the S-scan finds them.
// Batches (SB != nullptr): K texts concatenated, ends[0..K) their last
// positions; the L-scan starts from every text's sentinel, in text order.
inline void induce8(Dev &d, const u8 *T, u64 n, u32 *SA, const Buckets8 &B, Induce8Tables tb,
                    u8 *dig, u8 *PB, const u32 *DB, const u32 *SB = nullptr,
                    const u32 *ends = nullptr, u64 K = 1) {
    const Pred pr{T, SB, DB};
    u32 *head = tb.head, *tail = tb.tail, *fill = tb.fill, *counts = tb.counts;
    u32 *err = d.err;
    const u64 nn = n;
    // ---------------- L-scan ----------------
    d.zero(fill, 256 * sizeof(u32));
    if (!ends) {
        for_each(d, 1, [=] __device__(u64) {
            u32 c = T[nn - 1];
            SA[head[c]] = (u32)(nn - 1);  // the sentinel induces n-1 first
            fill[c] = 1;
        });
    } else {
        // text i's sentinel induces its last position; sentinels read in text order
        multisplit(
            d, K,
            [=] __device__(u64 i, u32 *key) {
                *key = T[ends[i]];
                return true;
            },
            [=] __device__(u64 i, u32 key, u32 gpos, u32 dstart) {
                SA[head[key] + (gpos - dstart)] = ends[i];
            },
            counts);
        for_each(d, 256, [=] __device__(u64 k) { fill[k] = counts[k]; });
    }
    for (int c = 0; c < 256; c++) {
        if (!B.cnt[c]) continue;
        const u64 hc = B.head[c], tc = B.head[c + 1], nl = B.nL[c], nlms = B.nLMS[c];
        const u32 cc = (u32)c;
        if (nl) chain_queue8(d, pr, SA, dig, QPos{hc, 1}, cc, fill + c, nl, tb.tmp);
        const u64 R = nl + nlms;
        if (c == 255) {
            // nothing can be induced above 255; the S-scan digits of the L part
            // (PB) still come from the cached predecessor bytes
            if (nl) {
                for_each(d, nl, [=] __device__(u64 j) {
                    u32 x = SA[hc + j];
                    u32 cp = dig[j];
                    PB[hc + j] = (u8)((pr.has(x) && cp < cc) ? cp : 255u);
                });
            }
            continue;
        }
        if (R == 0) continue;
        auto pos = [=] __device__(u64 j) -> u64 { return j < nl ? hc + j : tc - nlms + (j - nl); };
        int ps = multisplit2(
            d, R,
            [=] __device__(u64 j, u32 *key) {
                u32 k;
                if (j < nl) {
                    // the queue part has its predecessor bytes cached in dig
                    u32 x = SA[hc + j];
                    u32 cp = dig[j];
                    bool h = pr.has(x);
                    k = (h && cp > cc) ? cp : 0u;
                    PB[hc + j] = (u8)((h && cp < cc) ? cp : 255u);  // its S-scan digit
                } else {
                    // an LMS entry: its predecessor (L-type, so > c) is in PB
                    u32 cp = PB[pos(j)];
                    k = cp > cc ? cp : 0u;
                }
                dig[j] = (u8)k;  // 0 = not induced (valid digits are > c >= 0)
                *key = k;
                return k != 0;
            },
            [=] __device__(u64 j, u32 *key) {
                u32 k = dig[j];
                *key = k;
                return k != 0;
            },
            [=] __device__(u64 j, u32 key, u32 gpos, u32 dstart) {
                u32 x = SA[pos(j)];
                u64 dst = (u64)head[key] + fill[key] + (gpos - dstart);
                if (dst < nn) SA[dst] = x - 1;
                else atomicOr(err, 4u);
            },
            counts, "induce8", 6.0 * R, 5.0 * R);
        u64 *slot = d.pslot(ps);
        for_each(d, 256, [=] __device__(u64 k) {
            fill[k] += counts[k];
            if (slot && counts[k]) atomicAdd((unsigned long long *)slot, 4ull * counts[k]);
        });
    }
    // ---------------- S-scan ----------------
    d.zero(fill, 256 * sizeof(u32));
    for (int c = 255; c >= 0; c--) {
        if (!B.cnt[c]) continue;
        const u64 hc = B.head[c], tc = B.head[c + 1], nl = B.nL[c], ns = B.cnt[c] - B.nL[c];
        const u32 cc = (u32)c;
        if (ns) chain_queue8(d, pr, SA, dig, QPos{tc - 1, -1}, cc, fill + c, ns, tb.tmp);
        const u64 R = ns + nl;
        if (c == 0 || R == 0) continue;  // nothing can be induced below 0
        auto pos = [=] __device__(u64 j) -> u64 { return j < ns ? tc - 1 - j : hc + nl - 1 - (j - ns); };
        int ps = multisplit2(
            d, R,
            [=] __device__(u64 j, u32 *key) {
                u32 k;
                if (j < ns) {
                    u32 x = SA[pos(j)];
                    u32 cp = dig[j];
                    k = (pr.has(x) && cp < cc) ? cp : 255u;
                } else {
                    k = PB[pos(j)];  // stored by the L-scan
                }
                dig[j] = (u8)k;  // 255 = not induced (valid digits are < c <= 255)
                *key = k;
                return k != 255;
            },
            [=] __device__(u64 j, u32 *key) {
                u32 k = dig[j];
                *key = k;
                return k != 255;
            },
            [=] __device__(u64 j, u32 key, u32 gpos, u32 dstart) {
                u32 x = SA[pos(j)];
                i64 dst = (i64)tail[key] - 1 - (i64)(fill[key] + (gpos - dstart));
                if (dst >= 0 && dst < (i64)nn) SA[dst] = x - 1;
                else atomicOr(err, 8u);
            },
            counts, "induce8", 6.0 * R, 5.0 * R);
        u64 *slot = d.pslot(ps);
        for_each(d, 256, [=] __device__(u64 k) {
            fill[k] += counts[k];
            if (slot && counts[k]) atomicAdd((unsigned long long *)slot, 4ull * counts[k]);
        });
    }
}

}  // namespace psk
