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

constexpr u32 CHAIN_SMALL = 8192;

// Predecessor of suffix x: none for x = 0 and, in a batch of concatenated
// texts, none at a text's first position (start bit vector SB).
struct Pred {
    const u8 *T;
    const u32 *SB;
    __device__ __forceinline__ bool has(u32 x) const { return x > 0 && !(SB && getbit(SB, x)); }
    __device__ __forceinline__ u32 byte(u32 x) const { return has(x) ? (u32)T[x - 1] : 0u; }
};

// Every queue entry x gets its predecessor byte cached in dig[j] (j = read
// index), gathered once: the chain step needs "T[x-1] == c" and the
// multi-split that follows needs T[x-1] as its digit.  A suffix without a
// predecessor caches 0; consumers test Pred::has themselves.

// One workgroup finishes a queue: generations until the queue is full (limit)
// or a generation is empty.  A lone chain is followed by one thread.
__global__ void k_chain_finish(Pred pr, u32 *SA, u8 *dig, QPos qp, u32 c, u64 gs, u64 ge,
                               u64 limit, u32 *err) {
    __shared__ u32 lds[NWAVES + 1];
    while (gs < ge) {
        u64 cnt = ge - gs;
        if (cnt == 1) {
            // a lone chain: follow it serially
            if (threadIdx.x == 0) {
                u64 j = gs;
                u32 x = SA[qp(j)];
                for (;;) {
                    u32 cp = pr.byte(x);
                    dig[j] = (u8)cp;
                    if (!(pr.has(x) && cp == c)) break;
                    if (j + 1 >= limit) { atomicOr(err, 1u); break; }
                    SA[qp(j + 1)] = x - 1;
                    x--;
                    j++;
                }
            }
            __syncthreads();
            return;
        }
        u64 newend = ge;
        for (u64 j0 = gs; j0 < ge; j0 += BLOCK) {
            u64 j = j0 + threadIdx.x;
            u32 x = 0;
            bool v = false;
            if (j < ge) {
                x = SA[qp(j)];
                u32 cp = pr.byte(x);
                dig[j] = (u8)cp;
                v = pr.has(x) && cp == c;
            }
            u32 tot;
            u32 ex = block_exclusive_sum((u32)v, lds, &tot);
            if (v) {
                if (newend + ex < limit) SA[qp(newend + ex)] = x - 1;
                else atomicOr(err, 1u);
            }
            newend += tot;
        }
        gs = ge;
        ge = newend;
        __syncthreads();
    }
}

// Fill one bucket's queue: positions qp(0..limit); generation 0 is
// [0, *d_start).  Caches every entry's predecessor byte in dig[0..limit).
// d_tmp: one device u32.
inline void chain_queue8(Dev &d, Pred pr, u32 *SA, u8 *dig, QPos qp, u32 c,
                         const u32 *d_start, u64 limit, u32 *d_tmp) {
    u64 gs = 0, ge = d.read_u32(d_start);
    u32 *err = d.err;
    while (ge > gs) {
        u64 cnt = ge - gs;
        if (cnt <= CHAIN_SMALL) {
            k_chain_finish<<<1, BLOCK, 0, d.s>>>(pr, SA, dig, qp, c, gs, ge, limit, err);
            PSK_LAUNCH_CHECK();
            return;
        }
        const u64 g0 = gs, e0 = ge, lim = limit;
        // gather the predecessor bytes of this generation (the only gather)
        int prec = d.pbegin("chain8_gather", 9.0 * cnt);
        for_each(d, cnt, [=] __device__(u64 i) { dig[g0 + i] = (u8)pr.byte(SA[qp(g0 + i)]); });
        d.pend(prec);
        if (ge >= limit) return;  // queue full: this was the last generation
        scan_exclusive(
            d, cnt,
            [=] __device__(u64 i) { return (u32)(dig[g0 + i] == c && pr.has(SA[qp(g0 + i)])); },
            [=] __device__(u64 i, u32 off, u32 v) {
                if (v) {
                    if (e0 + off < lim) SA[qp(e0 + off)] = SA[qp(g0 + i)] - 1;
                    else atomicOr(err, 2u);
                }
            },
            d_tmp);
        u32 produced = d.read_u32(d_tmp);
        gs = ge;
        ge += produced;
    }
}

// Per-bucket device tables: head, tail, fill (induce counters).
struct Induce8Tables {
    u32 *head, *tail, *fill, *counts, *tmp;
};

// Full induce (L-scan then S-scan) over SA for a byte text, from the LMS
// entries already placed at the bucket tails.  dig: scratch of >= n bytes.
// Batches (SB != nullptr): K texts concatenated, ends[0..K) their last
// positions; the L-scan starts from every text's sentinel, in text order.
inline void induce8(Dev &d, const u8 *T, u64 n, u32 *SA, const Buckets8 &B, Induce8Tables tb,
                    u8 *dig, const u32 *SB = nullptr, const u32 *ends = nullptr, u64 K = 1) {
    const Pred pr{T, SB};
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
        if (c == 255 || R == 0) continue;  // nothing can be induced above 255
        auto pos = [=] __device__(u64 j) -> u64 { return j < nl ? hc + j : tc - nlms + (j - nl); };
        int ps = multisplit2(
            d, R,
            [=] __device__(u64 j, u32 *key) {
                u32 x = SA[pos(j)];
                // the queue part has its predecessor bytes cached
                u32 cp = j < nl ? (u32)dig[j] : pr.byte(x);
                u32 k = (pr.has(x) && cp > cc) ? cp : 0u;
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
                u32 x = SA[pos(j)];
                u32 cp = j < ns ? (u32)dig[j] : pr.byte(x);
                u32 k = (pr.has(x) && cp < cc) ? cp : 255u;
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
