// psk_level8.hpp -- level 0 (byte text) driver: psacak8.rs:19-53 on the GPU.
#pragma once

#include "psk_induce8.hpp"
#include "psk_level32.hpp"

namespace psk {

// A batch of texts concatenated in one array: their last positions (ends,
// in text order) and bit vectors of last and first positions.
struct BatchInfo {
    const u32 *ends;
    u64 K;
    const u32 *endbits, *startbits;
};

// An LMS suffix as the final induce places it: its position and the window
// of the bytes before it (one 16-byte record, fetched by one load).
struct __align__(16) LmsRec {
    u64 x, w;
};

// Records of the LMS positions in text order (lms_text ascending, so the
// text reads of the windows are sequential).
template <bool ALIGNED>
__global__ void __launch_bounds__(BLOCK) k_lms_records(const u32 *lms_text, u32 m1, Pred pr,
                                                       LmsRec *R) {
    for (u64 p = (u64)blockIdx.x * BLOCK + threadIdx.x; p < m1; p += (u64)gridDim.x * BLOCK) {
        const u32 x = lms_text[p];
        R[p] = LmsRec{x, pr.window<ALIGNED>(x)};
    }
}

// The bucket of sorted LMS entry j: the largest c with st[c] <= j (st[256] =
// m1 > j), by binary search in LDS.
__device__ __forceinline__ u32 lms_bucket(const u32 *st, u32 j) {
    u32 lo = 0, hi = 256;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) >> 1;
        if (st[mid] <= j) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Sorted LMS suffixes to their bucket tails (put_lmssuffixes psacak8.rs:76-87)
// from records: entry j of SA1 (an index into the text-order LMS list) starts
// with the byte c whose LMS range [lstart[c], lstart[c+1]) holds j and goes to
// j + lmsoff[c]; one record gather gives its position and window.
__global__ void __launch_bounds__(BLOCK) k_place_lms_rec(const u32 *SA1, u32 m1, const u32 *lstart,
                                                         const u32 *lmsoff, const LmsRec *R,
                                                         u32 *SA, u64 *PW, u32 *err) {
    __shared__ u32 st[257], off[256];
    for (int i = threadIdx.x; i < 257; i += BLOCK) st[i] = lstart[i];
    off[threadIdx.x] = lmsoff[threadIdx.x];
    __syncthreads();
    for (u64 j = (u64)blockIdx.x * BLOCK + threadIdx.x; j < m1; j += (u64)gridDim.x * BLOCK) {
        const u32 q = SA1[j];
        if (q >= m1) {
            atomicOr(err, 512u);
            continue;
        }
        const LmsRec r = R[q];
        const u32 dst = (u32)j + off[lms_bucket(st, (u32)j)];
        SA[dst] = (u32)r.x;
        PW[dst] = r.w;
    }
}

// The same placement without records (lean contexts): the entry goes to
// j + lmsoff[c] as lms_text[SA1[j]], and its window (PW) is gathered from the
// text: two dependent random reads instead of one.
template <bool ALIGNED>
__global__ void __launch_bounds__(BLOCK) k_place_lms(const u32 *SA1, u32 m1, const u32 *lstart,
                                                     const u32 *lmsoff, const u32 *lms_text,
                                                     Pred pr, u32 *SA, u64 *PW, u32 *err) {
    __shared__ u32 st[257], off[256];
    for (int i = threadIdx.x; i < 257; i += BLOCK) st[i] = lstart[i];
    off[threadIdx.x] = lmsoff[threadIdx.x];
    __syncthreads();
    for (u64 j = (u64)blockIdx.x * BLOCK + threadIdx.x; j < m1; j += (u64)gridDim.x * BLOCK) {
        const u32 q = SA1[j];
        if (q >= m1) {
            atomicOr(err, 512u);
            continue;
        }
        const u32 p = lms_text[q];
        const u32 dst = (u32)j + off[lms_bucket(st, (u32)j)];
        SA[dst] = p;
        PW[dst] = pr.window<ALIGNED>(p);
    }
}

// Level-0 tables from the bucket statistics (count, L count, LMS count per
// byte), one workgroup: bucket heads and tails, zeroed induce counters, the
// LMS placement bases and offsets, the start of each byte's LMS range in the
// sorted LMS list (lstart, 257 entries) and the rank of each byte among those
// present (crank).  Keeps the tables off the host.
__global__ void __launch_bounds__(BLOCK) k_level0_tables(const u32 *stat, u32 *tables, u8 *crank,
                                                         u32 *lstart) {
    __shared__ u32 red[NWAVES + 1];
    const u32 c = threadIdx.x;
    const u32 cnt = stat[c], nlms = stat[512 + c];
    u32 tot;
    const u32 head = block_exclusive_sum(cnt, red, &tot);
    const u32 ls = block_exclusive_sum(nlms, red, &tot);
    const u32 rk = block_exclusive_sum(cnt ? 1u : 0u, red, &tot);
    tables[c] = head;
    tables[256 + c] = head + cnt;
    tables[512 + c] = 0;   // fill
    tables[768 + c] = 0;   // counts
    tables[1024 + c] = 0;  // tmp
    tables[1280 + c] = head + cnt - nlms;       // LMS placement base
    tables[1536 + c] = head + cnt - nlms - ls;  // sorted LMS j -> slot j + lmsoff
    lstart[c] = ls;
    if (c == 255) lstart[256] = ls + nlms;
    crank[c] = (u8)rk;
}

// Device bytes of the level-0 chain-sort induce beyond the bit vectors: LMS
// list and SA1, the bucket-pointer text, ranks, item lists, labels and
// scratch (20 n), the short round-1 keys of the L sort, and the larger of the
// LMS records and a chain sort over the larger type class (~43 bytes per
// item: working copies, group ids and labels, keys, radix and group-sort
// tables).
inline u64 chain0_bytes(u64 n, u64 m1, u64 nL) {
    const u64 A = std::max<u64>(nL, n - nL);
    return 3 * (n / 8 + 8) + 16 * (m1 + 1) + 20 * (n + 1) + 8 * m1 +
           std::max<u64>(28 * m1, 43 * A) + (64ull << 20);
}

// SA of the byte text T[0..n) into SA (device buffers).  With bi, T holds a
// batch of texts, each with its own sentinel, and SA receives the SA of the
// concatenation in which every text's suffixes are in their own order.
inline void psacak8_run(Runtime &rt, const u8 *T, u64 n, u32 *SA, const BatchInfo *bi = nullptr) {
    Dev &d = rt.d;
    rt.levels.clear();
    rt.max_rounds = 0;
    rt.doubling_gave_up = false;
    rt.captured_ok = false;
    if (n == 0) return;
    const u32 *EB = bi ? bi->endbits : nullptr;
    const u32 *SB = bi ? bi->startbits : nullptr;
    const u32 *ENDS = bi ? bi->ends : nullptr;
    const u64 K = bi ? bi->K : 1;
    if (n <= 3 && !bi) {  // psacak8.rs:23-26
        k_naive_small<u8><<<1, 1, 0, d.s>>>(T, (u32)n, SA);
        PSK_LAUNCH_CHECK();
        return;
    }
    ArenaScope mk(*d.a);
    rt.depth = 0;
    d.phase(nullptr);
    const u64 nwords = (n + 31) / 32;
    u32 *sbits = d.a->get<u32>(nwords + 1);
    u32 *lmsbits = d.a->get<u32>(nwords + 1);
    u32 *dbits = d.a->get<u32>(nwords + 1);
    compute_types(d, T, n, sbits, EB, dbits);

    // bucket statistics (Buckets::new psacak8.rs:355-374)
    u32 *stat = d.a->get<u32>(3 * 256 + 1);
    d.zero(stat, (3 * 256 + 1) * sizeof(u32));
    {
        // about 2048 workgroups, each of at most 2^15 words (2^20 positions)
        u64 wpb = (nwords + 2047) / 2048;
        wpb = std::min<u64>(std::max<u64>((wpb + BLOCK - 1) / BLOCK * BLOCK, BLOCK), 1u << 15);
        k_lms_hist8<<<ceil_div(nwords, wpb), BLOCK, 0, d.s>>>(T, n, sbits, lmsbits, nwords, wpb,
                                                             stat, stat + 256, stat + 512,
                                                             stat + 768, SB);
    }
    PSK_LAUNCH_CHECK();
    Buckets8 B;
    u32 host_stat[3 * 256 + 1];
    d.read(host_stat, stat, sizeof(host_stat));
    memcpy(B.cnt, host_stat, sizeof(B.cnt));
    memcpy(B.nL, host_stat + 256, sizeof(B.nL));
    memcpy(B.nLMS, host_stat + 512, sizeof(B.nLMS));
    const u32 m1 = host_stat[768];
    B.head[0] = 0;
    u64 nL_total = 0;
    for (int c = 0; c < 256; c++) {
        B.head[c + 1] = B.head[c] + B.cnt[c];
        nL_total += B.nL[c];
    }

    // device tables: head, tail, fill, counts, tmp, lmsbase, lmsoff; the LMS
    // ranges and the byte ranks
    u32 *tables = d.a->get<u32>(7 * 256 + 8);
    Induce8Tables tb{tables, tables + 256, tables + 512, tables + 768, tables + 1024};
    u32 *lmsbase = tables + 1280, *lmsoff = tables + 1536;
    u32 *lstart = d.a->get<u32>(257);
    u8 *crank = d.a->get<u8>(256);
    k_level0_tables<<<1, BLOCK, 0, d.s>>>(stat, tables, crank, lstart);
    PSK_LAUNCH_CHECK();

    u32 *lms_text = d.a->get<u32>((u64)m1 + 1);
    bits_to_positions(d, lmsbits, nwords, lms_text);
    // Many distinct bytes: the final induce runs as chain sorts
    // (final_induce32), whose cost does not grow with the alphabet, instead
    // of two multi-split steps per bucket, when its workspace stays within
    // the published bound.  The bucket steps' digit cache and byte windows
    // are then needed only by an induced naming fallback.
    u32 sigma = 0;
    for (int c = 0; c < 256; c++) sigma += B.cnt[c] != 0;
    const u32 csig = n < Runtime::CHAIN_SMALL_N ? rt.chain_sigma_small : rt.chain_sigma;
    const bool chain0 = !bi && csig && sigma >= csig && n < (1ull << 31) &&
                        chain0_bytes(n, m1, nL_total) <=
                            (rt.lms_records ? workspace_bytes8(n) : lean_bytes8(n));
    u8 *dig = chain0 ? nullptr : d.a->get<u8>(n);
    u64 *PW = chain0 ? nullptr : d.a->get<u64>(n);
    const Pred pr{T, SB, dbits};

    d.phase("level 0 types, buckets, LMS list");
    LevelStat st{n, m1, nL_total, 0, 0, 1};
    rt.levels.resize(1);
    record_first_lms(d, lms_text, m1, 0);
    rt.levels[0] = st;

    // ---- LMS-substring names and the reduced problem (psacak8.rs:33-48) ------
    u32 *SA1 = d.a->get<u32>((u64)m1 + 1);
    // placement records (filled after the recursion), held from here:
    // allocated after the recursion they would need a chunk of their own that
    // no other buffer uses, call after call
    LmsRec *R = rt.lms_records && !chain0 ? d.a->get<LmsRec>(m1) : nullptr;
    bool named = false;
    {
        // by sorting the substrings, bytes ranked among those present; in a
        // batch each text's last substring runs into its own sentinel
        LmsEnds ends;
        if (EB) {
            u32 *EBR = d.a->get<u32>(nwords + 1);
            scan_exclusive(
                d, nwords, [=] __device__(u64 w) { return (u32)__popc(EB[w]); },
                [=] __device__(u64 w, u32 ex, u32) { EBR[w] = ex; }, nullptr);
            ends.EB = EB;
            ends.EBR = EBR;
        }
        named = solve_reduced_direct(rt, ByteRanks{T, crank}, sigma, n, lms_text, m1, SA1, 0, st,
                                     ends);
    }
    if (!named) {
        // induced stage 1 (long repeated LMS substrings, or batch texts whose
        // index does not fit a key): LMS characters to bucket tails
        // (psacak8.rs:57-72), in text order within a bucket, then induce the
        // LMS substrings
        ArenaScope ms(*d.a);
        if (chain0) {
            dig = d.a->get<u8>(n);
            PW = d.a->get<u64>(n);
        }
        multisplit(
            d, m1,
            [=] __device__(u64 i, u32 *key) {
                *key = T[lms_text[i]];
                return true;
            },
            [=] __device__(u64 i, u32 key, u32 gpos, u32 dstart) {
                u32 p = lms_text[i];
                u32 dst = lmsbase[key] + (gpos - dstart);
                SA[dst] = p;
                PW[dst] = pr.window<false>(p);
            },
            nullptr);
        induce8(d, T, n, SA, B, tb, dig, PW, dbits, SB, ENDS, K, !rt.lms_records);

        // sorted LMS substrings = LMS entries in SA order (compact_left
        // psacak8.rs:35): LMS flags of the SA entries packed once, then compacted
        u32 *lms_sorted = d.a->get<u32>((u64)m1 + 1);
        {
            ArenaScope m2(*d.a);
            u32 *fl = d.a->get<u32>(nwords + 1);
            for_each_waves(d, n, [=] __device__(u64 i, bool live) {
                u64 b = ballot(live && getbit(lmsbits, SA[i]));
                const int l = lane_id();
                if (l == 0 && i < n) fl[i >> 5] = (u32)b;
                if (l == 32 && i < n) fl[i >> 5] = (u32)(b >> 32);
            });
            scan_exclusive(
                d, nwords, [=] __device__(u64 w) { return (u32)__popc(fl[w]); },
                [=] __device__(u64 w, u32 off, u32) {
                    u32 x = fl[w];
                    while (x) {
                        int b = __ffs(x) - 1;
                        x &= x - 1;
                        lms_sorted[off++] = SA[w * 32 + b];
                    }
                },
                nullptr);
            m2.reset();
        }
        u32 *NP = d.a->get<u32>(n / 2 + 1);
        solve_reduced(rt, T, lms_sorted, lms_text, lmsbits, nwords, NP, m1, SA1, 0, st, EB);
        ms.reset();
    }

    if (chain0) {
        // the text in bucket-pointer form (an L byte -> its bucket's head, an
        // S byte -> its tail - 1: naming.rs:263-274), the bytes' dense ranks
        // with the tail of each rank's bucket, and the L and S positions
        d.phase("level 0 recursion below");
        ArenaScope mc(*d.a);
        u32 *s0 = d.a->get<u32>(n), *rank0 = d.a->get<u32>(n);
        u32 *items = d.a->get<u32>(n);
        u32 *RK = d.a->get<u32>(n + 1), *cnt = d.a->get<u32>(n + 1);
        u32 *rtails = d.a->get<u32>(256);
        const u32 *hd = tables, *tl = tables + 256, *bc = stat;
        const u8 *cr = crank;
        for_each(d, 256, [=] __device__(u64 c) {
            if (bc[c]) rtails[cr[c]] = tl[c];
        });
        const u32 *sb = sbits;
        for_each(d, n, [=] __device__(u64 x) {
            const u32 c = T[x];
            s0[x] = getbit(sb, x) ? tl[c] - 1 : hd[c];
            rank0[x] = cr[c];
        });
        const u64 nL = nL_total;
        bit_list(d, n, [=] __device__(u64 w) { return ~sb[w]; }, items, nullptr);
        positions_of(d, n, [=] __device__(u64 w) { return sb[w]; }, items + nL);
        final_induce32(rt, s0, n, sbits, rank0, rtails, sigma, lms_text, m1, SA1, items,
                       (u32)nL, items + nL, (u32)(n - nL), RK, cnt, SA);
        mc.reset();
        d.phase("level 0 final induce");
        mk.reset();
        return;
    }

    d.phase("level 0 records after recursion");
    // ---- sorted LMS suffixes to bucket tails (put_lmssuffixes psacak8.rs:76-87)
    // SA1 is sorted, so its entries come grouped by first byte: the bucket of
    // entry j follows from the LMS counts, without reading T[p]
    if (m1) {
        // SA1 holds reduced indices (the recursion leaves the mapping to
        // this level)
        u64 blocks = std::min<u64>(ceil_div(m1, BLOCK), 65536);
        const bool al = (((uintptr_t)T) & 7u) == 0;
        if (R) {
            if (al) k_lms_records<true><<<(u32)blocks, BLOCK, 0, d.s>>>(lms_text, m1, pr, R);
            else k_lms_records<false><<<(u32)blocks, BLOCK, 0, d.s>>>(lms_text, m1, pr, R);
            PSK_LAUNCH_CHECK();
            k_place_lms_rec<<<(u32)blocks, BLOCK, 0, d.s>>>(SA1, m1, lstart, lmsoff, R, SA, PW,
                                                             d.err);
        } else if (al) {
            k_place_lms<true><<<(u32)blocks, BLOCK, 0, d.s>>>(SA1, m1, lstart, lmsoff, lms_text, pr,
                                                              SA, PW, d.err);
        } else {
            k_place_lms<false><<<(u32)blocks, BLOCK, 0, d.s>>>(SA1, m1, lstart, lmsoff, lms_text,
                                                               pr, SA, PW, d.err);
        }
        PSK_LAUNCH_CHECK();
    }
    d.phase("level 0 placement");
    induce8(d, T, n, SA, B, tb, dig, PW, dbits, SB, ENDS, K, !rt.lms_records);
    d.phase("level 0 final induce");
    mk.reset();
}

// SAs of K texts stored back to back in T: text i is T[off[i] .. off[i+1])
// (off: host array, off[0] = 0, off[K] = N <= 2^32-2).  SA[off[i] + r] is
// the r-th smallest suffix of text i, as a position inside text i.  All texts
// share every kernel launch: one level-0 run over the concatenation in which
// each text has its own sentinel, then a stable partition by text id.
inline void batch8_run(Runtime &rt, const u8 *T, const u64 *off, u64 K, u32 *SA) {
    Dev &d = rt.d;
    const u64 N = off[K];
    if (N == 0) return;
    ArenaScope mk(*d.a);
    const u64 nwords = (N + 31) / 32;
    u32 *EB = d.a->get<u32>(nwords + 1);
    u32 *SB = d.a->get<u32>(nwords + 1);
    d.zero(EB, (nwords + 1) * sizeof(u32));
    d.zero(SB, (nwords + 1) * sizeof(u32));
    std::vector<u32> hoff(K + 1), hends;
    hends.reserve(K);
    for (u64 i = 0; i <= K; i++) hoff[i] = (u32)off[i];
    for (u64 i = 0; i < K; i++)
        if (off[i + 1] > off[i]) hends.push_back((u32)(off[i + 1] - 1));
    u32 *doff = d.a->get<u32>(K + 1);
    u32 *dends = d.a->get<u32>(hends.size() + 1);
    d.upload(doff, hoff.data(), (K + 1) * sizeof(u32));
    d.upload(dends, hends.data(), hends.size() * sizeof(u32));
    const u64 KK = K;
    for_each(d, K, [=] __device__(u64 i) {
        u32 a = doff[i], b = doff[i + 1];
        if (b > a) {
            atomicOr(&SB[a >> 5], 1u << (a & 31));
            atomicOr(&EB[(b - 1) >> 5], 1u << ((b - 1) & 31));
        }
    });
    BatchInfo bi{dends, hends.size(), EB, SB};
    psacak8_run(rt, T, N, SA, &bi);
    // stable partition of the concatenation's SA by text id.  The text of a
    // position x is the last i with doff[i] <= x (empty texts share their
    // offset with the next text); a table of the text at every 4 KiB of
    // positions narrows the search to the texts that start in x's block.
    constexpr int BTS = 12;
    const u64 nb = (N + (1ull << BTS) - 1) >> BTS;
    u32 *bt = d.a->get<u32>(nb + 1);
    auto last_le = [=] __device__(u64 lo, u64 hi, u64 x) -> u64 {
        // largest i in [lo, hi) with doff[i] <= x, given doff[lo] <= x
        while (hi - lo > 1) {
            u64 mid = (lo + hi) / 2;
            if (doff[mid] <= x) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    for_each(d, nb, [=] __device__(u64 b) { bt[b] = (u32)last_le(0, KK, b << BTS); });
    // the values (positions inside their text) replace the SA entries in
    // place and the sort starts from SA, so an even number of passes leaves
    // the result in SA with no copy
    u64 *k0 = d.a->get<u64>(N), *k1 = d.a->get<u64>(N);
    u32 *v1 = d.a->get<u32>(N);
    for_each(d, N, [=] __device__(u64 j) {
        const u32 x = SA[j];
        const u64 b = x >> BTS;
        const u64 hi = b + 1 < nb ? (u64)bt[b + 1] + 1 : KK;
        const u64 i = last_le(bt[b], hi < KK ? hi : KK, x);
        k0[j] = i;
        SA[j] = x - doff[i];
    });
    const bool sw = radix_sort_kv(d, N, k0, SA, k1, v1, bit_width(K > 1 ? K - 1 : 0));
    if (sw) PSK_HIP(hipMemcpyAsync(SA, v1, N * sizeof(u32), hipMemcpyDeviceToDevice, d.s));
    mk.reset();
}

// psacak32 on a bucket-pointer text (device buffers).
inline void psacak32_run(Runtime &rt, const u32 *s, u64 m, u32 *SA) {
    Dev &d = rt.d;
    rt.levels.clear();
    rt.max_rounds = 0;
    rt.doubling_gave_up = false;
    rt.captured_ok = false;
    if (m == 0) return;
    ArenaScope mk(*d.a);
    u32 *sb = d.a->get<u32>((m + 31) / 32 + 1);
    compute_types(d, s, m, sb);
    rt.levels.resize(1);
    level32(rt, s, m, sb, SA, 0);
    mk.reset();
}

// Run boundaries of sorted symbol keys K[0..n) (values < n): hd[x] / tl[x]
// = first / last sorted index of symbol x.  The run starts are summed per
// workgroup and added to one of RUN_COUNTERS counters (by workgroup), so the
// count takes a few thousand atomics, not one per wave that meets a run
// start (those serialise on a shared address).
constexpr int RUN_COUNTERS = 64;
__global__ void __launch_bounds__(BLOCK) k_run_ends(const u64 *K, u64 n, u32 *hd, u32 *tl,
                                                    u32 *cs) {
    __shared__ u32 wc[NWAVES];
    u32 c = 0;
    for (u64 i = (u64)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (u64)gridDim.x * BLOCK) {
        const u64 x = K[i];
        if (i == 0 || K[i - 1] != x) {
            hd[x] = (u32)i;
            c++;
        }
        if (i + 1 == n || K[i + 1] != x) tl[x] = (u32)i;
    }
    #pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, WAVE);
    if (lane_id() == 0) wc[wave_id()] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 tot = 0;
        for (int w = 0; w < NWAVES; w++) tot += wc[w];
        if (tot) atomicAdd(cs + (blockIdx.x % RUN_COUNTERS), tot);
    }
}

// translate_sample (psacak32.rs:912-943) on the device: rank the symbols,
// L -> head of the symbol's bucket, S -> tail - 1.  Returns false (and writes
// the SA directly) if every symbol is distinct.
inline bool translate32_run(Runtime &rt, const u32 *t, u64 n, u32 *out, u32 *SA) {
    Dev &d = rt.d;
    ArenaScope mk(*d.a);
    u64 *k = d.a->get<u64>(n), *kt = d.a->get<u64>(n);
    u32 *v = d.a->get<u32>(n), *vt = d.a->get<u32>(n);
    u32 *first = d.a->get<u32>(n);
    u32 *ctr = d.a->get<u32>(2);
    u32 *mx = ctr + 1;
    d.zero(mx, sizeof(u32));
    // keys, and the largest symbol (the sort needs only its bits); a wave
    // raises the maximum only when it beats the value already stored, so the
    // same-address atomics stay few
    if ((((uintptr_t)t) & 15u) == 0) {
        // four symbols per lane with 16-byte accesses
        const u64 nq = (n + 3) / 4, nn = n;
        for_each_waves(d, nq, [=] __device__(u64 q, bool live) {
            u32 x4[4] = {0u, 0u, 0u, 0u};
            const u64 i = 4 * q;
            if (live) {
                if (i + 4 <= nn) {
                    const uint4 x = *(const uint4 *)(t + i);
                    x4[0] = x.x;
                    x4[1] = x.y;
                    x4[2] = x.z;
                    x4[3] = x.w;
                    *(ulonglong2 *)(k + i) = make_ulonglong2(x.x, x.y);
                    *(ulonglong2 *)(k + i + 2) = make_ulonglong2(x.z, x.w);
                    *(uint4 *)(v + i) = make_uint4((u32)i, (u32)i + 1, (u32)i + 2, (u32)i + 3);
                } else {
                    for (u64 j = i; j < nn; j++) {
                        x4[j - i] = t[j];
                        k[j] = t[j];
                        v[j] = (u32)j;
                    }
                }
            }
            const u32 m01 = x4[0] > x4[1] ? x4[0] : x4[1], m23 = x4[2] > x4[3] ? x4[2] : x4[3];
            const u32 wm = wave_max(m01 > m23 ? m01 : m23);
            if (lane_id() == 0 && wm > __atomic_load_n(mx, __ATOMIC_RELAXED)) atomicMax(mx, wm);
        });
    } else {
        for_each_waves(d, n, [=] __device__(u64 i, bool live) {
            const u32 x = live ? t[i] : 0u;
            if (live) {
                k[i] = x;
                v[i] = (u32)i;
            }
            const u32 wm = wave_max(x);
            if (lane_id() == 0 && wm > __atomic_load_n(mx, __ATOMIC_RELAXED)) atomicMax(mx, wm);
        });
    }
    const u32 maxv = d.read_u32(mx);
    const int bits = std::max(1, bit_width(maxv));
    bool sw = radix_sort_kv(d, n, k, v, kt, vt, bits);
    const u64 *K = sw ? kt : k;
    const u32 *V = sw ? vt : v;
    const u64 nn = n;
    if ((u64)maxv < n) {
        // symbols below n: first and last sorted index of each symbol in
        // tables indexed by the symbol (written at the run boundaries of the
        // sorted order), then every position looks its value up in text order,
        // instead of a random scatter of the whole text
        u32 *hd = first, *tl = d.a->get<u32>(n);
        u32 *cs = d.a->get<u32>(RUN_COUNTERS);
        d.zero(cs, RUN_COUNTERS * sizeof(u32));
        {
            const u32 g = (u32)std::min<u64>(ceil_div(n, BLOCK), 8192);
            k_run_ends<<<g, BLOCK, 0, d.s>>>(K, n, hd, tl, cs);
            PSK_LAUNCH_CHECK();
        }
        u32 csum[RUN_COUNTERS];
        d.read(csum, cs, sizeof(csum));
        u64 dsum = 0;
        for (int c = 0; c < RUN_COUNTERS; c++) dsum += csum[c];
        const u32 distinct = (u32)std::min<u64>(dsum, n);
        if (distinct >= n) {
            PSK_HIP(hipMemcpyAsync(SA, V, n * sizeof(u32), hipMemcpyDeviceToDevice, d.s));
            mk.reset();
            return false;
        }
        u32 *sb = d.a->get<u32>((n + 31) / 32 + 1);
        compute_types(d, t, n, sb);
        // four symbols per thread with 16-byte accesses where t and out allow
        const bool vec = ((((uintptr_t)t) | ((uintptr_t)out)) & 15u) == 0;
        const u64 nq = vec ? n / 4 : 0;
        for_each(d, nq + (n - 4 * nq), [=] __device__(u64 q) {
            if (q < nq) {
                const u64 p = 4 * q;
                const uint4 x = *(const uint4 *)(t + p);
                const u32 ty = (sb[p >> 5] >> (p & 31)) & 15u;
                *(uint4 *)(out + p) =
                    make_uint4(ty & 1u ? tl[x.x] : hd[x.x], ty & 2u ? tl[x.y] : hd[x.y],
                               ty & 4u ? tl[x.z] : hd[x.z], ty & 8u ? tl[x.w] : hd[x.w]);
            } else {
                const u64 p = 4 * nq + (q - nq);
                const u32 x = t[p];
                out[p] = getbit(sb, p) ? tl[x] : hd[x];
            }
        });
        mk.reset();
        return true;
    }
    // distinct count
    scan_exclusive(
        d, n, [=] __device__(u64 i) { return (u32)(i == 0 || K[i] != K[i - 1]); },
        [=] __device__(u64, u32, u32) {}, ctr);
    u32 distinct = d.read_u32(ctr);
    if (distinct >= n) {
        PSK_HIP(hipMemcpyAsync(SA, V, n * sizeof(u32), hipMemcpyDeviceToDevice, d.s));
        mk.reset();
        return false;
    }
    // head = first sorted index of the symbol's run; tail = last + 1
    scan_inclusive_max(
        d, n, [=] __device__(u64 i) { return (i == 0 || K[i] != K[i - 1]) ? (u32)i : 0u; },
        [=] __device__(u64 i, u32 s0, u32) { first[i] = s0; });
    u32 *sb = d.a->get<u32>((n + 31) / 32 + 1);
    compute_types(d, t, n, sb);
    // last sorted index of each run, keyed by the run's first index (runs
    // are contiguous: a plain store at each run's end)
    u32 *last = d.a->get<u32>(n);
    for_each(d, n, [=] __device__(u64 i) {
        if (i + 1 == nn || K[i + 1] != K[i]) last[first[i]] = (u32)i;
    });
    for_each(d, nn, [=] __device__(u64 i) {
        u32 p = V[i];
        out[p] = getbit(sb, p) ? last[first[i]] : first[i];
    });
    mk.reset();
    return true;
}

}  // namespace psk
