// psk_prim.hpp -- device-wide primitives used by every level of the
// suffix-array pipeline: tiled 3-phase scans (sum / max), stream compaction,
// a stable 8-bit multi-split pass (one LSD radix pass with caller-defined
// digit and destination), and an LSD radix sort of (u64 key, u32 value).
//
// All kernels use 256-thread workgroups (4 wave64s).  Tiles are processed in
// "striped" rounds of 256 consecutive items so global loads coalesce and the
// in-tile order is the item order (the passes are stable).
#pragma once

#include <chrono>

#include "psk_common.hpp"

namespace psk {

// ---------------------------------------------------------------------------
// host-side device handle: stream + arena + pinned readback buffer
// ---------------------------------------------------------------------------
struct Dev {
    hipStream_t s = nullptr;
    Arena *a = nullptr;
    u32 *pinned = nullptr;  // >= 4096 u32 of pinned host memory
    u32 *err = nullptr;     // device flag: an index left its range (internal error)
    Prof *prof = nullptr;

    int pbegin(const char *name, double bytes) { return prof ? prof->begin(s, name, bytes) : -1; }
    void pend(int r) { if (prof) prof->end(r); }
    u64 *pslot(int r) { return prof ? prof->dyn_slot(r) : nullptr; }
    void padd(int r, double bytes) { if (prof) prof->add(r, bytes); }

    u64 syncs = 0;  // host synchronisations with the stream (readbacks, staged writes)
    bool trace = false;
    // with trace: every host synchronisation prints the source line asking for it
    void sync_site(const char *fn, int line) {
        syncs++;
        if (trace) fprintf(stderr, "[psk] sync %s:%d\n", fn, line);
    }
    u32 read_u32(const u32 *d, const char *fn = __builtin_FUNCTION(), int line = __builtin_LINE()) {
        PSK_HIP(hipMemcpyAsync(pinned, d, sizeof(u32), hipMemcpyDeviceToHost, s));
        PSK_HIP(hipStreamSynchronize(s));
        sync_site(fn, line);
        return pinned[0];
    }
    void read(void *host, const void *d, u64 bytes, const char *fn = __builtin_FUNCTION(),
              int line = __builtin_LINE()) {
        PSK_HIP(hipMemcpyAsync(pinned, d, bytes, hipMemcpyDeviceToHost, s));
        PSK_HIP(hipStreamSynchronize(s));
        sync_site(fn, line);
        memcpy(host, pinned, bytes);
    }
    void write(void *d, const void *host, u64 bytes) {
        // staged through pinned memory; the caller may reuse `host` at once
        memcpy(pinned + 2048, host, bytes);
        PSK_HIP(hipMemcpyAsync(d, pinned + 2048, bytes, hipMemcpyHostToDevice, s));
        PSK_HIP(hipStreamSynchronize(s));
        syncs++;
    }
    // arbitrary-size host -> device copy (pageable source), synchronous
    void upload(void *d, const void *host, u64 bytes) {
        if (bytes) PSK_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, s));
        PSK_HIP(hipStreamSynchronize(s));
        syncs++;
    }
    void zero(void *d, u64 bytes) {
        if (bytes) PSK_HIP(hipMemsetAsync(d, 0, bytes, s));
    }
    void fill_u32(u32 *d, u32 v, u64 count);
    // PSACAK_AMD_TRACE=1 in the environment (trace above): refinement rounds
    // print their sizes to stderr (extra kernels and host syncs; diagnostics only)
    // with trace: synchronise and print the time since the previous mark
    double tmark = -1;
    void phase(const char *name) {
        if (!trace) return;
        (void)hipStreamSynchronize(s);
        const double t = (double)std::chrono::duration_cast<std::chrono::microseconds>(
                             std::chrono::steady_clock::now().time_since_epoch()).count() / 1000.0;
        if (name && tmark >= 0) fprintf(stderr, "[psk] phase %s: %.3f ms\n", name, t - tmark);
        tmark = t;
    }
};

template <typename F>
__global__ void k_for_each(F f, u64 n) {
    u64 i = (u64)blockIdx.x * BLOCK + threadIdx.x;
    u64 stride = (u64)gridDim.x * BLOCK;
    for (; i < n; i += stride) f(i);
}

// Every thread of the grid calls f(i, i < n) for its items i (wave-uniform
// trip count, so every lane reaches f's wave-wide operations): one item per
// thread up to the grid's size, a grid-stride loop beyond it.
template <typename F>
__global__ void k_for_each_waves(F f, u64 n) {
    const u64 stride = (u64)gridDim.x * BLOCK;
    for (u64 base = (u64)blockIdx.x * BLOCK; base < n; base += stride) f(base + threadIdx.x, base + threadIdx.x < n);
}

// Launch f(i) for i in [0, n) with a grid-stride loop.
template <typename F>
inline void for_each(Dev &d, u64 n, F f) {
    if (n == 0) return;
    u64 blocks = (n + BLOCK - 1) / BLOCK;
    if (blocks > 65536u * 4) blocks = 65536u * 4;
    k_for_each<<<(u32)blocks, BLOCK, 0, d.s>>>(f, n);
    PSK_LAUNCH_CHECK();
}

// Launch f(i, i < n) over [0, n) with wave-uniform control flow (grid capped
// like for_each: texts of up to 2^32-2 bytes stay below the grid limit).
template <typename F>
inline void for_each_waves(Dev &d, u64 n, F f) {
    if (n == 0) return;
    u64 blocks = (n + BLOCK - 1) / BLOCK;
    if (blocks > 65536u * 4) blocks = 65536u * 4;
    k_for_each_waves<<<(u32)blocks, BLOCK, 0, d.s>>>(f, n);
    PSK_LAUNCH_CHECK();
}

inline void Dev::fill_u32(u32 *dp, u32 v, u64 count) {
    for_each(*this, count, [=] __device__(u64 i) { dp[i] = v; });
}

// ---------------------------------------------------------------------------
// 3-phase scans.  in(i) gives the u32 value of item i; out(i, prefix, v) is
// called once per item with its exclusive prefix (sum scan) or inclusive
// running value (max scan).
// ---------------------------------------------------------------------------
constexpr int SCAN_ROUNDS = 16;
constexpr u64 SCAN_TILE = (u64)BLOCK * SCAN_ROUNDS;  // 4096 items per tile
constexpr int TSCAN_IPT = 16;                       // tile sums per thread

struct OpSum {
    __device__ static u32 id() { return 0u; }
    __device__ static u32 op(u32 a, u32 b) { return a + b; }
};
struct OpMax {
    __device__ static u32 id() { return 0u; }
    __device__ static u32 op(u32 a, u32 b) { return a > b ? a : b; }
};

template <class Op>
__device__ __forceinline__ u32 wave_incl(u32 x) {
    const int l = lane_id();
    #pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        u32 y = __shfl_up(x, o, WAVE);
        if (l >= o) x = Op::op(x, y);
    }
    return x;
}

// block inclusive scan of one value per thread; *agg = block aggregate.
template <class Op>
__device__ __forceinline__ u32 block_incl(u32 v, u32 *lds, u32 *agg) {
    u32 inc = wave_incl<Op>(v);
    if (lane_id() == WAVE - 1) lds[wave_id()] = inc;
    __syncthreads();
    u32 pre = Op::id();
    for (int w = 0; w < wave_id(); w++) pre = Op::op(pre, lds[w]);
    u32 a = Op::id();
    for (int w = 0; w < NWAVES; w++) a = Op::op(a, lds[w]);
    *agg = a;
    __syncthreads();
    return Op::op(pre, inc);
}

template <class Op, class In>
__global__ void k_scan_reduce(In in, u64 n, u32 *tile_agg) {
    __shared__ u32 lds[NWAVES + 1];
    u64 base = (u64)blockIdx.x * SCAN_TILE;
    u32 acc = Op::id();
    for (int r = 0; r < SCAN_ROUNDS; r++) {
        u64 i = base + (u64)r * BLOCK + threadIdx.x;
        if (i < n) acc = Op::op(acc, in(i));
    }
    u32 agg;
    block_incl<Op>(acc, lds, &agg);
    if (threadIdx.x == 0) tile_agg[blockIdx.x] = agg;
}

// one workgroup: exclusive scan (Op) of nt tile aggregates in place; *total
// (if non-null) receives the grand aggregate.
template <class Op>
__device__ u32 scan_tiles_wg(u32 *a, u32 nt, u32 *total) {
    __shared__ u32 lds[NWAVES + 1];
    u32 carry = Op::id();
    for (u64 base = 0; base < nt; base += (u64)BLOCK * TSCAN_IPT) {
        u64 t0 = base + (u64)threadIdx.x * TSCAN_IPT;
        u32 v[TSCAN_IPT];
        u32 loc = Op::id();
        #pragma unroll
        for (int k = 0; k < TSCAN_IPT; k++) {
            v[k] = (t0 + k < nt) ? a[t0 + k] : Op::id();
            loc = Op::op(loc, v[k]);
        }
        u32 agg;
        u32 inc = block_incl<Op>(loc, lds, &agg);
        // exclusive prefix of this thread = inclusive of previous thread
        u32 ex = __shfl_up(inc, 1, WAVE);
        if (lane_id() == 0) {
            ex = Op::id();
            for (int w = 0; w < wave_id(); w++) ex = Op::op(ex, lds[w]);
        }
        __syncthreads();  // lds read above before the next round writes it
        u32 run = Op::op(carry, ex);
        #pragma unroll
        for (int k = 0; k < TSCAN_IPT; k++) {
            if (t0 + k < nt) a[t0 + k] = run;
            run = Op::op(run, v[k]);
        }
        carry = Op::op(carry, agg);
    }
    if (total && threadIdx.x == 0) *total = carry;
    return carry;
}

template <class Op>
__global__ void k_scan_tiles(u32 *a, u32 nt, u32 *total) {
    scan_tiles_wg<Op>(a, nt, total);
}

// The items [base, base + SCAN_TILE) of a scan whose tiles before this one
// combine to tbase: coalesced loads into LDS, each thread scans SCAN_ROUNDS
// consecutive items, one block scan of the thread totals, and coalesced calls
// of out in index order.  Returns the tile's aggregate.  buf holds
// SCAN_TILE + SCAN_TILE / 32 words (rows padded by one word per 32 to spread
// the per-thread runs over the banks), lds NWAVES + 1.
template <class Op, bool INCLUSIVE, class In, class Out>
__device__ __forceinline__ u32 scan_apply_tile(In in, Out out, u64 n, u64 base, u32 tbase,
                                               u32 *buf, u32 *lds) {
    u32 v[SCAN_ROUNDS];
    #pragma unroll
    for (int r = 0; r < SCAN_ROUNDS; r++) {
        const u64 i = base + (u64)r * BLOCK + threadIdx.x;
        v[r] = (i < n) ? in(i) : Op::id();
        const u32 j = (u32)r * BLOCK + threadIdx.x;
        buf[j + (j >> 5)] = v[r];
    }
    __syncthreads();
    const u32 j0 = threadIdx.x * SCAN_ROUNDS;
    u32 loc = Op::id();
    #pragma unroll
    for (int k = 0; k < SCAN_ROUNDS; k++) {
        const u32 j = j0 + k;
        loc = Op::op(loc, buf[j + (j >> 5)]);
    }
    u32 agg;
    const u32 inc = block_incl<Op>(loc, lds, &agg);
    // exclusive prefix of this thread = inclusive of the previous thread
    u32 ex = __shfl_up(inc, 1, WAVE);
    if (lane_id() == 0) {
        ex = Op::id();
        for (int w = 0; w < wave_id(); w++) ex = Op::op(ex, lds[w]);
    }
    u32 run = Op::op(tbase, ex);
    #pragma unroll
    for (int k = 0; k < SCAN_ROUNDS; k++) {
        const u32 j = j0 + k, idx = j + (j >> 5);
        const u32 x = buf[idx];
        run = Op::op(run, x);
        // exclusive prefix for a sum scan: inclusive minus own value
        buf[idx] = INCLUSIVE ? run : run - x;
    }
    __syncthreads();
    #pragma unroll
    for (int r = 0; r < SCAN_ROUNDS; r++) {
        const u64 i = base + (u64)r * BLOCK + threadIdx.x;
        const u32 j = (u32)r * BLOCK + threadIdx.x;
        if (i < n) out(i, buf[j + (j >> 5)], v[r]);
    }
    return agg;
}

// One tile of SCAN_TILE items per workgroup (scan_apply_tile).
// FUSED: tile_off holds the tiles' aggregates, not their scan; each
// workgroup combines those of the tiles before it (few tiles: a read of at
// most SCAN_FUSE_MAX words from L2 instead of a single-workgroup launch), and
// the last one writes the grand aggregate to *total (if non-null).
// SKIP (sum scans for compactions, whose out acts only on items with v != 0):
// a tile whose items sum to 0 makes no calls of out; unfused, tile_off then
// holds nt + 1 entries (the last one the grand total, which block 0 copies
// to *total).
template <class Op, bool INCLUSIVE, bool FUSED, class In, class Out, bool SKIP = false>
__global__ void __launch_bounds__(BLOCK) k_scan_apply(In in, Out out, u64 n, const u32 *tile_off,
                                                      u32 *total) {
    __shared__ u32 buf[SCAN_TILE + SCAN_TILE / 32];
    __shared__ u32 lds[NWAVES + 1];
    const u64 base = (u64)blockIdx.x * SCAN_TILE;
    u32 tbase;
    if (FUSED) {
        u32 acc = Op::id();
        for (u32 i = threadIdx.x; i < blockIdx.x; i += BLOCK) acc = Op::op(acc, tile_off[i]);
        block_incl<Op>(acc, lds, &tbase);  // includes barriers
    } else {
        tbase = tile_off[blockIdx.x];
    }
    if (SKIP) {
        if (!FUSED && total && blockIdx.x == 0 && threadIdx.x == 0) *total = tile_off[gridDim.x];
        const u32 own = FUSED ? tile_off[blockIdx.x] : tile_off[blockIdx.x + 1] - tbase;
        if (own == 0) {
            if (FUSED && total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = tbase;
            return;
        }
    }
    const u32 agg = scan_apply_tile<Op, INCLUSIVE>(in, out, n, base, tbase, buf, lds);
    if (FUSED && total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = Op::op(tbase, agg);
}

// scans of at most this many tiles combine the tile aggregates inside the
// apply kernel (two launches instead of three; the single-workgroup tile scan
// it saves costs ~24 us, the extra L2 reads of the largest scans less)
constexpr u32 SCAN_FUSE_MAX = 8192;

// Exclusive sum scan; d_total (device, may be null) receives the sum.
template <class In, class Out>
inline void scan_exclusive(Dev &d, u64 n, In in, Out out, u32 *d_total) {
    if (n == 0) {
        if (d_total) d.zero(d_total, sizeof(u32));
        return;
    }
    ArenaScope mk(*d.a);
    u32 nt = ceil_div(n, SCAN_TILE);
    u32 *ta = d.a->get<u32>(nt);
    // one tile: the fused apply reads no tile aggregates, so no reduce launch
    if (nt > 1) k_scan_reduce<OpSum><<<nt, BLOCK, 0, d.s>>>(in, n, ta);
    PSK_LAUNCH_CHECK();
    if (nt <= SCAN_FUSE_MAX) {
        k_scan_apply<OpSum, false, true><<<nt, BLOCK, 0, d.s>>>(in, out, n, ta, d_total);
    } else {
        k_scan_tiles<OpSum><<<1, BLOCK, 0, d.s>>>(ta, nt, d_total);
        PSK_LAUNCH_CHECK();
        k_scan_apply<OpSum, false, false><<<nt, BLOCK, 0, d.s>>>(in, out, n, ta, nullptr);
    }
    PSK_LAUNCH_CHECK();
    mk.reset();
}

// Positions p < nbits whose bit is set in wordfn(p / 32), ascending, into
// out; *d_total (device, may be null) receives their number.  Per tile of
// SCAN_TILE words the words' counts are scanned wave by wave and every wave
// writes its positions as one contiguous run, 64 consecutive outputs per
// store: lane j finds the word holding output j by a binary search over the
// wave's offsets and the bit by a binary select (a thread writing its own
// word's positions would store at a stride of the word's count).
template <bool FUSED, class WF, class Out>
__global__ void __launch_bounds__(BLOCK) k_bits_emit(WF wordfn, u64 nbits, u64 nwords,
                                                     const u32 *ta, Out out, u32 *total) {
    __shared__ u32 lds[NWAVES + 1];
    const u64 base = (u64)blockIdx.x * SCAN_TILE;
    u32 run;
    if (FUSED) {
        u32 acc = 0;
        for (u32 i = threadIdx.x; i < blockIdx.x; i += BLOCK) acc += ta[i];
        block_incl<OpSum>(acc, lds, &run);  // includes barriers
    } else {
        run = ta[blockIdx.x];
    }
    const int l = lane_id(), w = wave_id();
    for (int r = 0; r < SCAN_ROUNDS; r++) {
        const u64 wb = base + (u64)r * BLOCK + (u64)w * WAVE;  // the wave's first word
        const u64 wi = wb + l;
        u32 x = wi < nwords ? wordfn(wi) : 0u;
        const u64 p0 = wi * 32;
        if (p0 >= nbits) x = 0;
        else if (nbits - p0 < 32) x &= (1u << (nbits - p0)) - 1u;
        const u32 c = (u32)__popc(x);
        const u32 inc = wave_inclusive_sum(c);
        const u32 start = inc - c;
        const u32 wtot = __shfl(inc, WAVE - 1, WAVE);
        if (l == 0) lds[w] = wtot;
        __syncthreads();
        u32 woff = run, btot = 0;
        #pragma unroll
        for (int w2 = 0; w2 < NWAVES; w2++) {
            const u32 t = lds[w2];
            woff += w2 < w ? t : 0u;
            btot += t;
        }
        __syncthreads();
        run += btot;
        for (u32 j0 = 0; j0 < wtot; j0 += WAVE) {  // wave-uniform trip count
            const u32 j = j0 + (u32)l;
            int src = 0;
            #pragma unroll
            for (int step = WAVE / 2; step >= 1; step >>= 1) {
                const u32 st = __shfl(start, src + step, WAVE);
                if (st <= j) src += step;
            }
            u32 xs = __shfl(x, src, WAVE);
            u32 k = j - __shfl(start, src, WAVE);
            u32 pos = 0;
            #pragma unroll
            for (int width = 16; width >= 1; width >>= 1) {
                const u32 cnt = (u32)__popc(xs & ((1u << width) - 1u));
                if (k >= cnt) {
                    k -= cnt;
                    xs >>= width;
                    pos += (u32)width;
                }
            }
            if (j < wtot) out((u64)woff + j, (u32)((wb + (u64)src) * 32 + pos));
        }
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total = run;
}

// out(j, p): the j-th position p (for lists with a capacity of their own).
template <class WF, class Out>
inline void bit_list_f(Dev &d, u64 nbits, WF wordfn, Out out, u32 *d_total) {
    const u64 nwords = (nbits + 31) / 32;
    if (nwords == 0) {
        if (d_total) d.zero(d_total, sizeof(u32));
        return;
    }
    ArenaScope mk(*d.a);
    const u32 nt = ceil_div(nwords, SCAN_TILE);
    u32 *ta = d.a->get<u32>(nt);
    if (nt > 1) {
        auto cnt = [=] __device__(u64 w) -> u32 {
            u32 x = wordfn(w);
            const u64 p0 = w * 32;
            if (nbits - p0 < 32) x &= (1u << (nbits - p0)) - 1u;
            return (u32)__popc(x);
        };
        k_scan_reduce<OpSum><<<nt, BLOCK, 0, d.s>>>(cnt, nwords, ta);
        PSK_LAUNCH_CHECK();
    }
    if (nt <= SCAN_FUSE_MAX) {
        k_bits_emit<true><<<nt, BLOCK, 0, d.s>>>(wordfn, nbits, nwords, ta, out, d_total);
    } else {
        k_scan_tiles<OpSum><<<1, BLOCK, 0, d.s>>>(ta, nt, nullptr);
        PSK_LAUNCH_CHECK();
        k_bits_emit<false><<<nt, BLOCK, 0, d.s>>>(wordfn, nbits, nwords, ta, out, d_total);
    }
    PSK_LAUNCH_CHECK();
    mk.reset();
}

template <class WF>
inline void bit_list(Dev &d, u64 nbits, WF wordfn, u32 *out, u32 *d_total) {
    bit_list_f(d, nbits, wordfn, [=] __device__(u64 j, u32 p) { out[j] = p; }, d_total);
}

// Exclusive sum scan for a compaction: out is called only for the items of
// tiles with a nonzero sum, so it must act only on items with v != 0 (a
// sparse predicate then costs one read of the items, not two).
template <class In, class Out>
inline void compact_exclusive(Dev &d, u64 n, In in, Out out, u32 *d_total) {
    if (n == 0) {
        if (d_total) d.zero(d_total, sizeof(u32));
        return;
    }
    ArenaScope mk(*d.a);
    u32 nt = ceil_div(n, SCAN_TILE);
    u32 *ta = d.a->get<u32>((u64)nt + 1);
    if (nt > 1) k_scan_reduce<OpSum><<<nt, BLOCK, 0, d.s>>>(in, n, ta);
    PSK_LAUNCH_CHECK();
    if (nt <= SCAN_FUSE_MAX) {
        if (nt == 1) {
            // no tile aggregate was written: the one tile is never skipped
            k_scan_apply<OpSum, false, true><<<1, BLOCK, 0, d.s>>>(in, out, n, ta, d_total);
        } else {
            k_scan_apply<OpSum, false, true, In, Out, true><<<nt, BLOCK, 0, d.s>>>(in, out, n, ta,
                                                                                  d_total);
        }
    } else {
        k_scan_tiles<OpSum><<<1, BLOCK, 0, d.s>>>(ta, nt, ta + nt);
        PSK_LAUNCH_CHECK();
        k_scan_apply<OpSum, false, false, In, Out, true><<<nt, BLOCK, 0, d.s>>>(in, out, n, ta,
                                                                               d_total);
    }
    PSK_LAUNCH_CHECK();
    mk.reset();
}

// Inclusive max scan.
template <class In, class Out>
inline void scan_inclusive_max(Dev &d, u64 n, In in, Out out) {
    if (n == 0) return;
    ArenaScope mk(*d.a);
    u32 nt = ceil_div(n, SCAN_TILE);
    u32 *ta = d.a->get<u32>(nt);
    if (nt > 1) k_scan_reduce<OpMax><<<nt, BLOCK, 0, d.s>>>(in, n, ta);
    PSK_LAUNCH_CHECK();
    if (nt <= SCAN_FUSE_MAX) {
        k_scan_apply<OpMax, true, true><<<nt, BLOCK, 0, d.s>>>(in, out, n, ta, nullptr);
    } else {
        k_scan_tiles<OpMax><<<1, BLOCK, 0, d.s>>>(ta, nt, nullptr);
        PSK_LAUNCH_CHECK();
        k_scan_apply<OpMax, true, false><<<nt, BLOCK, 0, d.s>>>(in, out, n, ta, nullptr);
    }
    PSK_LAUNCH_CHECK();
    mk.reset();
}

// ---------------------------------------------------------------------------
// Stable 8-bit multi-split.  For i in [0, n): dig(i, &digit) returns whether
// item i takes part and its digit (< 256).  Phase 1 counts digits per tile,
// phase 2 scans the digit-major count matrix, phase 3 calls
// put(i, digit, gpos, dstart) where gpos is the item's position in the
// stable order by digit and dstart the first position of its digit, so
// gpos - dstart is its stable rank among items of that digit.
// ---------------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8), and each XCD has its own L2.  xcd_tile maps blocks so that each
// XCD processes one contiguous range of tiles: tiles that write neighbouring
// output (adjacent digit segments, adjacent count-matrix entries) then share
// an L2, which merges their partial cache lines before they reach HBM.  The
// map is a bijection of [0, nt).
constexpr u32 NXCD = 8;
__device__ __forceinline__ u32 xcd_tile(u32 b, u32 nt) {
    u32 q = nt / NXCD, r = nt % NXCD;
    u32 x = b % NXCD, i = b / NXCD;
    return x * q + (x < r ? x : r) + i;
}

// ---------------------------------------------------------------------------
// Device-sized launches.  A count that only the device knows (the seeds of a
// level-0 bucket queue, its deep chain entries, a count matrix's size) is
// read by the kernels themselves, so the host needs no readback to size the
// launch: the grid comes from an upper bound the host does know, capped at
// DN_GRID workgroups, and the workgroups stride over the tiles below the
// device count.  (A launch whose bound is close to the device count, with
// latency-bound tiles, takes one workgroup per tile: cap = the tile count.)
// The cap is a multiple of NXCD, so every tile index a block
// strides over is congruent to the block index mod NXCD and xcd_tile keeps
// each XCD on one contiguous range of tiles.  A device count of 0 costs a
// launch of empty workgroups (a few microseconds).
// ---------------------------------------------------------------------------
#ifndef PSK_DN_GRID
#define PSK_DN_GRID 2048
#endif
constexpr u32 DN_GRID = PSK_DN_GRID;  // 2048: 8 workgroups per CU
inline u32 dn_grid(u64 tiles, u64 cap = DN_GRID) {
    u64 g = tiles < cap ? tiles : cap;
    g = (g + NXCD - 1) / NXCD * NXCD;
    return g ? (u32)g : NXCD;
}

template <class In>
__global__ void __launch_bounds__(BLOCK) k_scan_reduce_dn(In in, const u32 *dn, u32 *tile_agg) {
    __shared__ u32 lds[NWAVES + 1];
    const u64 n = *dn;
    const u64 nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    for (u64 t = blockIdx.x; t < nt; t += gridDim.x) {
        const u64 base = t * SCAN_TILE;
        u32 acc = 0;
        for (int r = 0; r < SCAN_ROUNDS; r++) {
            const u64 i = base + (u64)r * BLOCK + threadIdx.x;
            if (i < n) acc += in(i);
        }
        u32 agg;
        block_incl<OpSum>(acc, lds, &agg);  // ends with a barrier: lds is free again
        if (threadIdx.x == 0) tile_agg[t] = agg;
    }
}

// tail (optional): tail[*dn] receives the total as well (a scanned count
// matrix whose last entry's count is read as the difference to the next)
__global__ void k_scan_tiles_dn(u32 *a, const u32 *dn, u32 *total, u32 *tail) {
    const u64 n = *dn;
    const u32 c = scan_tiles_wg<OpSum>(a, (u32)((n + SCAN_TILE - 1) / SCAN_TILE), total);
    if (tail && threadIdx.x == 0) tail[n] = c;
}

template <class In, class Out>
__global__ void __launch_bounds__(BLOCK) k_scan_apply_dn(In in, Out out, const u32 *dn,
                                                         const u32 *tile_off) {
    __shared__ u32 buf[SCAN_TILE + SCAN_TILE / 32];
    __shared__ u32 lds[NWAVES + 1];
    const u64 n = *dn;
    const u64 nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    for (u64 t = blockIdx.x; t < nt; t += gridDim.x) {
        if (t != blockIdx.x) __syncthreads();  // the previous tile's reads of buf are done
        scan_apply_tile<OpSum, false>(in, out, n, t * SCAN_TILE, tile_off[t], buf, lds);
    }
}

// Exclusive sum scan over [0, *dn), *dn <= ub: as scan_exclusive, with d_total
// (may be null) and tail (may be null, see k_scan_tiles_dn) receiving the sum.
template <class In, class Out>
inline void scan_exclusive_dn(Dev &d, u64 ub, const u32 *dn, In in, Out out, u32 *d_total,
                              u32 *tail = nullptr) {
    ArenaScope mk(*d.a);
    const u64 ntb = std::max<u64>(ceil_div(ub, SCAN_TILE), 1);
    u32 *ta = d.a->get<u32>(ntb);
    const u32 g = dn_grid(ntb);
    k_scan_reduce_dn<<<g, BLOCK, 0, d.s>>>(in, dn, ta);
    PSK_LAUNCH_CHECK();
    k_scan_tiles_dn<<<1, BLOCK, 0, d.s>>>(ta, dn, d_total, tail);
    PSK_LAUNCH_CHECK();
    k_scan_apply_dn<<<g, BLOCK, 0, d.s>>>(in, out, dn, ta);
    PSK_LAUNCH_CHECK();
    mk.reset();
}

#ifndef PSK_MS_ROUNDS
#define PSK_MS_ROUNDS 16
#endif
constexpr int MS_ROUNDS = PSK_MS_ROUNDS;
constexpr u64 MS_TILE = (u64)BLOCK * MS_ROUNDS;

// LDS count of one digit per lane: up to LDS_COUNT_LEADERS distinct digits
// of the wave are added by one lane each (few distinct digits are the common
// case for byte buckets: a DNA bucket's predecessor bytes plus the "not
// induced" digit are five), the rest with plain atomics.
#ifndef PSK_LDS_LEADERS
#define PSK_LDS_LEADERS 8
#endif
constexpr int LDS_COUNT_LEADERS = PSK_LDS_LEADERS;
__device__ __forceinline__ void lds_count_wave(u32 *h, u32 x, bool valid) {
    bool pending = valid;
    #pragma unroll
    for (int it = 0; it < LDS_COUNT_LEADERS; it++) {
        const u64 act = ballot(pending);
        if (!act) return;
        const int leader = __ffsll((long long)act) - 1;
        const u32 xl = __shfl(x, leader, WAVE);
        const bool mine = pending && x == xl;
        const u64 peers = ballot(mine);
        if (lane_id() == leader) atomicAdd(&h[xl], (u32)__popcll(peers));
        if (mine) pending = false;
    }
    if (pending) atomicAdd(&h[x], 1u);
}

// Digit counts of tile t of nt (hist is digit-major: [digit][tile]).  h: 256
// words of LDS.
template <class Dig>
__device__ __forceinline__ void ms_hist_tile(Dig dig, u64 n, u32 *hist, u32 nt, u32 t, u32 *h) {
    h[threadIdx.x] = 0;
    __syncthreads();
    u64 base = (u64)t * MS_TILE;
    for (int r = 0; r < MS_ROUNDS; r++) {
        u64 i = base + (u64)r * BLOCK + threadIdx.x;
        u32 key = 0;
        bool v = (i < n) && dig(i, &key);
        lds_count_wave(h, key, v);
    }
    __syncthreads();
    hist[(u64)threadIdx.x * nt + t] = h[threadIdx.x];
}

template <class Dig>
__global__ void k_ms_hist(Dig dig, u64 n, u32 *hist, u32 nt) {
    __shared__ u32 h[256];
    ms_hist_tile(dig, n, hist, nt, xcd_tile(blockIdx.x, nt), h);
}

// Stable scatter of tile t of nt; off: the scanned count matrix.
template <class Dig, class Put>
__device__ __forceinline__ void ms_scatter_tile(Dig dig, Put put, u64 n, const u32 *off, u32 nt,
                                                u32 t, u32 *run, u32 *dst0, u32 (*wc)[256]) {
    run[threadIdx.x] = off[(u64)threadIdx.x * nt + t];
    dst0[threadIdx.x] = off[(u64)threadIdx.x * nt];
    #pragma unroll
    for (int w = 0; w < NWAVES; w++) wc[w][threadIdx.x] = 0;
    __syncthreads();
    u64 base = (u64)t * MS_TILE;
    const int w = wave_id();
    for (int r = 0; r < MS_ROUNDS; r++) {
        u64 i = base + (u64)r * BLOCK + threadIdx.x;
        u32 key = 0;
        bool v = (i < n) && dig(i, &key);
        u64 peers = match_key8(v, key);
        u32 rk = (u32)__popcll(peers & lanemask_lt());
        if (v && rk == 0) wc[w][key] = (u32)__popcll(peers);
        __syncthreads();
        if (v) {
            u32 pos = run[key] + rk;
            for (int w2 = 0; w2 < w; w2++) pos += wc[w2][key];
            put(i, key, pos, dst0[key]);
        }
        __syncthreads();
        {
            u32 s = 0;
            #pragma unroll
            for (int w2 = 0; w2 < NWAVES; w2++) { s += wc[w2][threadIdx.x]; wc[w2][threadIdx.x] = 0; }
            run[threadIdx.x] += s;
        }
        __syncthreads();
    }
}

template <class Dig, class Put>
__global__ void k_ms_scatter(Dig dig, Put put, u64 n, const u32 *off, u32 nt) {
    __shared__ u32 run[256];
    __shared__ u32 dst0[256];
    __shared__ u32 wc[NWAVES][256];
    ms_scatter_tile(dig, put, n, off, nt, xcd_tile(blockIdx.x, nt), run, dst0, wc);
}

// The same with the item count *dn read on the device (device-sized
// launches); k_ms_hist_dn also leaves the count matrix's size in *msize.
template <class Dig>
__global__ void k_ms_hist_dn(Dig dig, const u32 *dn, u32 *hist, u32 *msize) {
    __shared__ u32 h[256];
    const u64 n = *dn;
    const u32 nt = (u32)((n + MS_TILE - 1) / MS_TILE);
    if (blockIdx.x == 0 && threadIdx.x == 0) *msize = 256 * nt;
    for (u32 v = blockIdx.x; v < nt; v += gridDim.x) {
        if (v != blockIdx.x) __syncthreads();
        ms_hist_tile(dig, n, hist, nt, xcd_tile(v, nt), h);
    }
}

template <class Dig, class Put>
__global__ void k_ms_scatter_dn(Dig dig, Put put, const u32 *dn, const u32 *off) {
    __shared__ u32 run[256];
    __shared__ u32 dst0[256];
    __shared__ u32 wc[NWAVES][256];
    const u64 n = *dn;
    const u32 nt = (u32)((n + MS_TILE - 1) / MS_TILE);
    for (u32 v = blockIdx.x; v < nt; v += gridDim.x) {
        if (v != blockIdx.x) __syncthreads();
        ms_scatter_tile(dig, put, n, off, nt, xcd_tile(v, nt), run, dst0, wc);
    }
}

// Runs the three phases; if d_counts is non-null it receives the per-digit
// totals (256 u32) on the device.  dig_count is used by the counting phase
// and dig by the scatter phase; they must agree (the counting phase may
// cache its digits for the scatter phase to read back).
// Profiling: tag != nullptr names the two kernels "<tag>_count" and
// "<tag>_scatter" with bytes_count / bytes_scatter algorithmic bytes; the
// scatter's record index is returned (for data-dependent byte slots).
template <class DigC, class Dig, class Put>
inline int multisplit2(Dev &d, u64 n, DigC dig_count, Dig dig, Put put, u32 *d_counts,
                       const char *tag = nullptr, double bytes_count = 0,
                       double bytes_scatter = 0) {
    if (n == 0) {
        if (d_counts) d.zero(d_counts, 256 * sizeof(u32));
        return -1;
    }
    ArenaScope mk(*d.a);
    u32 nt = ceil_div(n, MS_TILE);
    u32 *hist = d.a->get<u32>((u64)256 * nt + 1);
    std::string tc = tag ? std::string(tag) + "_count" : std::string();
    std::string ts = tag ? std::string(tag) + "_scatter" : std::string();
    int pr = tag ? d.pbegin(tc.c_str(), bytes_count) : -1;
    k_ms_hist<<<nt, BLOCK, 0, d.s>>>(dig_count, n, hist, nt);
    PSK_LAUNCH_CHECK();
    d.pend(pr);
    u64 total = (u64)256 * nt;
    scan_exclusive(
        d, total, [=] __device__(u64 i) { return hist[i]; },
        [=] __device__(u64 i, u32 ex, u32) { hist[i] = ex; }, hist + total);
    if (d_counts) {
        // per-digit totals from the scanned matrix: start of d+1 minus start of d
        u32 *dc = d_counts;
        for_each(d, 256, [=] __device__(u64 k) {
            dc[k] = hist[(k + 1) * nt] - hist[k * nt];
        });
    }
    int ps = tag ? d.pbegin(ts.c_str(), bytes_scatter) : -1;
    k_ms_scatter<<<nt, BLOCK, 0, d.s>>>(dig, put, n, hist, nt);
    PSK_LAUNCH_CHECK();
    d.pend(ps);
    mk.reset();
    return ps;
}

template <class Dig, class Put>
inline int multisplit(Dev &d, u64 n, Dig dig, Put put, u32 *d_counts, const char *tag = nullptr,
                      double bytes_count = 0, double bytes_scatter = 0) {
    return multisplit2(d, n, dig, dig, put, d_counts, tag, bytes_count, bytes_scatter);
}

// Stable multi-split of [0, *dn) (*dn <= ub) with device-side sizes: no
// host readback (dig and put as in multisplit).
template <class Dig, class Put>
inline void multisplit_dn(Dev &d, u64 ub, const u32 *dn, Dig dig, Put put) {
    if (ub == 0) return;
    ArenaScope mk(*d.a);
    const u64 ntb = ceil_div(ub, MS_TILE);
    u32 *hist = d.a->get<u32>(256 * ntb + 1);
    u32 *msize = d.a->get<u32>(1);
    const u32 g = dn_grid(ntb);
    k_ms_hist_dn<<<g, BLOCK, 0, d.s>>>(dig, dn, hist, msize);
    PSK_LAUNCH_CHECK();
    scan_exclusive_dn(
        d, 256 * ntb, msize, [=] __device__(u64 i) { return hist[i]; },
        [=] __device__(u64 i, u32 ex, u32) { hist[i] = ex; }, nullptr);
    k_ms_scatter_dn<<<g, BLOCK, 0, d.s>>>(dig, put, dn, hist);
    PSK_LAUNCH_CHECK();
    mk.reset();
}

// ---------------------------------------------------------------------------
// LSD radix sort of (u64 key, u32 value) pairs on the low `bits` key bits,
// 8-bit digits.  Per pass: a per-tile digit count, one scan of the
// digit-major count matrix, and a scatter that first reorders its tile by
// digit in LDS (stable), so each digit's items leave the tile as one
// contiguous, coalesced run.
// Returns true if the sorted result ended in (kt, vt) rather than (k, v).
// ---------------------------------------------------------------------------
#ifndef PSK_RS_ROUNDS
#define PSK_RS_ROUNDS 8
#endif
#ifndef PSK_RS_CT
#define PSK_RS_CT 1
#endif
constexpr int RS_ROUNDS = PSK_RS_ROUNDS;
constexpr u64 RS_TILE = (u64)BLOCK * RS_ROUNDS;  // 2048 pairs per tile

// Counting: one block per RS_CT consecutive tiles (fewer, longer-lived
// workgroups; each digit's RS_CT counts are written as one contiguous run).
constexpr u32 RS_CT = PSK_RS_CT;
constexpr int RS_COPIES = 8;  // LDS copies of a tile's digit counts (power of two)
__global__ void __launch_bounds__(BLOCK) k_rs_count(const u64 *__restrict__ ki, u64 n, int shift,
                                                    u32 *hist, u32 nt) {
    // eight copies of each tile's counts (lane & 7), padded onto different
    // banks: lanes of a wave that share a digit (grouped keys) spread their
    // atomics over eight counters instead of serialising on one
    __shared__ u32 h[RS_CT][RS_COPIES][257];
    #pragma unroll
    for (u32 c = 0; c < RS_CT; c++)
        #pragma unroll
        for (int q = 0; q < RS_COPIES; q++) h[c][q][threadIdx.x] = 0;
    const u32 ng = (nt + RS_CT - 1) / RS_CT;
    const u32 t0 = xcd_tile(blockIdx.x, ng) * RS_CT;
    __syncthreads();
    // all loads of the block's tiles first (independent, in flight together)
    u32 dk[RS_CT][RS_ROUNDS];
    #pragma unroll
    for (u32 c = 0; c < RS_CT; c++)
        #pragma unroll
        for (int r = 0; r < RS_ROUNDS; r++) {
            u64 i = (u64)(t0 + c) * RS_TILE + (u64)r * BLOCK + threadIdx.x;
            dk[c][r] = i < n ? (u32)(ki[i] >> shift) & 255u : 256u;
        }
    // plain LDS atomics, except when the whole wave holds one digit (the high
    // digits of grouped keys): then one lane adds the wave's count, instead
    // of 64 lanes serialising on one counter
    #pragma unroll
    for (u32 c = 0; c < RS_CT; c++)
        #pragma unroll
        for (int r = 0; r < RS_ROUNDS; r++) {
            const u32 x = dk[c][r];
            const u32 x0 = __shfl(x, 0, WAVE);
            if (ballot(x == x0) == ~0ull) {
                if (lane_id() == 0 && x0 < 256u) atomicAdd(&h[c][0][x0], (u32)WAVE);
            } else if (x < 256u) {
                atomicAdd(&h[c][lane_id() & (RS_COPIES - 1)][x], 1u);
            }
        }
    __syncthreads();
    for (u32 c = 0; c < RS_CT && t0 + c < nt; c++) {
        u32 sum = 0;
        #pragma unroll
        for (int q = 0; q < RS_COPIES; q++) sum += h[c][q][threadIdx.x];
        hist[(u64)threadIdx.x * nt + t0 + c] = sum;
    }
}

// off: exclusive scan of the digit-major count matrix (256*nt + 1 entries,
// the last one the total), so the count of (digit d, tile t) is
// off[d*nt+t+1] - off[d*nt+t].  Ranking: every round's wave-level digit
// counts go to their own LDS slice, one pass per digit turns them into
// tile-local offsets in (round, wave) order, and the tile is staged in LDS
// in digit order before the coalesced write-out.
// SMALL: off holds the unscanned counts (at most RS_FUSE_SCAN tiles, an
// L2-resident matrix) and every workgroup derives its own offsets: the count
// of its digit in the tiles before it plus the digit's start, a block scan of
// the digits' totals.  No scan launches.
constexpr u32 RS_FUSE_SCAN = 64;
template <bool SMALL>
__global__ void __launch_bounds__(BLOCK) k_rs_scatter(const u64 *__restrict__ ki,
                                                      const u32 *__restrict__ vi,
                                                      u64 *__restrict__ ko, u32 *__restrict__ vo,
                                                      u64 n, int shift, const u32 *off, u32 nt) {
    // The (round, wave, digit) counts are dead once every item holds its
    // staging slot, so the staged values reuse their memory: 34 KB of LDS per
    // workgroup, four workgroups per CU (three with separate arrays; a 32 KB
    // layout with five measured slower)
    __shared__ u64 sk[RS_TILE];
    __shared__ union {
        unsigned short wcnt[RS_ROUNDS][NWAVES][256];
        u32 sv[RS_TILE];
    } sw;
    __shared__ u32 tstart[256], gbase[256];
    __shared__ u32 red[NWAVES + 1];
    auto &wcnt = sw.wcnt;
    const u32 t = xcd_tile(blockIdx.x, nt);
    const u32 dg = threadIdx.x;
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++)
        #pragma unroll
        for (int w = 0; w < NWAVES; w++) wcnt[r][w][dg] = 0;
    u32 my_start, my_base;  // this thread's digit: start in the tile, global start
    {
        u32 g, c;
        if (SMALL) {
            const u32 *row = off + (u64)dg * nt;
            u32 all = 0, pre = 0;
            for (u32 t2 = 0; t2 < nt; t2++) {
                const u32 x = row[t2];
                all += x;
                pre += t2 < t ? x : 0u;
            }
            c = row[t];
            u32 sum;
            g = block_exclusive_sum(all, red, &sum) + pre;  // includes barriers
        } else {
            u64 o = (u64)dg * nt + t;
            g = off[o];
            c = off[o + 1] - g;
        }
        u32 tot;
        my_start = block_exclusive_sum(c, red, &tot);  // includes a barrier
        my_base = g;
    }
    const u64 base = (u64)t * RS_TILE;
    const int w = wave_id();
    u64 key[RS_ROUNDS];
    u32 val[RS_ROUNDS];
    u32 dk[RS_ROUNDS];
    unsigned short rk[RS_ROUNDS];
    bool vv[RS_ROUNDS];
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        u64 i = base + (u64)r * BLOCK + threadIdx.x;
        vv[r] = i < n;
        key[r] = vv[r] ? ki[i] : 0ull;
        val[r] = vv[r] ? vi[i] : 0u;
    }
    __syncthreads();  // wcnt zeroed
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        dk[r] = (u32)(key[r] >> shift) & 255u;
        u64 peers = match_key8(vv[r], dk[r]);
        rk[r] = (unsigned short)__popcll(peers & lanemask_lt());
        if (vv[r] && rk[r] == 0) wcnt[r][w][dk[r]] = (unsigned short)__popcll(peers);
    }
    __syncthreads();
    {
        u32 run = my_start;
        #pragma unroll
        for (int r = 0; r < RS_ROUNDS; r++)
            #pragma unroll
            for (int w2 = 0; w2 < NWAVES; w2++) {
                u32 c = wcnt[r][w2][dg];
                wcnt[r][w2][dg] = (unsigned short)run;
                run += c;
            }
    }
    __syncthreads();
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) rk[r] = (unsigned short)(wcnt[r][w][dk[r]] + rk[r]);
    __syncthreads();  // every count read: sv takes their memory
    u32 *sv = sw.sv;
    tstart[dg] = my_start;
    gbase[dg] = my_base;
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        if (vv[r]) {
            sk[rk[r]] = key[r];
            sv[rk[r]] = val[r];
        }
    }
    __syncthreads();
    u64 cnt = n - base < RS_TILE ? n - base : RS_TILE;
    for (u32 j = threadIdx.x; j < cnt; j += BLOCK) {
        u64 k = sk[j];
        u32 d = (u32)(k >> shift) & 255u;
        u64 g = (u64)gbase[d] + (j - tstart[d]);
        ko[g] = k;
        vo[g] = sv[j];
    }
}

// Small sorts (n <= RS_SMALL_MAX): every pass in one workgroup and one
// launch.  A pass counts all digits in LDS, then ranks the tiles in order as
// k_rs_scatter does, carrying each digit's next output position from tile to
// tile.  The buffers swap after every pass, as in radix_sort_kv.  Tiles run
// one after another (a few microseconds each), so past a few tiles the
// multi-workgroup passes win despite their four launches each.
constexpr u64 RS_SMALL_MAX = 4 * RS_TILE;
__global__ void __launch_bounds__(BLOCK) k_rs_small(u64 *k, u32 *v, u64 *kt, u32 *vt, u32 n,
                                                    int passes) {
    __shared__ u64 sk[RS_TILE];
    __shared__ u32 sv[RS_TILE];
    __shared__ u32 tstart[256], gbase[256], hist[256];
    __shared__ unsigned short wcnt[RS_ROUNDS][NWAVES][256];
    __shared__ u32 red[NWAVES + 1];
    const u32 dg = threadIdx.x;
    const int w = wave_id();
    const u32 nt = (u32)((n + RS_TILE - 1) / RS_TILE);
    for (int pass = 0; pass < passes; pass++) {
        const int shift = 8 * pass;
        const u64 *ki = (pass & 1) ? kt : k;
        const u32 *vi = (pass & 1) ? vt : v;
        u64 *ko = (pass & 1) ? k : kt;
        u32 *vo = (pass & 1) ? v : vt;
        hist[dg] = 0;
        __syncthreads();
        for (u32 i = threadIdx.x; i < n; i += BLOCK) atomicAdd(&hist[(u32)(ki[i] >> shift) & 255u], 1u);
        __syncthreads();
        {
            u32 tot;
            gbase[dg] = block_exclusive_sum(hist[dg], red, &tot);  // includes a barrier
        }
        for (u32 t = 0; t < nt; t++) {
            #pragma unroll
            for (int r = 0; r < RS_ROUNDS; r++)
                #pragma unroll
                for (int w2 = 0; w2 < NWAVES; w2++) wcnt[r][w2][dg] = 0;
            const u64 base = (u64)t * RS_TILE;
            u64 key[RS_ROUNDS];
            u32 val[RS_ROUNDS];
            u32 dk[RS_ROUNDS];
            unsigned short rk[RS_ROUNDS];
            bool vv[RS_ROUNDS];
            #pragma unroll
            for (int r = 0; r < RS_ROUNDS; r++) {
                u64 i = base + (u64)r * BLOCK + threadIdx.x;
                vv[r] = i < n;
                key[r] = vv[r] ? ki[i] : 0ull;
                val[r] = vv[r] ? vi[i] : 0u;
            }
            __syncthreads();  // wcnt zeroed
            #pragma unroll
            for (int r = 0; r < RS_ROUNDS; r++) {
                dk[r] = (u32)(key[r] >> shift) & 255u;
                u64 peers = match_key8(vv[r], dk[r]);
                rk[r] = (unsigned short)__popcll(peers & lanemask_lt());
                if (vv[r] && rk[r] == 0) wcnt[r][w][dk[r]] = (unsigned short)__popcll(peers);
            }
            __syncthreads();
            u32 c = 0;
            #pragma unroll
            for (int r = 0; r < RS_ROUNDS; r++)
                #pragma unroll
                for (int w2 = 0; w2 < NWAVES; w2++) c += wcnt[r][w2][dg];
            u32 tot;
            const u32 ts = block_exclusive_sum(c, red, &tot);  // includes a barrier
            tstart[dg] = ts;
            {
                u32 run = ts;
                #pragma unroll
                for (int r = 0; r < RS_ROUNDS; r++)
                    #pragma unroll
                    for (int w2 = 0; w2 < NWAVES; w2++) {
                        u32 x = wcnt[r][w2][dg];
                        wcnt[r][w2][dg] = (unsigned short)run;
                        run += x;
                    }
            }
            __syncthreads();
            #pragma unroll
            for (int r = 0; r < RS_ROUNDS; r++) {
                if (vv[r]) {
                    u32 lp = (u32)wcnt[r][w][dk[r]] + rk[r];
                    sk[lp] = key[r];
                    sv[lp] = val[r];
                }
            }
            __syncthreads();
            const u32 cnt = n - base < RS_TILE ? (u32)(n - base) : (u32)RS_TILE;
            for (u32 j = threadIdx.x; j < cnt; j += BLOCK) {
                u64 kk = sk[j];
                u32 d = (u32)(kk >> shift) & 255u;
                u32 g = gbase[d] + (j - tstart[d]);
                ko[g] = kk;
                vo[g] = sv[j];
            }
            __syncthreads();  // sk, sv, tstart, gbase read
            gbase[dg] += c;
        }
        // the pass's output is the next pass's input, in this workgroup
        __threadfence_block();
        __syncthreads();
    }
}

inline bool radix_sort_kv(Dev &d, u64 n, u64 *k, u32 *v, u64 *kt, u32 *vt, int bits) {
    bool swapped = false;
    if (n == 0) return false;
    if (n <= RS_SMALL_MAX) {
        const int passes = (bits + 7) / 8;
        if (passes == 0) return false;
        int ps = d.pbegin("k_rs_small", 24.0 * (double)n * passes);
        k_rs_small<<<1, BLOCK, 0, d.s>>>(k, v, kt, vt, (u32)n, passes);
        PSK_LAUNCH_CHECK();
        d.pend(ps);
        return (passes & 1) != 0;
    }
    ArenaScope mk(*d.a);
    const u32 nt = ceil_div(n, RS_TILE);
    u32 *hist = d.a->get<u32>((u64)256 * nt + 1);
    const u64 total = (u64)256 * nt;
    for (int shift = 0; shift < bits; shift += 8) {
        const u64 *ki = swapped ? kt : k;
        const u32 *vi = swapped ? vt : v;
        u64 *ko = swapped ? k : kt;
        u32 *vo = swapped ? v : vt;
        int pc = d.pbegin("k_rs_count", 8.0 * n);
        k_rs_count<<<ceil_div(nt, RS_CT), BLOCK, 0, d.s>>>(ki, n, shift, hist, nt);
        PSK_LAUNCH_CHECK();
        d.pend(pc);
        const bool small = nt <= RS_FUSE_SCAN;
        if (!small)
            scan_exclusive(
                d, total, [=] __device__(u64 i) { return hist[i]; },
                [=] __device__(u64 i, u32 ex, u32) { hist[i] = ex; }, hist + total);
        int ps = d.pbegin("k_rs_scatter", 24.0 * n);
        if (small) k_rs_scatter<true><<<nt, BLOCK, 0, d.s>>>(ki, vi, ko, vo, n, shift, hist, nt);
        else k_rs_scatter<false><<<nt, BLOCK, 0, d.s>>>(ki, vi, ko, vo, n, shift, hist, nt);
        PSK_LAUNCH_CHECK();
        d.pend(ps);
        swapped = !swapped;
    }
    mk.reset();
    return swapped;
}

inline int bit_width(u64 x) {
    int b = 0;
    while (x) { b++; x >>= 1; }
    return b;
}

// ---------------------------------------------------------------------------
// Group sort.  Items [0, A) come in G contiguous groups (group g holds items
// [start[g], start[g+1]), groups in order), as the tie groups of a
// refinement round do.  Every group is sorted stably by the key bits
// [0, lob) and stays where it is, so the result equals a stable sort of the
// whole array by (group, low bits) -- without sorting on the group bits:
//  * groups of at most GS_LOCAL items are ranked by counting inside one
//    workgroup (each item counts the members that sort before it);
//  * larger groups go through stable LSD passes over the low bits alone,
//    with tiles that never cross a group boundary and a count matrix laid
//    out group by group ([group][digit][tile of the group]), so one global
//    scan gives every (digit, tile) its position inside its group.
// ---------------------------------------------------------------------------
constexpr u32 GS_LOCAL = 256;  // largest group ranked by counting
constexpr u32 GS_WIN = 1024;   // group starts per ranking workgroup

// Tiles of the large groups: tile t belongs to group tg[t], is that group's
// tile t - tb[g], and covers its items [start[g] + RS_TILE (t - tb[g]), ...).
struct GroupTiles {
    const u32 *start, *tb, *tg, *bias;
    __device__ __forceinline__ void tile(u32 t, u64 &base, u32 &cnt, u64 &mbase, u32 &mstride,
                                         u64 &obias) const {
        const u32 g = tg[t];
        const u32 j = t - tb[g];
        const u32 s0 = start[g], s1 = start[g + 1];
        const u32 nt = tb[g + 1] - tb[g];
        base = (u64)s0 + (u64)j * RS_TILE;
        const u64 rem = (u64)s1 - base;
        cnt = (u32)(rem < RS_TILE ? rem : RS_TILE);
        mbase = (u64)tb[g] * 256 + j;
        mstride = nt;
        obias = bias[g];
    }
};

// Ranking of the small groups whose first item lies in this workgroup's
// window; their items move to start[g] + rank (in place or into ko/vo).
__global__ void __launch_bounds__(BLOCK) k_gs_local(const u64 *K, const u32 *V, u64 *ko, u32 *vo,
                                                    u64 A, const u32 *gid, u32 gmask,
                                                    const u32 *start, u64 lomask) {
    constexpr u32 CAP = GS_WIN + GS_LOCAL;
    __shared__ u64 sk[CAP];
    __shared__ u32 sv[CAP];
    __shared__ unsigned short sg0[CAP], sg1[CAP];  // the item's group, local bounds
    const u64 w0 = (u64)blockIdx.x * GS_WIN;
    if (w0 >= A) return;
    const u64 w1 = w0 + GS_WIN < A ? w0 + GS_WIN : A;
    const u32 ga = gid[w0] & gmask;
    const u32 gfirst = start[ga] == w0 ? ga : ga + 1;
    const u32 glast = gid[w1 - 1] & gmask;
    if (gfirst > glast) return;  // one group covers the whole window
    const u64 lo = start[gfirst];
    u64 hi = start[glast + 1];
    // the last group may run past the window: a small one is ranked here
    // (at most GS_LOCAL - 1 items past it), a large one elsewhere
    if (hi - start[glast] > GS_LOCAL) hi = start[glast];
    const u32 nl = (u32)(hi - lo);
    for (u32 i = threadIdx.x; i < nl; i += BLOCK) {
        const u64 p = lo + i;
        const u32 g = gid[p] & gmask;
        const u32 a = start[g], b = start[g + 1];
        if (b - a <= GS_LOCAL) {
            sk[i] = K[p];
            sv[i] = V[p];
            sg0[i] = (unsigned short)(a - lo);
            sg1[i] = (unsigned short)(b - lo);
        } else {
            sg0[i] = sg1[i] = 0xFFFF;  // a large group's item: not ranked here
        }
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < nl; i += BLOCK) {
        const u32 a = sg0[i];
        if (a == 0xFFFF) continue;
        const u32 b = sg1[i];
        const u64 key = sk[i], mine = key & lomask;
        u32 r = 0;
        for (u32 j = a; j < b; j++) {
            const u64 x = sk[j] & lomask;
            r += (x < mine) || (x == mine && j < i);
        }
        ko[lo + a + r] = key;
        vo[lo + a + r] = sv[i];
    }
}

// Digit counts of one group tile per workgroup: hist[mbase + d * mstride].
template <class Tiles>
__global__ void __launch_bounds__(BLOCK) k_gs_count(const u64 *__restrict__ ki, int shift,
                                                    u32 *hist, u32 nt, Tiles tiles) {
    __shared__ u32 h[RS_COPIES][257];  // copies as in k_rs_count
    #pragma unroll
    for (int q = 0; q < RS_COPIES; q++) h[q][threadIdx.x] = 0;
    const u32 t = xcd_tile(blockIdx.x, nt);
    u64 base, mbase, ob;
    u32 cnt, ms;
    tiles.tile(t, base, cnt, mbase, ms, ob);
    __syncthreads();
    for (u32 r = 0; r < RS_ROUNDS; r++) {
        const u32 j = r * BLOCK + threadIdx.x;
        const u32 x = j < cnt ? (u32)(ki[base + j] >> shift) & 255u : 256u;
        const u32 x0 = __shfl(x, 0, WAVE);
        if (ballot(x == x0) == ~0ull) {
            if (lane_id() == 0 && x0 < 256u) atomicAdd(&h[0][x0], (u32)WAVE);
        } else if (x < 256u) {
            atomicAdd(&h[lane_id() & (RS_COPIES - 1)][x], 1u);
        }
    }
    __syncthreads();
    u32 sum = 0;
    #pragma unroll
    for (int q = 0; q < RS_COPIES; q++) sum += h[q][threadIdx.x];
    hist[mbase + (u64)threadIdx.x * ms] = sum;
}

// Stable scatter of one group tile (the ranking of k_rs_scatter): off holds
// the exclusive scan of the group-ordered count matrix (plus the total), so
// the count of (digit, tile) is the difference to the next entry.
template <class Tiles>
__global__ void __launch_bounds__(BLOCK) k_gs_scatter(const u64 *__restrict__ ki,
                                                      const u32 *__restrict__ vi,
                                                      u64 *__restrict__ ko, u32 *__restrict__ vo,
                                                      int shift, const u32 *off, u32 nt,
                                                      Tiles tiles) {
    // LDS laid out as in k_rs_scatter (sv over the dead counts)
    __shared__ u64 sk[RS_TILE];
    __shared__ union {
        unsigned short wcnt[RS_ROUNDS][NWAVES][256];
        u32 sv[RS_TILE];
    } sw;
    __shared__ u32 tstart[256];
    __shared__ u64 gbase[256];
    __shared__ u32 red[NWAVES + 1];
    auto &wcnt = sw.wcnt;
    const u32 t = xcd_tile(blockIdx.x, nt);
    u64 base, mbase, ob;
    u32 cnt, ms;
    tiles.tile(t, base, cnt, mbase, ms, ob);
    const u32 dg = threadIdx.x;
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++)
        #pragma unroll
        for (int w = 0; w < NWAVES; w++) wcnt[r][w][dg] = 0;
    u32 my_start;
    u64 my_base;
    {
        // a group's entries are tb*256 + d*nt_g + j: the entry after (d, j)
        // is (d, j+1), or (d+1, 0), or the next group's first entry (the
        // appended total after the last group) -- always the next index
        const u64 o = mbase + (u64)dg * ms;
        const u32 g = off[o];
        const u32 cn = off[o + 1] - g;
        u32 tot;
        my_start = block_exclusive_sum(cn, red, &tot);  // includes a barrier
        my_base = (u64)g + ob;
    }
    const int w = wave_id();
    u64 key[RS_ROUNDS];
    u32 val[RS_ROUNDS];
    u32 dk[RS_ROUNDS];
    unsigned short rk[RS_ROUNDS];
    bool vv[RS_ROUNDS];
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const u32 j = (u32)r * BLOCK + threadIdx.x;
        vv[r] = j < cnt;
        key[r] = vv[r] ? ki[base + j] : 0ull;
        val[r] = vv[r] ? vi[base + j] : 0u;
    }
    __syncthreads();  // wcnt zeroed
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        dk[r] = (u32)(key[r] >> shift) & 255u;
        u64 peers = match_key8(vv[r], dk[r]);
        rk[r] = (unsigned short)__popcll(peers & lanemask_lt());
        if (vv[r] && rk[r] == 0) wcnt[r][w][dk[r]] = (unsigned short)__popcll(peers);
    }
    __syncthreads();
    {
        u32 run = my_start;
        #pragma unroll
        for (int r = 0; r < RS_ROUNDS; r++)
            #pragma unroll
            for (int w2 = 0; w2 < NWAVES; w2++) {
                u32 c = wcnt[r][w2][dg];
                wcnt[r][w2][dg] = (unsigned short)run;
                run += c;
            }
    }
    __syncthreads();
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) rk[r] = (unsigned short)(wcnt[r][w][dk[r]] + rk[r]);
    __syncthreads();  // every count read: sv takes their memory
    u32 *sv = sw.sv;
    tstart[dg] = my_start;
    gbase[dg] = my_base;
    #pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        if (vv[r]) {
            sk[rk[r]] = key[r];
            sv[rk[r]] = val[r];
        }
    }
    __syncthreads();
    for (u32 j = threadIdx.x; j < cnt; j += BLOCK) {
        u64 k = sk[j];
        u32 d = (u32)(k >> shift) & 255u;
        u64 g = gbase[d] + (j - tstart[d]);
        ko[g] = k;
        vo[g] = sv[j];
    }
}

// Group sort of (k, v)[0..A) with G groups whose ids are gid[i] & gmask
// (dense, non-decreasing, 0 <= id < G): each group sorted stably by the key
// bits [0, lob).  NT and LI: the tiles (RS_TILE items) and the items of the
// groups of more than GS_LOCAL items, as the relabel that made the groups
// counted them.  Returns true if the result is in (kt, vt).
inline bool group_sort_kv(Dev &d, u64 A, u32 G, const u32 *gid, u32 gmask, u64 *k, u32 *v,
                          u64 *kt, u32 *vt, int lob, u32 NT, u32 LI) {
    if (A == 0 || G == 0) return false;
    ArenaScope mk(*d.a);
    u32 *start = d.a->get<u32>((u64)G + 1);
    const u64 AA = A;
    const u32 GG = G;
    for_each(d, AA, [=] __device__(u64 i) {
        const u32 g = gid[i] & gmask;
        if (i == 0 || (gid[i - 1] & gmask) != g) start[g] = (u32)i;
        if (i + 1 == AA) start[GG] = (u32)AA;
    });
    u32 *tb = nullptr, *bias = nullptr;
    if (NT) {
        // tiles of the large groups, and each large group's shift from its
        // place among the large groups' items to its place in the array
        tb = d.a->get<u32>((u64)G + 1);
        bias = d.a->get<u32>((u64)G + 1);
        u32 *tbl = tb, *bs = bias;
        scan_exclusive(
            d, GG,
            [=] __device__(u64 g) {
                const u32 s = start[g + 1] - start[g];
                return s > GS_LOCAL ? (u32)((s + RS_TILE - 1) / RS_TILE) : 0u;
            },
            [=] __device__(u64 g, u32 ex, u32 v) {
                tbl[g] = ex;
                if (g + 1 == GG) tbl[GG] = ex + v;
            },
            nullptr);
        scan_exclusive(
            d, GG,
            [=] __device__(u64 g) {
                const u32 s = start[g + 1] - start[g];
                return s > GS_LOCAL ? s : 0u;
            },
            [=] __device__(u64 g, u32 ex, u32) { bs[g] = start[g] - ex; }, nullptr);
    }
    const u32 hc[2] = {NT, LI};
    const int passes = NT ? (lob + 7) / 8 : 0;
    const bool swapped = (passes & 1) != 0;
    const u64 lomask = lob >= 64 ? ~0ull : ((1ull << lob) - 1);
    {
        int pr = d.pbegin("k_gs_local", 32.0 * (double)(A - hc[1]));
        k_gs_local<<<ceil_div(A, GS_WIN), BLOCK, 0, d.s>>>(k, v, swapped ? kt : k, swapped ? vt : v,
                                                          A, gid, gmask, start, lomask);
        PSK_LAUNCH_CHECK();
        d.pend(pr);
    }
    if (NT) {
        u32 *tg = d.a->get<u32>(NT);
        for_each(d, NT, [=] __device__(u64 t) {
            // last group whose first tile is at or before t
            u32 lo = 0, hi = GG;
            while (hi - lo > 1) {
                const u32 mid = (lo + hi) >> 1;
                if (tb[mid] <= (u32)t) lo = mid;
                else hi = mid;
            }
            tg[t] = lo;
        });
        const GroupTiles tiles{start, tb, tg, bias};
        const u64 total = (u64)256 * NT;
        u32 *hist = d.a->get<u32>(total + 1);
        const double items = (double)hc[1];
        bool sw = false;
        for (int p = 0; p < passes; p++) {
            const int shift = 8 * p;
            const u64 *ki = sw ? kt : k;
            const u32 *vi = sw ? vt : v;
            u64 *ko = sw ? k : kt;
            u32 *vo = sw ? v : vt;
            int pc = d.pbegin("k_gs_count", 8.0 * items);
            k_gs_count<<<NT, BLOCK, 0, d.s>>>(ki, shift, hist, NT, tiles);
            PSK_LAUNCH_CHECK();
            d.pend(pc);
            scan_exclusive(
                d, total, [=] __device__(u64 i) { return hist[i]; },
                [=] __device__(u64 i, u32 ex, u32) { hist[i] = ex; }, hist + total);
            int ps = d.pbegin("k_gs_scatter", 24.0 * items);
            k_gs_scatter<<<NT, BLOCK, 0, d.s>>>(ki, vi, ko, vo, shift, hist, NT, tiles);
            PSK_LAUNCH_CHECK();
            d.pend(ps);
            sw = !sw;
        }
    }
    return swapped;
}

// Trace line of a refinement round: items and groups before and after, and
// the items left tied by the size of their group (log2 classes).  gid: group
// id of every kept item (masked by gmask), contiguous groups.
inline void trace_round(Dev &d, const char *what, int round, u64 items, int bits, const u32 *gid,
                        u64 kept, u32 gmask, u64 groups) {
    if (!d.trace) return;
    constexpr int NC = 24;
    u64 cls[NC] = {0};
    if (kept && groups) {
        ArenaScope mk(*d.a);
        u32 *h = d.a->get<u32>(NC);
        u32 *st = d.a->get<u32>(groups), *en = d.a->get<u32>(groups);
        d.zero(h, NC * sizeof(u32));
        const u64 kk = kept, gg = groups;
        // first and last item of every group (dense ids), then the sizes
        for_each(d, kk, [=] __device__(u64 i) {
            const u32 g = gid[i] & gmask;
            if (g >= gg) return;
            if (!i || (gid[i - 1] & gmask) != g) st[g] = (u32)i;
            if (i + 1 == kk || (gid[i + 1] & gmask) != g) en[g] = (u32)i;
        });
        // one atomic per class and wave (same-address atomics serialise)
        k_for_each_waves<<<ceil_div(gg, BLOCK), BLOCK, 0, d.s>>>(
            [=] __device__(u64 g, bool live) {
                const u64 sz = live ? (u64)en[g] - st[g] + 1 : 0;
                int c = 0;
                while ((2ull << c) <= sz && c < NC - 1) c++;
                for (int k = 0; k < NC; k++) {
                    const u64 sum = wave_inclusive_sum<u64>(live && c == k ? sz : 0);
                    if (lane_id() == WAVE - 1 && sum) atomicAdd(&h[k], (u32)sum);
                }
            },
            gg);
        PSK_LAUNCH_CHECK();
        u32 hh[NC];
        d.read(hh, h, sizeof(hh));
        for (int c = 0; c < NC; c++) cls[c] = hh[c];
        mk.reset();
    }
    fprintf(stderr, "[psk] %s round %d: %llu items, %d key bits (%d passes) -> %llu tied in %llu groups;",
            what, round, (unsigned long long)items, bits, (bits + 7) / 8, (unsigned long long)kept,
            (unsigned long long)groups);
    for (int c = 0; c < NC; c++)
        if (cls[c]) fprintf(stderr, " <%llu:%llu", 2ull << c, (unsigned long long)cls[c]);
    fprintf(stderr, "\n");
}

}  // namespace psk
