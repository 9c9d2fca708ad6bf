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
//
// Windows.  Inducing x-1 from x needs T[x-1]; the suffix induced next from
// x-1 needs T[x-2], and so on down the text, a random gather each.  Every SA
// slot therefore carries a window (PW, one u64 per slot) of the bytes before
// its entry: T[x-1], T[x-2], ... (up to 7).  A placed LMS suffix gets its
// window from one gather; an induced entry x-1 inherits the window of x
// shifted by one byte.  The chain of inductions from an LMS position down to
// the previous one spans one LMS substring, so a window covers it whenever
// that substring is at most 7 bytes long (97% of them for DNA); a window that
// has run out is refilled from the text when it is needed.
#pragma once

#include "psk_types.hpp"

namespace psk {

struct Buckets8 {
    u32 cnt[256], nL[256], nLMS[256];
    u64 head[257];
};

// Window of the bytes before an entry x: bits 0..55 hold T[x-1], T[x-2], ...
// (byte i at bits 8(i-1)), bits 56..58 the number v of valid bytes, bit 59
// is set when position x-v has no predecessor (the text, or in a batch the
// entry's own text, starts there).  v = 0 without bit 59: not known yet.
struct Win {
    static constexpr u64 BYTES = (1ull << 56) - 1;
    __device__ __forceinline__ static u32 v(u64 w) { return (u32)(w >> 56) & 7u; }
    __device__ __forceinline__ static bool e(u64 w) { return (w >> 59) & 1ull; }
    __device__ __forceinline__ static bool known(u64 w) { return (w >> 56) != 0; }
    // byte i >= 1 (T[x-i]); i <= v
    __device__ __forceinline__ static u32 b(u64 w, u32 i) { return (u32)(w >> (8 * (i - 1))) & 255u; }
    __device__ __forceinline__ static u64 make(u64 bytes, u32 v, bool e) {
        return (bytes & BYTES) | ((u64)v << 56) | ((u64)e << 59);
    }
    // window of x - t from the window of x (t <= v)
    __device__ __forceinline__ static u64 shift(u64 w, u32 t) {
        return make((w & BYTES) >> (8 * t), v(w) - t, e(w));
    }
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
    // Window of x gathered from the text: up to 7 bytes, stopping where a
    // position has no predecessor.  With ALIGNED (T 8-byte aligned) the bytes
    // come from the (at most two) aligned u64 words that hold them.
    template <bool ALIGNED>
    __device__ __forceinline__ u64 window(u32 x) const {
        u32 v = x < 7u ? x : 7u;  // bytes before x in the array
        bool e = x <= 7u;
        if (SB && v) {
            // a text start s in [x-v+1, x] leaves the bytes T[s..x-1]
            u32 lo = x - v + 1;
            u32 wa = lo >> 5, wb = x >> 5;
            u32 mb = SB[wb] & (0xFFFFFFFFu >> (31 - (x & 31)));
            if (wb == wa) mb &= 0xFFFFFFFFu << (lo & 31);
            u32 s = 0;
            bool found = false;
            if (mb) {
                s = wb * 32 + (31 - __clz(mb));
                found = true;
            } else if (wa != wb) {
                u32 ma = SB[wa] & (0xFFFFFFFFu << (lo & 31));
                if (ma) {
                    s = wa * 32 + (31 - __clz(ma));
                    found = true;
                }
            }
            if (found) {
                v = x - s;
                e = true;
            }
        }
        u64 bytes = 0;
        if (v) {
            if (ALIGNED) {
                const u64 *Q = (const u64 *)T;
                u32 lo = x - v;
                u32 wa = lo >> 3, wb = (x - 1) >> 3;
                u64 q0 = Q[wa];
                u64 q1 = wb != wa ? Q[wb] : 0ull;
                #pragma unroll
                for (u32 i = 1; i <= 7; i++) {
                    if (i <= v) {
                        u32 pos = x - i;
                        u64 q = (pos >> 3) == wa ? q0 : q1;
                        bytes |= ((q >> (8 * (pos & 7))) & 255ull) << (8 * (i - 1));
                    }
                }
            } else {
                #pragma unroll
                for (u32 i = 1; i <= 7; i++)
                    if (i <= v) bytes |= (u64)T[x - i] << (8 * (i - 1));
            }
        }
        return Win::make(bytes, v, e);
    }
    __device__ __forceinline__ u64 window_any(u32 x, bool aligned) const {
        return aligned ? window<true>(x) : window<false>(x);
    }
};

// Window of the entry x - t (1 <= t <= r) of a chain: seed x whose window is
// w, followed by its run of r bytes c; sb = T[x-r-1] when x - r has a
// predecessor (has_sb).
__device__ __forceinline__ u64 chain_window(u64 w, u32 r, u32 c, bool has_sb, u32 sb, u32 t) {
    const u32 v = Win::v(w);
    if (r < v) return Win::shift(w, t);
    // the run covers the window: T[x-t-1 .. x-r] = c, then sb or the text start
    const u32 k = r - t;
    const u32 vv = k < 7u ? k : 7u;
    u64 bytes = (u64)c * 0x0101010101010101ull;
    bytes &= vv ? (~0ull >> (64 - 8 * vv)) : 0ull;
    if (vv < 7u) {
        if (has_sb) return Win::make(bytes | ((u64)sb << (8 * vv)), vv + 1, false);
        return Win::make(bytes, vv, true);
    }
    return Win::make(bytes, 7, false);
}

// Fill one bucket's queue: positions qp(0..limit); generation 0 (the entries
// induced from other buckets) is [0, *d_start).  Every further entry is a
// chain: seed x_k followed by x_k-1, x_k-2, ... while the predecessor byte is
// c, so seed k contributes r_k entries, its run of c's before it.  The queue
// is breadth-first: all depth-1 entries in seed order, then depth 2, ...;
// the entry (k, depth t) sits at a + sum_{t'<t} #{r >= t'} + #{k' < k :
// r_k' >= t}, i.e. the entries emitted seed by seed, stably sorted by depth.
// One pass measures the runs, one scan gives emission offsets, and one stable
// multi-split by depth places them (two passes if a run exceeds 256).
// Leaves every queue entry's window known (refilled where it had run out)
// and its predecessor byte in dig[0..limit), or `none` if it has no
// predecessor: 0 for the L-scan, 255 for the S-scan, a byte that the scan
// never induces from (0 is never > c, 255 never < c).  d_tmp: 2 u32.
// A seed with a chain (r > 0): its position, run and window.
struct ChainSeed {
    u32 x, r;
    u64 w;
};

// The run of c's before seed x (window w, refilled if it had run out) and
// the byte before the run: sb | 256 if x - r has a predecessor, else 0.
__device__ __forceinline__ u32 seed_run(const Pred &pr, u32 x, u64 w, u32 c, u32 *sb) {
    const u32 v = Win::v(w);
    *sb = 0;
    if (!v || Win::b(w, 1) != c) return 0;
    // inside the window, or beyond it from the decided bits
    u32 r = 1;
    while (r < v && Win::b(w, r + 1) == c) r++;
    if (r < v) {
        *sb = Win::b(w, r + 1) | 256u;
    } else if (!Win::e(w)) {
        r = pr.run_before(x);
        if (pr.has(x - r)) *sb = pr.byte(x - r) | 256u;
    }
    return r;
}

constexpr int CQ_ROUNDS = 16;
constexpr u64 CQ_TILE = (u64)BLOCK * CQ_ROUNDS;  // queue seeds per tile
// Chain entries of depth <= CQ_DIRECT go straight to their queue slots; the
// rest (runs longer than CQ_DIRECT: repeats, low-complexity stretches) are
// emitted and split by depth.  Per-seed ranks of the CQ_DIRECT depths share
// one 64-bit block scan: 9-bit fields (<= 256 seeds per scan step).
constexpr u32 CQ_DIRECT = 7;

// Runs of one tile of seeds [t CQ_TILE, ...) of a queue with a seeds: every
// seed's predecessor byte goes to dig (`none` without one), and the seeds
// with a chain are compacted, in seed order, to the start of the tile's own
// range of cs; tact[t] counts them.  The tile's entry counts go to the
// count matrix cm (nt tiles per row): row d-1 (d <= CQ_DIRECT) holds the
// seeds with r >= d, row CQ_DIRECT the entries deeper than CQ_DIRECT.
// da (optional): the seed count is *da (device-sized launch: the workgroups
// stride over its tiles), and block 0 leaves the count matrix's size in
// *msize.
__global__ void __launch_bounds__(BLOCK) k_chain_runs(Pred pr, bool aligned, const u32 *SA, u64 *PW,
                                                      u8 *dig, QPos qp, u32 c, u64 a, u32 none,
                                                      ChainSeed *cs, u32 *tact, u32 *cm, u32 nt,
                                                      u32 *maxr, const u32 *da = nullptr,
                                                      u32 *msize = nullptr) {
    __shared__ u32 red[NWAVES + 1];
    __shared__ u32 sums[CQ_DIRECT + 1];
    if (da) {
        a = *da;
        nt = (u32)((a + CQ_TILE - 1) / CQ_TILE);
        if (msize && blockIdx.x == 0 && threadIdx.x == 0) *msize = (CQ_DIRECT + 1) * nt;
    }
    for (u32 t = blockIdx.x; t < nt; t += gridDim.x) {
    if (t != blockIdx.x) __syncthreads();  // the previous tile's sums are written
    if (threadIdx.x <= CQ_DIRECT) sums[threadIdx.x] = 0;
    const u64 base = (u64)t * CQ_TILE;
    u32 nact = 0, mx = 0, deep = 0;
    u32 ge[CQ_DIRECT];
    #pragma unroll
    for (u32 q = 0; q < CQ_DIRECT; q++) ge[q] = 0;
    for (int rr = 0; rr < CQ_ROUNDS; rr++) {
        const u64 k = base + (u64)rr * BLOCK + threadIdx.x;
        u32 r = 0, x = 0;
        u64 w = 0;
        if (k < a) {
            const u64 slot = qp(k);
            x = SA[slot];
            w = PW[slot];
            if (!Win::known(w)) {
                w = pr.window_any(x, aligned);
                PW[slot] = w;
            }
            dig[k] = (u8)(Win::v(w) ? Win::b(w, 1) : none);
            u32 sb;
            r = seed_run(pr, x, w, c, &sb);
        }
        u32 tot;
        const u32 ex = block_exclusive_sum(r ? 1u : 0u, red, &tot);  // includes barriers
        if (r) cs[base + nact + ex] = ChainSeed{x, r, w};
        nact += tot;
        #pragma unroll
        for (u32 q = 0; q < CQ_DIRECT; q++) ge[q] += r > q ? 1u : 0u;
        deep += r > CQ_DIRECT ? r - CQ_DIRECT : 0u;
        mx = r > mx ? r : mx;
    }
    #pragma unroll
    for (u32 q = 0; q <= CQ_DIRECT; q++) {
        const u32 v = wave_inclusive_sum(q < CQ_DIRECT ? ge[q] : deep);
        if (lane_id() == WAVE - 1 && v) atomicAdd(&sums[q], v);
    }
    __syncthreads();
    if (threadIdx.x <= CQ_DIRECT) cm[(u64)threadIdx.x * nt + t] = sums[threadIdx.x];
    if (threadIdx.x == 0) tact[t] = nact;
    // same-address atomics serialise in L2: only raise the maximum when this
    // wave beats the value already stored
    const u32 wm = wave_max(mx);
    if (lane_id() == 0 && wm > __atomic_load_n(maxr, __ATOMIC_RELAXED)) atomicMax(maxr, wm);
    }
}

// Chain entries of one tile's compacted seeds.  co is the exclusive scan of
// the count matrix: co[(d-1) nt + t] is the queue offset (past the a seeds)
// of tile t's first depth-d entry, co[CQ_DIRECT nt + t] that of its first
// deep entry, and deep entries start at d0.  An entry of depth <= CQ_DIRECT
// is written to its slot (SA, PW, dig); a deeper one is emitted as (depth -
// CQ_DIRECT - 1, entry, predecessor byte, window) at its index past d0, in
// seed order, for the split by depth.
// da (optional): device-sized launch as in k_chain_runs, with d0 = hv[2]
// and the total of the count matrix in hv[1] (checked against the queue's
// chain entries); block 0 leaves the number of deep entries in hv[3].
__global__ void __launch_bounds__(BLOCK) k_chain_emit(Pred pr, const ChainSeed *cs, const u32 *tact,
                                                      const u32 *co, u32 nt, u64 d0, u32 c,
                                                      u32 none, QPos qp, u64 a, u64 limit, u32 *SA,
                                                      u64 *PW, u8 *dig, u32 *ek, u32 *ev, u8 *eb,
                                                      u64 *ew, u32 *err, const u32 *da = nullptr,
                                                      u32 *hv = nullptr) {
    __shared__ u64 red64[NWAVES + 1];
    __shared__ u32 red[NWAVES + 1];
    if (da) {
        a = *da;
        nt = (u32)((a + CQ_TILE - 1) / CQ_TILE);
        d0 = hv[2];
        const bool bad = a > limit || (u64)hv[1] != limit - a || d0 > limit - a;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            hv[3] = bad ? 0u : (u32)(limit - a - d0);
            if (bad) atomicOr(err, 2u);  // chain entries do not fill their bucket part
        }
        if (bad) return;
    }
    for (u32 t = blockIdx.x; t < nt; t += gridDim.x) {
    if (t != blockIdx.x) __syncthreads();
    const u64 base = (u64)t * CQ_TILE;
    const u32 na = tact[t];
    u64 off[CQ_DIRECT];
    #pragma unroll
    for (u32 q = 0; q < CQ_DIRECT; q++) off[q] = co[(u64)q * nt + t];
    u64 doff = co[(u64)CQ_DIRECT * nt + t];
    const u64 nchain = limit - a;
    for (u32 b0 = 0; b0 < na; b0 += BLOCK) {
        const u32 i = b0 + threadIdx.x;
        ChainSeed sd{0, 0, 0};
        if (i < na) sd = cs[base + i];
        const u32 r = sd.r;
        u64 flags = 0;
        #pragma unroll
        for (u32 q = 0; q < CQ_DIRECT; q++) flags |= (u64)(r > q ? 1u : 0u) << (9 * q);
        u64 tot;
        const u64 ex = block_exclusive_sum(flags, red64, &tot);  // includes barriers
        const u32 dr = r > CQ_DIRECT ? r - CQ_DIRECT : 0u;
        u32 dtot;
        const u32 dex = block_exclusive_sum(dr, red, &dtot);
        if (r) {
            u32 sb;
            seed_run(pr, sd.x, sd.w, c, &sb);
            const bool hs = sb >> 8;
            const u32 last = hs ? sb & 255u : none;
            const u32 rd = r < CQ_DIRECT ? r : CQ_DIRECT;
            for (u32 q = 0; q < rd; q++) {
                const u64 pos = off[q] + ((ex >> (9 * q)) & 511u);
                if (pos >= d0 || a + pos >= limit) {
                    atomicOr(err, 2u);
                    continue;
                }
                const u64 slot = qp(a + pos);
                SA[slot] = sd.x - (q + 1);
                PW[slot] = chain_window(sd.w, r, c, hs, sb & 255u, q + 1);
                dig[a + pos] = (u8)(q + 1 < r ? c : last);
            }
            if (dr) {
                const u64 o = doff - d0 + dex;
                if (doff < d0 || o + dr > nchain - d0) {
                    atomicOr(err, 2u);
                } else {
                    for (u32 u = 0; u < dr; u++) {
                        const u32 dep = CQ_DIRECT + 1 + u;
                        ek[o + u] = u;
                        ev[o + u] = sd.x - dep;
                        eb[o + u] = dep < r ? (u8)c : (u8)last;
                        ew[o + u] = chain_window(sd.w, r, c, hs, sb & 255u, dep);
                    }
                }
            }
        }
        #pragma unroll
        for (u32 q = 0; q < CQ_DIRECT; q++) off[q] += (tot >> (9 * q)) & 511u;
        doff += dtot;
    }
    }
}

inline void chain_queue8(Dev &d, Pred pr, bool aligned, u32 *SA, u64 *PW, u8 *dig, QPos qp,
                         u32 c, const u32 *d_start, u64 limit, u32 *d_tmp, u32 none) {
    const u64 a = d.read_u32(d_start);
    u32 *err = d.err;
    if (a > limit) fail(PSK_EDEVICE, "queue seed count exceeds its bucket part");
    if (a == 0) {
        if (limit) fail(PSK_EDEVICE, "chain entries do not fill their bucket part");
        return;
    }
    (void)d_tmp;
    ArenaScope mk(*d.a);
    const u32 nt = ceil_div(a, CQ_TILE);
    ChainSeed *cs = d.a->get<ChainSeed>(a);
    const u64 rows = (u64)(CQ_DIRECT + 1) * nt;
    u32 *tact = d.a->get<u32>(nt), *cm = d.a->get<u32>(rows), *co = d.a->get<u32>(rows);
    u32 *hc = d.a->get<u32>(4);  // largest run, entries in all, first deep entry
    d.zero(hc, 4 * sizeof(u32));
    const u64 nchain = limit - a;
    {
        int prec = d.pbegin("chain8_runs", 13.0 * a);
        k_chain_runs<<<nt, BLOCK, 0, d.s>>>(pr, aligned, SA, PW, dig, qp, c, a, none, cs, tact, cm,
                                            nt, hc);
        PSK_LAUNCH_CHECK();
        d.pend(prec);
    }
    if (nchain == 0) return;
    // queue offsets of every (depth, tile) and of the deep entries
    const u64 drow = (u64)CQ_DIRECT * nt;
    scan_exclusive(
        d, rows, [=] __device__(u64 i) { return cm[i]; },
        [=] __device__(u64 i, u32 ex, u32) {
            co[i] = ex;
            if (i == drow) hc[2] = ex;
        },
        hc + 1);
    u32 hv[3];
    d.read(hv, hc, sizeof(hv));
    const u32 R = hv[0];
    const u64 d0 = hv[2];
    if ((u64)hv[1] != nchain || d0 > nchain)
        fail(PSK_EDEVICE, "chain entries do not fill their bucket part");
    const u64 ndeep = nchain - d0;
    u32 *ek = nullptr, *ev = nullptr;
    u8 *eb = nullptr;
    u64 *ew = nullptr;
    if (ndeep) {
        ek = d.a->get<u32>(ndeep);
        ev = d.a->get<u32>(ndeep);
        eb = d.a->get<u8>(ndeep);
        ew = d.a->get<u64>(ndeep);
    }
    {
        int pe = d.pbegin("chain8_emit", 16.0 * a + 13.0 * (double)nchain);
        k_chain_emit<<<nt, BLOCK, 0, d.s>>>(pr, cs, tact, co, nt, d0, c, none, qp, a, limit, SA, PW,
                                            dig, ek, ev, eb, ew, err);
        PSK_LAUNCH_CHECK();
        d.pend(pe);
    }
    if (!ndeep) return;
    // deep entries in breadth-first order: a stable split by depth
    const u64 lim = limit, base = a + d0;
    if (R - CQ_DIRECT <= 256) {
        multisplit(
            d, ndeep,
            [=] __device__(u64 j, u32 *key) {
                *key = ek[j];
                return true;
            },
            [=] __device__(u64 j, u32, u32 gpos, u32) {
                u64 q = base + gpos;
                if (q < lim) {
                    const u64 slot = qp(q);
                    SA[slot] = ev[j];
                    PW[slot] = ew[j];
                    dig[q] = eb[j];
                } else {
                    atomicOr(err, 2u);
                }
            },
            nullptr);
    } else {
        // long runs: stable LSD sort of (depth, emission index)
        u64 *k0 = d.a->get<u64>(ndeep), *k1 = d.a->get<u64>(ndeep);
        u32 *v0 = d.a->get<u32>(ndeep), *v1 = d.a->get<u32>(ndeep);
        for_each(d, ndeep, [=] __device__(u64 j) {
            k0[j] = ek[j];
            v0[j] = (u32)j;
        });
        bool sw = radix_sort_kv(d, ndeep, k0, v0, k1, v1, bit_width(R - CQ_DIRECT - 1));
        const u32 *V = sw ? v1 : v0;
        for_each(d, ndeep, [=] __device__(u64 i) {
            u32 j = V[i];
            const u64 slot = qp(base + i);
            SA[slot] = ev[j];
            PW[slot] = ew[j];
            dig[base + i] = eb[j];
        });
    }
}

// chain_queue8 without host synchronisation, for texts whose runs of equal
// bytes are all at most 256 long (every deep entry's depth fits one 8-bit
// multi-split digit): the seed count a = *d_start, the count matrix's size,
// the deep entries' start and count stay on the device, and the launches are
// device-sized from the bucket part's size (limit >= a).  The queue's fill
// is checked on the device (err bit 2).
inline void chain_queue8_dn(Dev &d, Pred pr, bool aligned, u32 *SA, u64 *PW, u8 *dig, QPos qp,
                            u32 c, const u32 *d_start, u64 limit, u32 none) {
    if (limit == 0) return;
    u32 *err = d.err;
    ArenaScope mk(*d.a);
    const u64 ntb = ceil_div(limit, CQ_TILE);
    ChainSeed *cs = d.a->get<ChainSeed>(limit);
    const u64 rows = (u64)(CQ_DIRECT + 1) * ntb;
    u32 *tact = d.a->get<u32>(ntb), *cm = d.a->get<u32>(rows), *co = d.a->get<u32>(rows);
    // largest run, entries in all, first deep entry, deep entries, matrix size
    u32 *hv = d.a->get<u32>(8);
    d.zero(hv, 8 * sizeof(u32));
    u32 *msize = hv + 4;
    // seeds are most of a large bucket part: one workgroup per tile of the
    // bound (the gathers of a tile are latency-bound)
    const u32 g = dn_grid(ntb, ntb);
    {
        int prec = d.pbegin("chain8_runs", 13.0 * limit);
        k_chain_runs<<<g, BLOCK, 0, d.s>>>(pr, aligned, SA, PW, dig, qp, c, 0, none, cs, tact, cm,
                                           0, hv, d_start, msize);
        PSK_LAUNCH_CHECK();
        d.pend(prec);
    }
    // queue offsets of every (depth, tile) and of the deep entries: the deep
    // row starts at CQ_DIRECT / (CQ_DIRECT + 1) of the matrix
    u32 *h2 = hv + 2;
    scan_exclusive_dn(
        d, rows, msize, [=] __device__(u64 i) { return cm[i]; },
        [=] __device__(u64 i, u32 ex, u32) {
            co[i] = ex;
            if (i == (u64)*msize / (CQ_DIRECT + 1) * CQ_DIRECT) *h2 = ex;
        },
        hv + 1);
    u32 *ek = d.a->get<u32>(limit), *ev = d.a->get<u32>(limit);
    u8 *eb = d.a->get<u8>(limit);
    u64 *ew = d.a->get<u64>(limit);
    {
        int pe = d.pbegin("chain8_emit", 29.0 * limit);
        k_chain_emit<<<g, BLOCK, 0, d.s>>>(pr, cs, tact, co, 0, 0, c, none, qp, 0, limit, SA, PW,
                                           dig, ek, ev, eb, ew, err, d_start, hv);
        PSK_LAUNCH_CHECK();
        d.pend(pe);
    }
    // deep entries in breadth-first order: a stable split by depth, after the
    // seeds and the entries of depth <= CQ_DIRECT
    const u64 lim = limit;
    const u32 *dd0 = hv + 2;
    multisplit_dn(
        d, limit, hv + 3,
        [=] __device__(u64 j, u32 *key) {
            *key = ek[j];
            return true;
        },
        [=] __device__(u64 j, u32, u32 gpos, u32) {
            const u64 q = (u64)*d_start + *dd0 + gpos;
            if (q < lim) {
                const u64 slot = qp(q);
                SA[slot] = ev[j];
                PW[slot] = ew[j];
                dig[q] = eb[j];
            } else {
                atomicOr(err, 2u);
            }
        });
    mk.reset();
}

// The same queue fill without host synchronisation, for texts whose runs of
// equal bytes are all at most 256 long (so every depth fits one 8-bit
// multi-split digit): the seed count a = *d_start stays on the device, and
// every launch takes the host-known maximum shape (limit items; the items
// past a, or past the limit - a chain entries, take no part).  d_tmp: 2 u32.
inline void chain_queue8_dev(Dev &d, Pred pr, bool aligned, u32 *SA, u64 *PW, u8 *dig, QPos qp,
                             u32 c, const u32 *d_start, u64 limit, u32 *d_tmp, u32 none) {
    if (limit == 0) return;
    u32 *err = d.err;
    ArenaScope mk(*d.a);
    u32 *run = d.a->get<u32>(limit + 1);
    unsigned short *stop = d.a->get<unsigned short>(limit + 1);  // byte | has << 8
    const u64 lim = limit;
    {
        int prec = d.pbegin("chain8_runs", 22.0 * limit);
        for_each(d, lim, [=] __device__(u64 k) {
            const u64 a = *d_start;
            if (k >= a) return;
            u32 r = 0;
            const u64 slot = qp(k);
            u32 x = SA[slot];
            u64 w = PW[slot];
            if (!Win::known(w)) {
                w = pr.window_any(x, aligned);
                PW[slot] = w;
            }
            const u32 v = Win::v(w);
            const u32 cp = v ? Win::b(w, 1) : none;
            dig[k] = (u8)cp;
            u32 sb = 0;
            if (v && cp == c) {
                r = 1;
                while (r < v && Win::b(w, r + 1) == c) r++;
                if (r < v) {
                    sb = Win::b(w, r + 1) | 256u;
                } else if (!Win::e(w)) {
                    r = pr.run_before(x);
                    if (pr.has(x - r)) sb = pr.byte(x - r) | 256u;
                }
            }
            run[k] = r;
            stop[k] = (unsigned short)sb;
        });
        d.pend(prec);
    }
    // emission offsets of the seeds' chains; their total must fill the
    // queue: a + total == limit
    u32 *eo = d.a->get<u32>(limit + 1);
    u32 *tot = d_tmp + 1;
    scan_exclusive(
        d, lim, [=] __device__(u64 k) { return k < *d_start ? run[k] : 0u; },
        [=] __device__(u64 k, u32 ex, u32) { eo[k] = ex; }, tot);
    u32 *ek = d.a->get<u32>(limit), *ev = d.a->get<u32>(limit);
    u8 *eb = d.a->get<u8>(limit);
    u64 *ew = d.a->get<u64>(limit);
    for_each(d, lim, [=] __device__(u64 k) {
        const u64 a = *d_start;
        if (k == 0 && (a > lim || (u64)*tot + a != lim)) atomicOr(err, 1024u);
        if (k >= a) return;
        u32 r = run[k];
        if (!r) return;
        const u64 slot = qp(k);
        u32 x = SA[slot], o = eo[k];
        if ((u64)o + r > lim) return;  // flagged above
        u64 w = PW[slot];
        const u32 st = stop[k];
        const bool hs = st >> 8;
        const u32 sb = st & 255u;
        for (u32 t = 1; t <= r; t++) {
            ek[o + t - 1] = t - 1;
            ev[o + t - 1] = x - t;
            eb[o + t - 1] = t < r ? (u8)c : (u8)(hs ? sb : none);
            ew[o + t - 1] = chain_window(w, r, c, hs, sb, t);
        }
    });
    // breadth-first order: a stable multi-split of the emitted entries by depth
    multisplit(
        d, lim,
        [=] __device__(u64 j, u32 *key) {
            if (j >= lim - *d_start) return false;
            *key = ek[j];
            return true;
        },
        [=] __device__(u64 j, u32, u32 gpos, u32) {
            const u64 q = *d_start + gpos;
            if (q < lim) {
                const u64 slot = qp(q);
                SA[slot] = ev[j];
                PW[slot] = ew[j];
                dig[q] = eb[j];
            } else {
                atomicOr(err, 2u);
            }
        },
        nullptr);
}

// Longest run of equal bytes in the text is at most 256?  From the decided
// bits DB (bit p: T[p] != T[p+1] or p ends a text): a run ends at each
// decided position, and starts after the previous one.
inline bool runs_at_most_256(Dev &d, const u32 *DB, u64 n, u32 *d_flag) {
    const u64 nwords = (n + 31) / 32;
    d.zero(d_flag, sizeof(u32));
    for_each(d, nwords, [=] __device__(u64 w) {
        const u32 x = DB[w];
        if (!x) return;
        // the run ending at the lowest decided position of this word starts
        // after the previous decided position (or at the text start)
        const u64 p = w * 32 + (u64)(__ffs(x) - 1);
        u64 q = w;
        bool found = false;
        u64 prev = 0;
        for (int back = 0; back < 10 && q > 0; back++) {
            const u32 y = DB[--q];
            if (y) {
                prev = q * 32 + (u64)(31 - __clz(y));
                found = true;
                break;
            }
        }
        if (!found && q > 0) {
            atomicOr(d_flag, 1u);  // more than 9 words without a decided bit
            return;
        }
        const u64 len = found ? p - prev : p + 1;
        if (len > 256) atomicOr(d_flag, 1u);
    });
    return d.read_u32(d_flag) == 0;
}

// Per-bucket device tables: head, tail, fill (induce counters).
struct Induce8Tables {
    u32 *head, *tail, *fill, *counts, *tmp;
};

// Full induce (L-scan then S-scan) over SA for a byte text, from the LMS
// entries already placed at the bucket tails with their windows in PW (one
// u64 per SA slot).  dig: scratch of >= n bytes.
// Batches (SB != nullptr): K texts concatenated, ends[0..K) their last
// positions; the L-scan starts from every text's sentinel, in text order.
// lean: large queues take their exact sizes (readbacks) instead of the
// sync-free fill, whose seed and deep-entry arrays are sized by the bucket
// part (~33 bytes per entry): a server's worker contexts keep their
// workspace to what a job strictly needs.
inline void induce8(Dev &d, const u8 *T, u64 n, u32 *SA, const Buckets8 &B, Induce8Tables tb,
                    u8 *dig, u64 *PW, const u32 *DB, const u32 *SB = nullptr,
                    const u32 *ends = nullptr, u64 K = 1, bool lean = false) {
    const Pred pr{T, SB, DB};
    const bool aligned = (((uintptr_t)T) & 7u) == 0;
    u32 *head = tb.head, *tail = tb.tail, *fill = tb.fill, *counts = tb.counts;
    // Queues fill without host synchronisation when no run of equal bytes
    // exceeds 256 (every chain depth fits one multi-split digit): small ones
    // (large alphabets) from per-seed launches of the bucket part's shape,
    // large ones from tiles of seeds with device-side counts; otherwise with
    // the exact sizes read back
    constexpr u64 SYNC_FREE_MAX = 1ull << 22;
    const bool short_runs = runs_at_most_256(d, DB, n, tb.tmp);
    auto queue = [&](QPos qp, u32 c, const u32 *d_start, u64 limit, u32 none) {
        if (short_runs && limit < SYNC_FREE_MAX)
            chain_queue8_dev(d, pr, aligned, SA, PW, dig, qp, c, d_start, limit, tb.tmp, none);
        else if (short_runs && !lean)
            chain_queue8_dn(d, pr, aligned, SA, PW, dig, qp, c, d_start, limit, none);
        else
            chain_queue8(d, pr, aligned, SA, PW, dig, qp, c, d_start, limit, tb.tmp, none);
    };
    u32 *err = d.err;
    const u64 nn = n;
    // ---------------- L-scan ----------------
    d.zero(fill, 256 * sizeof(u32));
    if (!ends) {
        for_each(d, 1, [=] __device__(u64) {
            u32 c = T[nn - 1];
            SA[head[c]] = (u32)(nn - 1);  // the sentinel induces n-1 first
            PW[head[c]] = pr.window<false>((u32)(nn - 1));
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
                u32 slot = head[key] + (gpos - dstart);
                SA[slot] = ends[i];
                PW[slot] = pr.window<false>(ends[i]);
            },
            counts);
        for_each(d, 256, [=] __device__(u64 k) { fill[k] = counts[k]; });
    }
    for (int c = 0; c < 256; c++) {
        if (!B.cnt[c]) continue;
        const u64 hc = B.head[c], tc = B.head[c + 1], nl = B.nL[c], nlms = B.nLMS[c];
        const u32 cc = (u32)c;
        if (nl) queue(QPos{hc, 1}, cc, fill + c, nl, 0u);
        const u64 R = nl + nlms;
        if (c == 255 || R == 0) continue;  // nothing can be induced above 255
        auto pos = [=] __device__(u64 j) -> u64 { return j < nl ? hc + j : tc - nlms + (j - nl); };
        int ps = multisplit2(
            d, R,
            [=] __device__(u64 j, u32 *key) {
                // predecessor byte: cached for the queue part; from the window
                // for an LMS entry (which always has a predecessor).  No
                // predecessor reads 0, which is never > c.
                u32 cp = j < nl ? (u32)dig[j] : Win::b(PW[pos(j)], 1);
                u32 k = cp > cc ? cp : 0u;
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
                const u64 src = pos(j);
                u32 x = SA[src];
                u64 w = PW[src];
                u64 dst = (u64)head[key] + fill[key] + (gpos - dstart);
                if (dst < nn) {
                    SA[dst] = x - 1;
                    PW[dst] = Win::shift(w, 1);
                } else {
                    atomicOr(err, 4u);
                }
            },
            counts, "induce8", 10.0 * R, 5.0 * R);
        u64 *slot = d.pslot(ps);
        for_each(d, 256, [=] __device__(u64 k) {
            fill[k] += counts[k];
            if (slot && counts[k]) atomicAdd((unsigned long long *)slot, 20ull * counts[k]);
        });
    }
    // ---------------- S-scan ----------------
    d.zero(fill, 256 * sizeof(u32));
    for (int c = 255; c >= 0; c--) {
        if (!B.cnt[c]) continue;
        const u64 hc = B.head[c], tc = B.head[c + 1], nl = B.nL[c], ns = B.cnt[c] - B.nL[c];
        const u32 cc = (u32)c;
        if (ns) queue(QPos{tc - 1, -1}, cc, fill + c, ns, 255u);
        const u64 R = ns + nl;
        if (c == 0 || R == 0) continue;  // nothing can be induced below 0
        auto pos = [=] __device__(u64 j) -> u64 { return j < ns ? tc - 1 - j : hc + nl - 1 - (j - ns); };
        int ps = multisplit2(
            d, R,
            [=] __device__(u64 j, u32 *key) {
                u32 k;
                if (j < ns) {
                    // cached predecessor byte, 255 if none
                    u32 cp = dig[j];
                    k = cp < cc ? cp : 255u;
                } else {
                    // an L entry: its window is known since its bucket's L-scan
                    u64 w = PW[pos(j)];
                    k = (Win::v(w) && Win::b(w, 1) < cc) ? Win::b(w, 1) : 255u;
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
                const u64 src = pos(j);
                u32 x = SA[src];
                u64 w = PW[src];
                i64 dst = (i64)tail[key] - 1 - (i64)(fill[key] + (gpos - dstart));
                if (dst >= 0 && dst < (i64)nn) {
                    SA[dst] = x - 1;
                    PW[dst] = Win::shift(w, 1);
                } else {
                    atomicOr(err, 8u);
                }
            },
            counts, "induce8", 6.0 * R, 5.0 * R);
        u64 *slot = d.pslot(ps);
        for_each(d, 256, [=] __device__(u64 k) {
            fill[k] += counts[k];
            if (slot && counts[k]) atomicAdd((unsigned long long *)slot, 20ull * counts[k]);
        });
    }
}

}  // namespace psk
