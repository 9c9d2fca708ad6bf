// psk_relabel.hpp -- relabel and compaction of one sorted refinement round,
// shared by the chain sort of the integer levels and the direct sort of LMS
// substrings.
//
// For sorted keys K (hi = the item's group, lo = what refines it) and items V:
//   g0(i) = first index of i's group, g1(i) = its last index,
//   s0(i) = first index of i's subgroup (items with an equal key),
//   label(i) = base + (s0 - g0)    (first-slot labels), or
//              base - (g1 - s0)    (last-slot labels),
//   where base is the group's label (hi itself, or glab[hi] for dense ids),
//   keep(i)  = i's subgroup has >= 2 items (and, with a terminated bit, is
//              not terminated) -> compacted, in order, into nxt.
// One reduce pass (per tile: max of group starts, max of subgroup starts,
// number kept, min of group ends, number of kept subgroups) and one apply pass.
#pragma once

#include "psk_naming.hpp"

namespace psk {

struct RelIn {
    const u64 *K;
    u64 A;
    int B;
    bool tbit;  // bit 0 of a key set: the item is finished (never kept)
    u64 kmask = ~0ull;  // the sorted bits of a key (the rest rides along)
    // bit 0 of a key set: the item is alone in its subgroup (finished) even
    // where its neighbours carry the same key -- their order is already
    // final (the sort was stable over input in final order)
    bool solo = false;
    // flags of item i: group start, subgroup start, kept, group end
    __device__ __forceinline__ void flags(u64 i, bool &gs, bool &ss, bool &kp, bool &ge) const {
        bool ns;
        flags(i, gs, ss, kp, ge, ns);
    }
    // and ns: item i + 1 is in i's subgroup
    __device__ __forceinline__ void flags(u64 i, bool &gs, bool &ss, bool &kp, bool &ge,
                                          bool &ns) const {
        u64 k = K[i] & kmask, hi = k >> B;
        u64 kp_ = i ? K[i - 1] & kmask : ~0ull;
        u64 kn_ = i + 1 < A ? K[i + 1] & kmask : ~0ull;
        const bool so = solo && (k & 1u);
        gs = i == 0 || (kp_ >> B) != hi;
        ss = i == 0 || kp_ != k || so;
        ge = i + 1 == A || (kn_ >> B) != hi;
        ns = i + 1 < A && kn_ == k && !so;
        kp = (!ss || ns) && !(tbit && (k & 1u)) && !so;
    }
};

// Run offsets.  In the final chain sorts of a level a position's label is
// read by a later round only through an item of the same type run (chains
// stay inside their run up to the terminal), at offset h' >= 2h from it, and
// after the L sort only at the first L position of a run (the S chains'
// terminals).  An item's run offset (same-type positions right before it,
// capped at RUN_CAP) rides in the top bits of its key, above the sorted
// bits, and of its group id between rounds.
constexpr int RUN_KEY_SHIFT = 59;
constexpr int RUN_GID_SHIFT = 27;
constexpr u32 RUN_GID_MASK = (1u << RUN_GID_SHIFT) - 1u;
constexpr u32 RUN_CAP = 31;

// Run offset of position p in the type bits sb (bit set = S-type), for an
// S item (s) or an L item; positions before 0 end a run.
__device__ __forceinline__ u32 run_offset(const u32 *sb, u32 p, bool s) {
    if (p == 0) return 0;
    const u32 e = p - 1, we = e >> 5, be = e & 31u;
    const u32 w0 = we ? sb[we - 1] : (s ? 0u : 0xFFFFFFFFu);
    const u64 c = ((u64)sb[we] << 32) | w0;
    const u32 y = (u32)(c >> (be + 1));  // bit 31 = position p-1, bit 0 = p-32
    const u32 off = (u32)__clz(s ? ~y : y);
    return off < RUN_CAP ? off : RUN_CAP;
}

// What a round does with an item's new label.  ChainSink: positions' labels
// (the first subgroup of a group keeps its label, so it is not rewritten);
// with SA, an item that is not kept (alone in its subgroup: its label is its
// final slot + 1) also writes its SA entry p.  Items leave the round in
// sorted order, so those writes land close together.  With run offsets in
// the keys (kmin = 2h of the round, sitems: the S sort), a label no later
// round reads is not written: an S item below offset kmin, an L item at an
// offset in [1, kmin); an L item at offset 0 is read only by the S sort,
// which sees final labels, so it is written once, when it finishes (first
// subgroup or not: earlier rounds skipped it).  wall: an earlier round
// skipped writes and this one carries no offsets, so every label is written.
// FinishSink: an item that is not kept records its subgroup's slot, and the
// subgroup's first item marks that slot as a group start.
struct ChainSink {
    u32 *RK;
    bool lastconv;
    u32 *SA;
    u64 m;
    u32 *err;
    u32 kmin = 0;  // 0: no run offsets in the keys
    bool sitems = false;
    bool wall = false;
    __device__ __forceinline__ void operator()(u32 p, u32 nl, bool first_sub, bool kept, bool,
                                               u64 key) const {
        const bool changed = lastconv || !first_sub;
        bool w;
        if (kmin) {
            const u32 off = (u32)(key >> RUN_KEY_SHIFT);
            w = (!sitems && off == 0) ? !kept : (off >= kmin && changed);
        } else {
            w = changed || wall;
        }
        if (w) RK[p] = nl;
        if (SA && !kept) {
            if (nl >= 1 && nl <= m) SA[nl - 1] = p;
            else atomicOr(err, 64u);
        }
    }
    // group id of a kept item for the next round (with its run offset)
    __device__ __forceinline__ u32 gid(u32 q, u64 key) const {
        return kmin ? q | ((u32)(key >> RUN_KEY_SHIFT) << RUN_GID_SHIFT) : q;
    }
};
struct FinishSink {
    u32 *L, *F;
    __device__ __forceinline__ void operator()(u32 p, u32 nl, bool, bool kept, bool sstart,
                                               u64) const {
        if (!kept) {
            L[p] = nl;
            if (sstart) F[nl] = 1;
        }
    }
    __device__ __forceinline__ u32 gid(u32 q, u64) const { return q; }
};

constexpr int REL_ROUNDS = 16;
constexpr u64 REL_TILE = (u64)BLOCK * REL_ROUNDS;

// The four inclusive block scans of a relabel round at once (max of group
// starts, max of subgroup starts, sum of kept items, sum of kept subgroups):
// four independent wave scans and one LDS exchange, two barriers instead of
// eight.  agg receives the four block aggregates.
__device__ __forceinline__ void rel_scan4(u32 &g, u32 &s, u32 &k, u32 &q, u32 (*lds)[NWAVES],
                                          u32 *agg) {
    g = wave_incl<OpMax>(g);
    s = wave_incl<OpMax>(s);
    k = wave_incl<OpSum>(k);
    q = wave_incl<OpSum>(q);
    const int w = wave_id();
    if (lane_id() == WAVE - 1) {
        lds[0][w] = g;
        lds[1][w] = s;
        lds[2][w] = k;
        lds[3][w] = q;
    }
    __syncthreads();
    u32 pg = 0, ps = 0, pk = 0, pq = 0, ag = 0, as = 0, ak = 0, aq = 0;
    #pragma unroll
    for (int w2 = 0; w2 < NWAVES; w2++) {
        const u32 x0 = lds[0][w2], x1 = lds[1][w2], x2 = lds[2][w2], x3 = lds[3][w2];
        if (w2 < w) {
            pg = pg > x0 ? pg : x0;
            ps = ps > x1 ? ps : x1;
            pk += x2;
            pq += x3;
        }
        ag = ag > x0 ? ag : x0;
        as = as > x1 ? as : x1;
        ak += x2;
        aq += x3;
    }
    __syncthreads();
    g = g > pg ? g : pg;
    s = s > ps ? s : ps;
    k += pk;
    q += pq;
    agg[0] = ag;
    agg[1] = as;
    agg[2] = ak;
    agg[3] = aq;
}

__global__ void k_rel_reduce(RelIn in, u32 *tg, u32 *ts, u32 *tk, u32 *te, u32 *tq) {
    __shared__ u32 l1[NWAVES + 1];
    __shared__ u32 l4x[4][NWAVES];
    u64 base = (u64)blockIdx.x * REL_TILE;
    u32 g = 0, s = 0, k = 0, q = 0, e = 0xFFFFFFFFu;
    for (int r = 0; r < REL_ROUNDS; r++) {
        u64 i = base + (u64)r * BLOCK + threadIdx.x;
        if (i < in.A) {
            bool gs, ss, kp, ge;
            in.flags(i, gs, ss, kp, ge);
            if (gs) g = (u32)i;
            if (ss) s = (u32)i;
            k += kp;
            q += kp && ss;
            if (ge && (u32)i < e) e = (u32)i;
        }
    }
    u32 ag4[4];
    rel_scan4(g, s, k, q, l4x, ag4);
    const u32 a1 = ag4[0], a2 = ag4[1], a3 = ag4[2], a4 = ag4[3];
    // min of group ends: minimum over the block
    u32 x = e;
    #pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        u32 y = __shfl_xor(x, o, WAVE);
        x = x < y ? x : y;
    }
    if (lane_id() == 0) l1[wave_id()] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 m = 0xFFFFFFFFu;
        for (int w = 0; w < NWAVES; w++) m = m < l1[w] ? m : l1[w];
        tg[blockIdx.x] = a1;
        ts[blockIdx.x] = a2;
        tk[blockIdx.x] = a3;
        te[blockIdx.x] = m;
        tq[blockIdx.x] = a4;
    }
}

// one workgroup: a[i] = min(a[i+1 ..]) (exclusive suffix minimum)
__device__ void rmin_tiles_wg(u32 *a, u32 nt) {
    __shared__ u32 wt[NWAVES];
    __shared__ u32 carry_sh;
    const int l = lane_id(), w = wave_id();
    if (threadIdx.x == 0) carry_sh = 0xFFFFFFFFu;
    __syncthreads();
    for (i64 base = (i64)((nt - 1) / BLOCK) * BLOCK; base >= 0; base -= BLOCK) {
        u64 i = (u64)base + threadIdx.x;
        u32 v = i < nt ? a[i] : 0xFFFFFFFFu;
        u32 x = v;
        #pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            u32 y = __shfl_down(x, o, WAVE);
            if (l + o < WAVE) x = x < y ? x : y;
        }
        if (l == 0) wt[w] = x;
        __syncthreads();
        u32 carry = carry_sh;
        u32 after = carry;
        for (int w2 = NWAVES - 1; w2 > w; w2--) after = after < wt[w2] ? after : wt[w2];
        u32 xn = __shfl_down(x, 1, WAVE);
        u32 ex = (l + 1 < WAVE) ? (xn < after ? xn : after) : after;
        if (i < nt) a[i] = ex;
        __syncthreads();
        if (threadIdx.x == 0) {
            u32 m = carry;
            for (int w2 = 0; w2 < NWAVES; w2++) m = m < wt[w2] ? m : wt[w2];
            carry_sh = m;
        }
        __syncthreads();
    }
}

// The tile scans of one relabel round in one launch, a workgroup each:
// exclusive max of group starts (tg) and of subgroup starts (ts), exclusive
// sums of kept items (tk, total to cnt[0]) and of kept subgroups (tq, total
// to cnt[1]), and, with last-slot labels, the exclusive suffix minimum of
// group ends (te).  Grid: 4 blocks, or 5 with te.  cnt[2..3] (the next
// round's large groups, added up by the apply) are cleared.
__global__ void k_rel_tiles(u32 *tg, u32 *ts, u32 *tk, u32 *tq, u32 *te, u32 nt, u32 *cnt) {
    switch (blockIdx.x) {
        case 0:
            scan_tiles_wg<OpMax>(tg, nt, nullptr);
            if (threadIdx.x == 0) cnt[2] = cnt[3] = 0;
            break;
        case 1: scan_tiles_wg<OpMax>(ts, nt, nullptr); break;
        case 2: scan_tiles_wg<OpSum>(tk, nt, cnt); break;
        case 3: scan_tiles_wg<OpSum>(tq, nt, cnt + 1); break;
        default: rmin_tiles_wg(te, nt); break;
    }
}

// Relabel and compact one round.  hi of a key is the item's group: its
// label itself (glab == nullptr) or a dense group id whose label is glab[hi].
// Kept items (not yet alone in their subgroup) go to nxt in sorted order, with
// the dense id of their subgroup among the kept ones in gnx and that
// subgroup's label in glnx[id].  The kept subgroups of more than GS_LOCAL
// items (the next round's large groups, which the group sort takes through
// LSD passes in RS_TILE-item tiles) add their tiles to cnt[2] and their items
// to cnt[3], so the next round is sized by the one readback of cnt.
template <class Sink>
__global__ void k_rel_apply(RelIn in, const u32 *V, Sink sink, u32 *nxt, const u32 *tg,
                            const u32 *ts, const u32 *tk, const u32 *te, const u32 *tq,
                            bool lastconv, u64 mm, const u32 *glab, u32 *gnx, u32 *glnx,
                            u32 *err, u32 *cnt) {
    __shared__ u32 l4x[4][NWAVES];
    __shared__ u32 wt[NWAVES];
    __shared__ u32 g1sh[REL_TILE];
    const u64 base = (u64)blockIdx.x * REL_TILE;
    const int l = lane_id(), w = wave_id();
    if (lastconv) {
        // g1 of every tile item: inclusive suffix minimum of group ends,
        // rounds right to left, seeded with the tiles to the right
        u32 carry = te[blockIdx.x];
        for (int r = REL_ROUNDS - 1; r >= 0; r--) {
            u64 i = base + (u64)r * BLOCK + threadIdx.x;
            u32 e = 0xFFFFFFFFu;
            if (i < in.A) {
                bool gs, ss, kp, ge;
                in.flags(i, gs, ss, kp, ge);
                if (ge) e = (u32)i;
            }
            u32 x = e;
            #pragma unroll
            for (int o = 1; o < WAVE; o <<= 1) {
                u32 y = __shfl_down(x, o, WAVE);
                if (l + o < WAVE) x = x < y ? x : y;
            }
            if (l == 0) wt[w] = x;
            __syncthreads();
            u32 after = carry;
            for (int w2 = NWAVES - 1; w2 > w; w2--) after = after < wt[w2] ? after : wt[w2];
            g1sh[r * BLOCK + threadIdx.x] = x < after ? x : after;
            u32 rm = carry;
            for (int w2 = 0; w2 < NWAVES; w2++) rm = rm < wt[w2] ? rm : wt[w2];
            __syncthreads();
            carry = rm;
        }
    }
    u32 rg = tg[blockIdx.x], rs = ts[blockIdx.x], rk = tk[blockIdx.x], rq = tq[blockIdx.x];
    for (int r = 0; r < REL_ROUNDS; r++) {
        u64 i = base + (u64)r * BLOCK + threadIdx.x;
        bool gs = false, ss = false, kp = false, ge = false, ns = false;
        if (i < in.A) in.flags(i, gs, ss, kp, ge, ns);
        u32 g0 = gs ? (u32)i : 0u, s0 = ss ? (u32)i : 0u;
        u32 kk = kp ? 1u : 0u, qq = (kp && ss) ? 1u : 0u;
        u32 ag4[4];
        rel_scan4(g0, s0, kk, qq, l4x, ag4);
        const u32 a1 = ag4[0], a2 = ag4[1], a3 = ag4[2], a4 = ag4[3];
        g0 = g0 > rg ? g0 : rg;
        s0 = s0 > rs ? s0 : rs;
        kk += rk;
        qq += rq;
        if (i < in.A) {
            const u64 kraw = in.K[i];
            u64 hi = (kraw & in.kmask) >> in.B;
            u64 lab = glab ? (u64)glab[hi] : hi;
            if (lab > mm) {
                atomicOr(err, 32u);
            } else {
                u32 p = V[i];
                u32 nl = lastconv ? (u32)lab - (g1sh[r * BLOCK + threadIdx.x] - s0)
                                  : (u32)lab + (s0 - g0);
                sink(p, nl, s0 == g0, kp, ss, kraw);
                if (kp) {
                    nxt[kk - 1] = p;
                    gnx[kk - 1] = sink.gid(qq - 1, kraw);
                    if (ss) glnx[qq - 1] = nl;
                }
            }
        }
        {
            // the last item of a kept subgroup knows its size
            const u32 sz = (kp && !ns) ? (u32)(i + 1 - s0) : 0u;
            const bool big = sz > GS_LOCAL;
            if (ballot(big)) {
                u32 tl = big ? (u32)((sz + RS_TILE - 1) / RS_TILE) : 0u, it = big ? sz : 0u;
                #pragma unroll
                for (int o = WAVE / 2; o > 0; o >>= 1) {
                    tl += __shfl_xor(tl, o, WAVE);
                    it += __shfl_xor(it, o, WAVE);
                }
                if (lane_id() == 0) {
                    atomicAdd(cnt + 2, tl);
                    atomicAdd(cnt + 3, it);
                }
            }
        }
        rg = rg > a1 ? rg : a1;
        rs = rs > a2 ? rs : a2;
        rk += a3;
        rq += a4;
    }
}

}  // namespace psk
