// psk_types.hpp -- L/S typing and LMS extraction on the GPU.
//
// Reference semantics (common.rs:43-68): T[n-1] is L-type; position i is
// S-type iff T[i] < T[i+1], or T[i] == T[i+1] and i+1 is S-type; i is LMS iff
// it is S-type, i > 0 and i-1 is L-type.
//
// GPU formulation: a position is "decided" when T[i] != T[i+1] (its type is
// T[i] < T[i+1]); an undecided position takes the type of the nearest decided
// position to its right.  Each thread owns TY_WPT 32-position words and
// builds (decided, value) bit masks; the carry into a word is the type of the
// first position of the next word, resolved with wave ballots inside a
// workgroup and with a right-to-left pass over workgroup aggregates.
// Output: S-type and LMS bit vectors (bit b of word w = position 32w+b).
#pragma once

#include "psk_prim.hpp"

namespace psk {

constexpr int TY_WPT = 4;
constexpr u64 TY_POS_PER_BLOCK = (u64)BLOCK * TY_WPT * 32;

// A batch of texts concatenated into one array carries an end bit vector
// (bit set at each text's last position): a text end is always L-type, as if
// followed by its own sentinel, so no carry crosses a text boundary.
__device__ __forceinline__ void apply_ends(const u32 *endbits, u64 p0, u64 n, u32 &D, u32 &V) {
    if (endbits && p0 < n) {
        u32 e = endbits[p0 >> 5];
        D |= e;
        V &= ~e;
    }
}

template <class C>
__device__ __forceinline__ void type_masks(const C *t, u64 n, u64 p0, u32 &D, u32 &V,
                                           const u32 *endbits) {
    D = 0;
    V = 0;
    C a = (p0 < n) ? t[p0] : (C)0;
    #pragma unroll 8
    for (int b = 0; b < 32; b++) {
        u64 p = p0 + b;
        if (p + 1 < n) {
            C c = t[p + 1];
            if (a != c) {
                D |= 1u << b;
                V |= (u32)(a < c) << b;
            }
            a = c;
        } else {
            D |= 1u << b;  // p == n-1 is L-type; p >= n is padding (L)
        }
    }
    apply_ends(endbits, p0, n, D, V);
}

// Byte texts: 32 bytes (+1) per word with dword loads when the text is
// 4-byte aligned and the word is not the last one.
template <>
__device__ __forceinline__ void type_masks<u8>(const u8 *t, u64 n, u64 p0, u32 &D, u32 &V,
                                               const u32 *endbits) {
    if (p0 + 33 <= n && (((uintptr_t)t) & 3u) == 0) {
        const u32 *q = (const u32 *)(t + p0);
        u32 w[9];
        #pragma unroll
        for (int k = 0; k < 8; k++) w[k] = q[k];
        w[8] = t[p0 + 32];
        D = 0;
        V = 0;
        #pragma unroll
        for (int b = 0; b < 32; b++) {
            u32 a = (w[b >> 2] >> (8 * (b & 3))) & 255u;
            u32 c = (w[(b + 1) >> 2] >> (8 * ((b + 1) & 3))) & 255u;
            D |= (u32)(a != c) << b;
            V |= (u32)(a < c) << b;
        }
        apply_ends(endbits, p0, n, D, V);
        return;
    }
    D = 0;
    V = 0;
    u32 a = (p0 < n) ? t[p0] : 0u;
    for (int b = 0; b < 32; b++) {
        u64 p = p0 + b;
        if (p + 1 < n) {
            u32 c = t[p + 1];
            D |= (u32)(a != c) << b;
            V |= (u32)(a < c) << b;
            a = c;
        } else {
            D |= 1u << b;
        }
    }
    apply_ends(endbits, p0, n, D, V);
}

// Type bits of one word given the type of the position right after it.
__device__ __forceinline__ u32 resolve_word(u32 D, u32 V, u32 carry) {
    u32 S = 0, cur = carry & 1u;
    #pragma unroll
    for (int b = 31; b >= 0; b--) {
        if ((D >> b) & 1u) cur = (V >> b) & 1u;
        S |= cur << b;
    }
    return S;
}

// For every thread: the `first` value of the nearest higher-numbered thread of
// the workgroup whose `has` is set, or block_carry if there is none.
__device__ __forceinline__ u32 carry_from_right(bool has, u32 first, u32 block_carry, u32 *lh,
                                                u32 *lf) {
    const int l = lane_id(), w = wave_id();
    u64 m = ballot(has);
    u32 wf = __shfl(first, m ? (__ffsll((long long)m) - 1) : 0, WAVE);
    if (l == 0) {
        lh[w] = m != 0;
        lf[w] = wf;
    }
    __syncthreads();
    u32 wc = block_carry;
    for (int w2 = NWAVES - 1; w2 > w; w2--)
        if (lh[w2]) wc = lf[w2];
    u64 higher = (l == WAVE - 1) ? 0ull : (m & ~((2ull << l) - 1ull));
    int src = higher ? (__ffsll((long long)higher) - 1) : 0;
    u32 f = __shfl(first, src, WAVE);
    __syncthreads();
    return higher ? f : wc;
}

// Mask sources for the aggregate / final kernels: word w -> (D, V).
// From the text (byte texts, dword loads), or from mask words that
// k_type_masks wrote (wider symbols).
template <class C>
struct MasksFromText {
    const C *t;
    u64 n;
    const u32 *endbits;
    __device__ void operator()(u64 w, u32 &D, u32 &V) const { type_masks(t, n, w * 32, D, V, endbits); }
};
struct MasksFromWords {
    const u32 *Dw, *Vw;
    u64 nwords;
    __device__ void operator()(u64 w, u32 &D, u32 &V) const {
        if (w < nwords) {
            D = Dw[w];
            V = Vw[w];
        } else {
            D = 0xFFFFFFFFu;  // padding positions are L-type
            V = 0;
        }
    }
};

// D and V words of t with wave-wide loads: per step the 64 lanes read 64
// consecutive symbols (coalesced) and two ballots give two words; lane l
// keeps word l of its wave's 64.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_type_masks(const C *t, u64 n, u64 nwords,
                                                     const u32 *endbits, u32 *Dw, u32 *Vw) {
    const int l = lane_id();
    const u64 wbase = ((u64)blockIdx.x * NWAVES + wave_id()) * WAVE;
    u32 myD = 0, myV = 0;
    #pragma unroll 4
    for (int s = 0; s < WAVE / 2; s++) {
        u64 p = wbase * 32 + (u64)s * 64 + l;
        u32 c0 = p < n ? (u32)t[p] : 0u;
        u32 c1 = __shfl_down(c0, 1, WAVE);
        if (l == WAVE - 1) c1 = p + 1 < n ? (u32)t[p + 1] : 0u;
        bool last = p + 1 >= n;  // p == n-1 is L-type, p >= n padding (L)
        u64 bd = ballot(last || c0 != c1);
        u64 bv = ballot(!last && c0 < c1);
        if (l == 2 * s) {
            myD = (u32)bd;
            myV = (u32)bv;
        } else if (l == 2 * s + 1) {
            myD = (u32)(bd >> 32);
            myV = (u32)(bv >> 32);
        }
    }
    u64 w = wbase + l;
    if (w < nwords) {
        apply_ends(endbits, w * 32, n, myD, myV);
        Dw[w] = myD;
        Vw[w] = myV;
    }
}

// D and V words with 16-byte loads (text 16-byte aligned): per step lane l
// holds P = 16 / sizeof(C) consecutive symbols of the wave's 64 P, the
// symbol after its last one comes from lane l + 1 (lane 63 loads it), and
// the 32 / P lanes of a word OR their bit groups together.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_type_masks_v(const C *t, u64 n, u64 nwords,
                                                       const u32 *endbits, u32 *Dw, u32 *Vw) {
    constexpr int P = 16 / (int)sizeof(C);
    constexpr int LPW = 32 / P;
    static_assert(P * (int)sizeof(C) == 16 && LPW * P == 32, "symbol width");
    const int l = lane_id();
    const u64 nsteps = (n + 64 * P - 1) / (64 * P);
    const u64 nw = (u64)gridDim.x * NWAVES;
    for (u64 st = (u64)blockIdx.x * NWAVES + wave_id(); st < nsteps; st += nw) {
        const u64 p0 = (st * WAVE + l) * P;
        C c[P];
        if (p0 + P <= n) {
            const uint4 v = *(const uint4 *)(t + p0);
            __builtin_memcpy(c, &v, 16);
        } else {
            #pragma unroll
            for (int j = 0; j < P; j++) c[j] = p0 + j < n ? t[p0 + j] : (C)0;
        }
        C nx = (C)__shfl_down((u32)c[0], 1, WAVE);
        if (l == WAVE - 1) nx = p0 + P < n ? t[p0 + P] : (C)0;
        u32 dm = 0, vm = 0;
        #pragma unroll
        for (int j = 0; j < P; j++) {
            const C a = c[j], b = j + 1 < P ? c[j + 1] : nx;
            const bool last = p0 + j + 1 >= n;  // p == n-1 is L-type, p >= n padding (L)
            dm |= (u32)(last || a != b) << j;
            vm |= (u32)(!last && a < b) << j;
        }
        u32 D = dm << (P * (l % LPW)), V = vm << (P * (l % LPW));
        #pragma unroll
        for (int o = 1; o < LPW; o <<= 1) {
            D |= __shfl_xor(D, o, WAVE);
            V |= __shfl_xor(V, o, WAVE);
        }
        const u64 w = (st * WAVE + l) / LPW;
        if (l % LPW == 0 && w < nwords) {
            apply_ends(endbits, w * 32, n, D, V);
            Dw[w] = D;
            Vw[w] = V;
        }
    }
}

template <class M>
__global__ void k_type_agg(M masks, u8 *blk_has, u8 *blk_first) {
    __shared__ u32 lh[NWAVES], lf[NWAVES];
    u64 w0 = ((u64)blockIdx.x * BLOCK + threadIdx.x) * TY_WPT;
    bool has = false;
    u32 first = 0;
    for (int k = 0; k < TY_WPT; k++) {
        u32 D, V;
        masks(w0 + k, D, V);
        if (!has && D) {
            has = true;
            first = (V >> (__ffs(D) - 1)) & 1u;
        }
    }
    u64 m = ballot(has);
    u32 f = __shfl(first, m ? (__ffsll((long long)m) - 1) : 0, WAVE);
    if (lane_id() == 0) {
        lh[wave_id()] = m != 0;
        lf[wave_id()] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u8 bh = 0, bf = 0;
        for (int w = 0; w < NWAVES; w++)
            if (lh[w]) { bh = 1; bf = (u8)lf[w]; break; }
        blk_has[blockIdx.x] = bh;
        blk_first[blockIdx.x] = bf;
    }
}

// One workgroup: carry[b] = type of the first position of block b+1 (0 for
// the last block), from the block aggregates, right to left.
constexpr int TYC_IPT = 16;
__global__ void k_type_carry(const u8 *blk_has, const u8 *blk_first, u32 nb, u8 *carry) {
    __shared__ u32 lh[NWAVES], lf[NWAVES];
    __shared__ u32 chunk_carry;
    if (threadIdx.x == 0) chunk_carry = 0;
    __syncthreads();
    const u64 CH = (u64)BLOCK * TYC_IPT;
    u64 nchunks = (nb + CH - 1) / CH;
    for (u64 ci = nchunks; ci-- > 0;) {
        u64 b0 = ci * CH + (u64)threadIdx.x * TYC_IPT;
        bool has = false;
        u32 first = 0;
        for (int k = 0; k < TYC_IPT; k++) {
            u64 b = b0 + k;
            if (b < nb && blk_has[b]) { has = true; first = blk_first[b]; break; }
        }
        u32 cr = chunk_carry;
        u32 c = carry_from_right(has, first, cr, lh, lf);
        for (int k = TYC_IPT - 1; k >= 0; k--) {
            u64 b = b0 + k;
            if (b < nb) {
                carry[b] = (u8)c;
                if (blk_has[b]) c = blk_first[b];
            }
        }
        // aggregate of this chunk for the chunk on its left: lowest thread with has
        u64 m = ballot(has);
        u32 wf = __shfl(first, m ? (__ffsll((long long)m) - 1) : 0, WAVE);
        if (lane_id() == 0) {
            lh[wave_id()] = m != 0;
            lf[wave_id()] = wf;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 0; w < NWAVES; w++)
                if (lh[w]) { chunk_carry = lf[w]; break; }
        }
        __syncthreads();
    }
}

template <class M>
__global__ void k_type_final(M masks, const u8 *blk_carry, u32 *sbits, u64 nwords, u32 *dbits) {
    __shared__ u32 lh[NWAVES], lf[NWAVES];
    u64 w0 = ((u64)blockIdx.x * BLOCK + threadIdx.x) * TY_WPT;
    u32 D[TY_WPT], V[TY_WPT];
    bool has = false;
    u32 first = 0;
    #pragma unroll
    for (int k = 0; k < TY_WPT; k++) {
        masks(w0 + k, D[k], V[k]);
        if (!has && D[k]) {
            has = true;
            first = (V[k] >> (__ffs(D[k]) - 1)) & 1u;
        }
    }
    u32 c = carry_from_right(has, first, blk_carry[blockIdx.x], lh, lf);
    #pragma unroll
    for (int k = TY_WPT - 1; k >= 0; k--) {
        u32 S = resolve_word(D[k], V[k], c);
        c = S & 1u;
        if (w0 + k < nwords) {
            sbits[w0 + k] = S;
            if (dbits) dbits[w0 + k] = D[k];
        }
    }
}

// S-type bit vector of t[0..n) into sbits[0 .. ceil(n/32)).
// dbits (optional): bit p set where T[p] != T[p+1] or p ends a text (the
// "decided" mask); run lengths of equal symbols are read from it.
template <class C>
inline void compute_types(Dev &d, const C *t, u64 n, u32 *sbits, const u32 *endbits = nullptr,
                          u32 *dbits = nullptr) {
    if (n == 0) return;
    u64 nwords = (n + 31) / 32;
    u32 nb = ceil_div(nwords, (u64)BLOCK * TY_WPT);
    ArenaScope mk(*d.a);
    u8 *bh = d.a->get<u8>(nb);
    u8 *bf = d.a->get<u8>(nb);
    u8 *bc = d.a->get<u8>(nb);
    auto run = [&](auto masks) {
        k_type_agg<<<nb, BLOCK, 0, d.s>>>(masks, bh, bf);
        PSK_LAUNCH_CHECK();
        k_type_carry<<<1, BLOCK, 0, d.s>>>(bh, bf, nb, bc);
        PSK_LAUNCH_CHECK();
        k_type_final<<<nb, BLOCK, 0, d.s>>>(masks, bc, sbits, nwords, dbits);
        PSK_LAUNCH_CHECK();
    };
    const bool vec = (sizeof(C) == 1 || sizeof(C) == 4) && (((uintptr_t)t) & 15u) == 0;
    if (sizeof(C) == 1 && !vec) {
        run(MasksFromText<C>{t, n, endbits});
    } else {
        u32 *Dw = d.a->get<u32>(nwords), *Vw = d.a->get<u32>(nwords);
        if (vec) {
            const u64 steps = ceil_div(n, (u64)WAVE * (16 / sizeof(C)));
            const u64 blocks = std::min<u64>(ceil_div(steps, (u64)NWAVES), 65536);
            k_type_masks_v<C><<<(u32)blocks, BLOCK, 0, d.s>>>(t, n, nwords, endbits, Dw, Vw);
        } else {
            k_type_masks<C><<<ceil_div(nwords, (u64)BLOCK), BLOCK, 0, d.s>>>(t, n, nwords, endbits,
                                                                           Dw, Vw);
        }
        PSK_LAUNCH_CHECK();
        run(MasksFromWords{Dw, Vw, nwords});
    }
    mk.reset();
}

__device__ __forceinline__ u32 getbit(const u32 *bits, u64 i) { return (bits[i >> 5] >> (i & 31)) & 1u; }

__device__ __forceinline__ u32 lms_word(const u32 *sbits, u64 w) {
    u32 S = sbits[w];
    u32 prev = w ? sbits[w - 1] : 0xFFFFFFFFu;  // position 0 is never LMS
    return S & ~((S << 1) | (prev >> 31));
}

// Level-0 (byte) LMS bits + bucket statistics.  Per byte value c:
// cnt[c] occurrences, nL[c] L-type occurrences, nLMS[c] LMS occurrences;
// *m1 = number of LMS positions.  (psacak8.rs:355-374 Buckets::new)
constexpr u64 H_F = 21;  // bits per packed LDS counter field
// startbits (batches): a text's first position has no predecessor in its
// text, so it is never LMS.
// A workgroup takes wpb consecutive words (wpb * 32 < 2^H_F positions, so
// the packed LDS fields cannot overflow): few workgroups, so few global
// atomics on the 3 x (distinct bytes) + 1 result counters, which every
// workgroup adds to.
__global__ void __launch_bounds__(BLOCK) k_lms_hist8(const u8 *t, u64 n, const u32 *sbits,
                                                     u32 *lmsbits, u64 nwords, u64 wpb, u32 *cnt,
                                                     u32 *nL, u32 *nLMS, u32 *m1,
                                                     const u32 *startbits) {
    // four copies (lane & 3), padded onto different banks: a DNA text's
    // wave holds four byte values, whose lanes would otherwise serialise
    // on four counters
    __shared__ unsigned long long h[4][257];
    __shared__ u32 red[NWAVES + 1];
    #pragma unroll
    for (int q = 0; q < 4; q++) h[q][threadIdx.x] = 0;
    __syncthreads();
    const u64 w0 = (u64)blockIdx.x * wpb;
    const u64 w1 = w0 + wpb < nwords ? w0 + wpb : nwords;
    const bool vec = (((uintptr_t)t) & 15u) == 0;
    u32 lcount = 0;
    for (u64 w = w0 + threadIdx.x; w < w1; w += BLOCK) {
        u32 S = sbits[w];
        u32 lms = lms_word(sbits, w);
        if (startbits) lms &= ~startbits[w];
        lmsbits[w] = lms;
        lcount += __popc(lms);
        const u64 p0 = w * 32;
        const u32 lim = (u32)((n - p0) < 32 ? (n - p0) : 32);
        u32 word[8];
        if (vec && lim == 32) {
            const uint4 a0 = *(const uint4 *)(t + p0), a1 = *(const uint4 *)(t + p0 + 16);
            word[0] = a0.x; word[1] = a0.y; word[2] = a0.z; word[3] = a0.w;
            word[4] = a1.x; word[5] = a1.y; word[6] = a1.z; word[7] = a1.w;
        } else {
            #pragma unroll
            for (int q = 0; q < 8; q++) word[q] = 0;
            for (u32 b = 0; b < lim; b++) word[b >> 2] |= (u32)t[p0 + b] << (8 * (b & 3));
        }
        #pragma unroll
        for (u32 b = 0; b < 32; b++) {
            if (b < lim) {
                const u32 c = (word[b >> 2] >> (8 * (b & 3))) & 255u;
                unsigned long long v = 1ull | ((unsigned long long)(~(S >> b) & 1u) << H_F) |
                                       ((unsigned long long)((lms >> b) & 1u) << (2 * H_F));
                atomicAdd(&h[lane_id() & 3][c], v);
            }
        }
    }
    u32 tot;
    block_exclusive_sum(lcount, red, &tot);
    if (threadIdx.x == 0 && tot) atomicAdd(m1, tot);
    __syncthreads();
    unsigned long long v = h[0][threadIdx.x] + h[1][threadIdx.x] + h[2][threadIdx.x] + h[3][threadIdx.x];
    if (v) {
        const unsigned long long M = (1ull << H_F) - 1;
        atomicAdd(&cnt[threadIdx.x], (u32)(v & M));
        u32 l = (u32)((v >> H_F) & M), s = (u32)((v >> (2 * H_F)) & M);
        if (l) atomicAdd(&nL[threadIdx.x], l);
        if (s) atomicAdd(&nLMS[threadIdx.x], s);
    }
}

// LMS bits for a u32 level; *m1 += number of LMS positions.  Grid-stride
// over the words, so few workgroups add to the one counter.
__global__ void __launch_bounds__(BLOCK) k_lms32(const u32 *sbits, u32 *lmsbits, u64 nwords,
                                                 u32 *m1) {
    __shared__ u32 red[NWAVES + 1];
    u32 lcount = 0;
    for (u64 w = (u64)blockIdx.x * BLOCK + threadIdx.x; w < nwords; w += (u64)gridDim.x * BLOCK) {
        u32 lms = lms_word(sbits, w);
        lmsbits[w] = lms;
        lcount += __popc(lms);
    }
    u32 tot;
    block_exclusive_sum(lcount, red, &tot);
    if (threadIdx.x == 0 && tot) atomicAdd(m1, tot);
}

// Positions of set bits (ascending) -> out[0..count).
inline void bits_to_positions(Dev &d, const u32 *bits, u64 nwords, u32 *out) {
    bit_list(d, nwords * 32, [=] __device__(u64 w) { return bits[w]; }, out, nullptr);
}

}  // namespace psk
