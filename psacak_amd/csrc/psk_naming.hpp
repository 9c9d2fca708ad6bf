// psk_naming.hpp -- LMS-substring naming, reduced string, translation to
// bucket-pointer form, inverse shortcut and permutation.
//
// Reference: name_lmssubstrings (naming.rs:20-259), its equality rule
// UintFingerprint::equals (naming.rs:388-408), translate (naming.rs:263-274),
// permutate_lmssuffixes (naming.rs:413-430), the inverse shortcut
// (psacak8.rs:41-47, psacak32.rs:44-50).
//
// GPU form: the LMS substring of an LMS position p runs to the next LMS
// position inclusive (the crate's fingerprint length, naming.rs:277-306),
// found from the LMS bit vector; the last LMS substring runs into the sentinel
// and equals nothing.  One thread per sorted item computes its length and
// packed first symbols and compares them with its predecessor's (symbols past
// the packed prefix are compared in the text); an exclusive scan of the
// difference flags gives the names.
#pragma once

#include "psk_types.hpp"

namespace psk {

template <class C>
__device__ __forceinline__ bool lms_substr_equal(const C *T, u32 a, u32 b, u32 len) {
    for (u32 k = 0; k < len; k++)
        if (T[a + k] != T[b + k]) return false;
    return true;
}

// Any set bit of bits in positions [p, q)?
__device__ __forceinline__ bool range_has_bit(const u32 *bits, u32 p, u32 q) {
    for (u32 w = p >> 5; w <= ((q - 1) >> 5); w++) {
        u32 x = bits[w];
        if (w == (p >> 5)) x &= 0xFFFFFFFFu << (p & 31);
        if (w == ((q - 1) >> 5)) x &= 0xFFFFFFFFu >> (31 - ((q - 1) & 31));
        if (x) return true;
    }
    return false;
}

// First symbols of an LMS substring packed into one word: 8 bytes of a byte
// text, 2 symbols of a u32 text (an LMS substring of a u32 level that has a
// next LMS position holds at least 3 symbols).
template <class C> struct LmsFp;
template <> struct LmsFp<u8> {
    static constexpr u32 cap = 8;
    __device__ static u64 pack(const u8 *T, u32 p, u32 len) {
        u64 f = 0;
        u32 e = len < cap ? len : cap;
        #pragma unroll
        for (u32 k = 0; k < cap; k++) f = (f << 8) | (k < e ? (u64)T[p + k] : 0ull);
        return f;
    }
};
template <> struct LmsFp<u32> {
    static constexpr u32 cap = 2;
    __device__ static u64 pack(const u32 *T, u32 p, u32) { return ((u64)T[p] << 32) | T[p + 1]; }
};

// LMS substring of the LMS position p (naming.rs:277-306: it runs to the
// next LMS position inclusive): its length, or 0 when it is unique by
// construction -- there is no next LMS position (it runs into the sentinel),
// or (batches) it runs past its text's last position (EB).
template <class C>
__device__ __forceinline__ u32 lms_len(const u32 *lmsbits, u64 nwords, const u32 *EB, u32 p) {
    u64 w = p >> 5;
    u32 x = lmsbits[w] & ~((2u << (p & 31)) - 1u);
    while (x == 0 && ++w < nwords) x = lmsbits[w];
    if (x == 0) return 0;
    u32 q = (u32)(w * 32 + (u32)(__ffs(x) - 1));
    if (EB && range_has_bit(EB, p, q)) return 0;
    return q - p + 1;
}

// diff[j] = 1 iff the j-th sorted LMS substring differs from the (j-1)-th
// (UintFingerprint::equals, naming.rs:388-408: same length, same symbols).
// One item per thread; a block also evaluates the item before its first.
template <class C>
__global__ void __launch_bounds__(BLOCK) k_name_diff(const C *T, const u32 *lms_sorted, u32 m1,
                                                     const u32 *lmsbits, u64 nwords,
                                                     const u32 *EB, u8 *diff) {
    __shared__ u32 sl[BLOCK + 1];
    __shared__ u64 sf[BLOCK + 1];
    const u64 j = (u64)blockIdx.x * BLOCK + threadIdx.x;
    u32 p = 0, len = 0;
    u64 fp = 0;
    if (j < m1) {
        p = lms_sorted[j];
        len = lms_len<C>(lmsbits, nwords, EB, p);
        if (len) fp = LmsFp<C>::pack(T, p, len);
    }
    sl[threadIdx.x + 1] = len;
    sf[threadIdx.x + 1] = fp;
    if (threadIdx.x == 0) {
        u32 l0 = 0;
        u64 f0 = 0;
        if (j > 0 && j < m1) {
            u32 a = lms_sorted[j - 1];
            l0 = lms_len<C>(lmsbits, nwords, EB, a);
            if (l0) f0 = LmsFp<C>::pack(T, a, l0);
        }
        sl[0] = l0;
        sf[0] = f0;
    }
    __syncthreads();
    if (j >= m1) return;
    if (j == 0) { diff[0] = 1; return; }
    const u32 la = sl[threadIdx.x];
    bool dif = la == 0 || len == 0 || la != len || sf[threadIdx.x] != fp;
    if (!dif && len > LmsFp<C>::cap) {
        u32 a = lms_sorted[j - 1];
        const u32 c = LmsFp<C>::cap;
        dif = !lms_substr_equal(T, a + c, p + c, len - c);
    }
    diff[j] = dif ? 1 : 0;
}

// Returns k (number of distinct names).  reduced[r] = name of the r-th LMS
// position in text order; tails[c] = number of LMS substrings with name <= c.
// Names go first to a position-indexed table NP[p/2] (LMS positions are at
// least 2 apart), then to text order by one monotone gather.
// Batches: EB marks each text's last position; an LMS substring that would
// run past it reaches that text's sentinel and equals nothing.
template <class C>
inline u32 name_lms(Dev &d, const C *T, const u32 *lms_sorted, const u32 *lms_text,
                    const u32 *lmsbits, u64 nwords, u32 *NP, u32 m1, u8 *diff, u32 *reduced,
                    u32 *tails, u32 *d_k, const u32 *EB = nullptr) {
    if (m1 == 0) return 0;
    k_name_diff<C><<<ceil_div(m1, BLOCK), BLOCK, 0, d.s>>>(T, lms_sorted, m1, lmsbits, nwords, EB,
                                                          diff);
    PSK_LAUNCH_CHECK();
    scan_exclusive(
        d, m1, [=] __device__(u64 j) { return (u32)diff[j]; },
        [=] __device__(u64 j, u32 ex, u32 v) {
            u32 name = ex + v - 1;
            NP[lms_sorted[j] >> 1] = name;
            if (j + 1 == m1 || diff[j + 1]) tails[name] = (u32)j + 1;
        },
        d_k);
    for_each(d, m1, [=] __device__(u64 r) { reduced[r] = NP[lms_text[r] >> 1]; });
    return d.read_u32(d_k);
}

// translate (naming.rs:263-274): L-type name c -> bucket head (tails[c-1] or
// 0), S-type -> tails[c] - 1.  sbits are the types of the name string.
// Four names per thread with 16-byte accesses when both arrays allow them.
inline void translate_names(Dev &d, const u32 *names, u32 *out, u32 m1, const u32 *sbits,
                            const u32 *tails) {
    auto tr = [=] __device__(u32 c, u32 s) { return s ? tails[c] - 1 : (c ? tails[c - 1] : 0u); };
    const bool vec = ((((uintptr_t)names) | ((uintptr_t)out)) & 15u) == 0;
    const u64 nq = vec ? m1 / 4 : 0;
    for_each(d, nq + (m1 - 4 * nq), [=] __device__(u64 q) {
        if (q < nq) {
            const u64 i = 4 * q;
            const uint4 c = *(const uint4 *)(names + i);
            const u32 t = (sbits[i >> 5] >> (i & 31)) & 15u;
            *(uint4 *)(out + i) = make_uint4(tr(c.x, t & 1u), tr(c.y, (t >> 1) & 1u),
                                             tr(c.z, (t >> 2) & 1u), tr(c.w, t >> 3));
        } else {
            const u64 i = 4 * nq + (q - nq);
            out[i] = tr(names[i], getbit(sbits, i));
        }
    });
}

// k == m1: every name is unique, so the reduced text is its own inverse SA
// (psacak8.rs:41-47): SA1[reduced[i]] = i.
inline void inverse_names(Dev &d, const u32 *reduced, u32 m1, u32 *SA1) {
    for_each(d, m1, [=] __device__(u64 i) { SA1[reduced[i]] = (u32)i; });
}

// Brute-force SA for n <= 3 (saca_tiny, common.rs:13-25).
template <class C>
__global__ void k_naive_small(const C *t, u32 n, u32 *sa) {
    u32 idx[3] = {0, 1, 2};
    for (u32 i = 1; i < n; i++) {
        for (u32 j = i; j > 0; j--) {
            u32 a = idx[j - 1], b = idx[j];
            // is suffix b < suffix a ?
            u32 x = a, y = b;
            while (x < n && y < n && t[x] == t[y]) { x++; y++; }
            bool less = (y == n) ? (x != n) : (x != n && t[y] < t[x]);
            if (!less) break;
            idx[j - 1] = b;
            idx[j] = a;
        }
    }
    for (u32 i = 0; i < n; i++) sa[i] = idx[i];
}

}  // namespace psk
