// psk_naming.hpp -- LMS-substring naming, reduced string, translation to
// bucket-pointer form, inverse shortcut and permutation.
//
// Reference: name_lmssubstrings (naming.rs:20-259), its equality rule
// UintFingerprint::equals (naming.rs:388-408), translate (naming.rs:263-274),
// permutate_lmssuffixes (naming.rs:413-430), the inverse shortcut
// (psacak8.rs:41-47, psacak32.rs:44-50).
//
// GPU form: the LMS substring of an LMS position p runs to the next LMS
// position inclusive (the crate's fingerprint length, naming.rs:277-306); the
// last LMS substring runs into the sentinel and equals nothing.  One thread
// per adjacent sorted pair compares lengths, then symbols; an exclusive scan
// of the difference flags gives the names; each name is scattered to the
// text-order index of its position, which a rank table R[p/2] provides
// (LMS positions are at least 2 apart).
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

// Returns k (number of distinct names).  reduced[r] = name of the r-th LMS
// position in text order; tails[c] = number of LMS substrings with name <= c.
// Batches: EB marks each text's last position; an LMS substring that would
// run past it reaches that text's sentinel and equals nothing.
template <class C>
inline u32 name_lms(Dev &d, const C *T, const u32 *lms_sorted, const u32 *lms_text,
                    const u32 *R, u32 m1, u32 *diff, u32 *reduced, u32 *tails, u32 *d_k,
                    const u32 *EB = nullptr) {
    if (m1 == 0) return 0;
    for_each(d, m1, [=] __device__(u64 j) {
        if (j == 0) { diff[0] = 1; return; }
        u32 a = lms_sorted[j - 1], b = lms_sorted[j];
        u32 ra = R[a >> 1], rb = R[b >> 1];
        if (ra + 1 == m1 || rb + 1 == m1) { diff[j] = 1; return; }  // sentinel substring
        u32 na = lms_text[ra + 1], nb = lms_text[rb + 1];
        if (EB && (range_has_bit(EB, a, na) || range_has_bit(EB, b, nb))) { diff[j] = 1; return; }
        u32 la = na - a + 1, lb = nb - b + 1;
        diff[j] = (la != lb || !lms_substr_equal(T, a, b, la)) ? 1u : 0u;
    });
    scan_exclusive(
        d, m1, [=] __device__(u64 j) { return diff[j]; },
        [=] __device__(u64 j, u32 ex, u32 v) {
            u32 name = ex + v - 1;
            u32 p = lms_sorted[j];
            reduced[R[p >> 1]] = name;
            if (j + 1 == m1 || diff[j + 1]) tails[name] = (u32)j + 1;
        },
        d_k);
    return d.read_u32(d_k);
}

// translate (naming.rs:263-274): L-type name c -> bucket head (tails[c-1] or
// 0), S-type -> tails[c] - 1.  sbits are the types of the name string.
inline void translate_names(Dev &d, u32 *reduced, u32 m1, const u32 *sbits, const u32 *tails) {
    for_each(d, m1, [=] __device__(u64 i) {
        u32 c = reduced[i];
        reduced[i] = getbit(sbits, i) ? tails[c] - 1 : (c ? tails[c - 1] : 0u);
    });
}

// k == m1: every name is unique, so the reduced text is its own inverse SA.
inline void inverse_names(Dev &d, const u32 *reduced, u32 m1, u32 *SA1) {
    for_each(d, m1, [=] __device__(u64 i) { SA1[reduced[i]] = (u32)i; });
}

// SA1 (indices into the text-order LMS list) -> LMS positions, in place.
inline void permute_lms(Dev &d, u32 *SA1, u32 m1, const u32 *lms_text) {
    for_each(d, m1, [=] __device__(u64 j) { SA1[j] = lms_text[SA1[j]]; });
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
