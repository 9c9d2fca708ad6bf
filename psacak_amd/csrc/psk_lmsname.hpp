// psk_lmsname.hpp -- LMS-substring naming by sorting the substrings directly.
//
// The names of a level (naming.rs:20-259) only depend on the order and the
// equality of its LMS substrings (UintFingerprint::equals, naming.rs:388-408:
// same length, same symbols).  Instead of inducing that order over every
// suffix of the level, this sorts the m1 substrings themselves by packed
// symbol ranks:
//
//   digit i of the substring at LMS position p (next LMS position q) is
//     rank(T[p+i]) + 1   for i <= q - p     (the symbols, terminal included)
//     sigma + 1          for i == q - p + 1 (a terminator above every symbol)
//     0                  after it;
//   the last substring runs to the text's end and has digit 0 (the
//   sentinel, below every symbol) in place of the terminator, so it is
//   unique and smaller than any substring it prefixes.
//
// A terminator above every symbol puts a substring after every longer
// substring it prefixes, which is the induced (LMS-substring) order: where a
// substring ends at its terminal LMS position, a longer one with the same
// symbols has an L-type position (an S-type one would be an LMS position and
// end it there), and an L-type suffix is smaller than an S-type suffix with
// the same first symbol.  Equal digit strings up to the terminator are equal
// substrings.
//
// Rounds: each sorts the tied, unfinished substrings by (dense id of their
// tie group, next D digits, finished bit) with the LSD radix sort and splits
// the groups; a substring is finished once its terminator has been compared or
// it is alone in its group.  Finished substrings record the slot of their
// group in the final order; a scan over the group-start marks turns slots into
// names.  Long repeated substrings need many rounds: after max_rounds the
// caller falls back to induced sorting.
#pragma once

#include "psk_relabel.hpp"

namespace psk {

// Symbol ranks in [0, sigma) by position.
struct ByteRanks {
    const u8 *T;
    const u8 *crank;  // rank of each byte value among the bytes present
    __device__ __forceinline__ u32 operator()(u64 pos) const { return crank[T[pos]]; }
};
struct WordRanks {
    const u32 *r;
    __device__ __forceinline__ u32 operator()(u64 pos) const { return r[pos]; }
};

// Histogram of LMS-substring lengths (number of symbols, terminal included),
// the last bin collecting everything longer.
constexpr int LEN_BINS = 64;
__global__ void k_lms_lens(const u32 *lms_text, u32 m1, u64 m, u32 *hist) {
    __shared__ u32 h[LEN_BINS];
    if (threadIdx.x < LEN_BINS) h[threadIdx.x] = 0;
    __syncthreads();
    for (u64 r = (u64)blockIdx.x * BLOCK + threadIdx.x; r < m1; r += (u64)gridDim.x * BLOCK) {
        u64 p = lms_text[r];
        u64 len = r + 1 < m1 ? (u64)lms_text[r + 1] - p + 1 : m - p;
        atomicAdd(&h[len < LEN_BINS - 1 ? len : LEN_BINS - 1], 1u);
    }
    __syncthreads();
    if (threadIdx.x < LEN_BINS && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Digits of round 1: radix passes vs. the substrings left for later rounds.
inline int lms_first_digits(const u32 *hist, u32 m1, int Db) {
    // unterminated after D digits: length >= D
    auto rem = [&](int D) {
        double s = 0;
        for (int b = 0; b < LEN_BINS; b++)
            if (b >= D || b == LEN_BINS - 1) s += hist[b];
        return s;
    };
    double best = 1e300;
    int bestD = 1;
    for (int P = 1; P <= 8; P++) {
        int D = ((8 * P < 63 ? 8 * P : 63) - 1) / Db;
        if (D < 1) continue;
        double r1 = rem(D);
        int Bg = bit_width((u64)r1 + 1);
        int D2 = (62 - Bg) / Db;
        int P2 = (Bg + D2 * Db + 1 + 7) / 8;
        double r2 = rem(D + D2);
        // per item and round: the radix passes plus about two passes' worth of
        // key building and relabelling
        double cost = (double)m1 * (P + 2) + r1 * (P2 + 2) + r2 * 10;
        if (cost < best) {
            best = cost;
            bestD = D;
        }
    }
    return bestD;
}

// Rounds after which a quarter of the substrings still tied means long
// repeated substrings (induced sorting is then the cheaper way).
constexpr int LMS_TIED_CHECK_ROUND = 4;

// Names of the m1 LMS substrings of a text of length m (LMS positions
// lms_text, ascending): reduced[r] = name of the r-th, tails[c] = number of
// substrings with name <= c, *k = number of names.  Returns false (outputs
// undefined) if ties remain after max_rounds, or if more than a quarter of
// the substrings are still tied after LMS_TIED_CHECK_ROUND rounds.
template <class RF>
inline bool name_lms_direct(Dev &d, RF rk, u64 sigma, u64 m, const u32 *lms_text, u32 m1,
                            u32 *reduced, u32 *tails, u32 *k, int max_rounds) {
    if (m1 == 0) {
        *k = 0;
        return true;
    }
    Arena::Mark mk = d.a->mark();
    const int Db = bit_width(sigma + 1);
    u32 *cnt = d.a->get<u32>(LEN_BINS + 2);
    d.zero(cnt, (LEN_BINS + 2) * sizeof(u32));
    {
        u64 blocks = std::min<u64>(ceil_div(m1, BLOCK), 4096);
        k_lms_lens<<<(u32)blocks, BLOCK, 0, d.s>>>(lms_text, m1, m, cnt);
        PSK_LAUNCH_CHECK();
    }
    u32 hist[LEN_BINS];
    d.read(hist, cnt, sizeof(hist));
    int D = lms_first_digits(hist, m1, Db);

    u32 *act = d.a->get<u32>(m1), *vt = d.a->get<u32>(m1);
    u32 *gid = d.a->get<u32>(m1), *gid2 = d.a->get<u32>(m1);
    u32 *glab = d.a->get<u32>((u64)m1 / 2 + 1), *glab2 = d.a->get<u32>((u64)m1 / 2 + 1);
    u64 *keys = d.a->get<u64>(m1), *kt = d.a->get<u64>(m1);
    u32 *L = d.a->get<u32>(m1), *F = d.a->get<u32>(m1);
    u32 *err = d.err;
    d.zero(F, (u64)m1 * sizeof(u32));
    for_each(d, m1, [=] __device__(u64 i) { act[i] = (u32)i; });

    u32 A = m1;
    u64 off = 0;  // digits compared so far by every active substring
    int Bg = 0;
    int rounds = 0;
    const u64 term = sigma + 1, mm = m;
    const u32 mm1 = m1;
    while (A > 0) {
        if (rounds == max_rounds || (rounds >= LMS_TIED_CHECK_ROUND && A > m1 / 4)) {
            d.a->reset(mk);
            return false;
        }
        rounds++;
        const bool first = rounds == 1;
        if (!first) D = (62 - Bg) / Db;
        const int LB = D * Db + 1;  // digits + finished bit; keys keep <= 63 bits
        const u64 o = off;
        const int DD = D, DB = Db;
        const u32 *ac = act, *gi = gid;
        for_each(d, A, [=] __device__(u64 i) {
            u32 r = ac[i];
            u64 p = lms_text[r];
            bool last = r + 1 == mm1;
            u64 len = last ? mm - p : (u64)lms_text[r + 1] - p + 1;
            u64 td = last ? 0ull : term;
            u64 acc = 0;
            for (int j = 0; j < DD; j++) {
                u64 idx = o + (u64)j;
                u64 dg = idx < len ? (u64)rk(p + idx) + 1 : (idx == len ? td : 0ull);
                acc = (acc << DB) | dg;
            }
            u64 hi = first ? 0ull : (u64)gi[i];
            keys[i] = (hi << LB) | (acc << 1) | (len < o + (u64)DD ? 1ull : 0ull);
        });
        bool sw = radix_sort_kv(d, A, keys, act, kt, vt, Bg + LB);
        const u64 *K = sw ? kt : keys;
        const u32 *V = sw ? vt : act;
        u32 *nxt = sw ? act : vt;
        {
            Arena::Mark m3 = d.a->mark();
            const u32 nt = ceil_div(A, REL_TILE);
            u32 *tg = d.a->get<u32>(nt), *ts = d.a->get<u32>(nt);
            u32 *tk = d.a->get<u32>(nt), *te = d.a->get<u32>(nt);
            u32 *tq = d.a->get<u32>(nt);
            RelIn in{K, A, LB, true};
            k_rel_reduce<<<nt, BLOCK, 0, d.s>>>(in, tg, ts, tk, te, tq);
            PSK_LAUNCH_CHECK();
            k_scan_tiles<OpMax><<<1, BLOCK, 0, d.s>>>(tg, nt, nullptr);
            k_scan_tiles<OpMax><<<1, BLOCK, 0, d.s>>>(ts, nt, nullptr);
            k_scan_tiles<OpSum><<<1, BLOCK, 0, d.s>>>(tk, nt, cnt);
            k_scan_tiles<OpSum><<<1, BLOCK, 0, d.s>>>(tq, nt, cnt + 1);
            PSK_LAUNCH_CHECK();
            int pr = d.pbegin("lms_name_relabel", 28.0 * A);
            k_rel_apply<<<nt, BLOCK, 0, d.s>>>(in, V, FinishSink{L, F}, nxt, tg, ts, tk, te, tq,
                                               false, (u64)m1, first ? nullptr : glab, gid2,
                                               glab2, err);
            PSK_LAUNCH_CHECK();
            d.pend(pr);
            d.a->reset(m3);
        }
        u32 ag[2];
        d.read(ag, cnt, sizeof(ag));
        A = ag[0];
        if (nxt != act) std::swap(act, vt);
        std::swap(gid, gid2);
        std::swap(glab, glab2);
        Bg = bit_width(ag[1] ? ag[1] - 1 : 0);
        off += (u64)D;
    }
    // slots -> names: the marked group starts, counted in slot order
    u32 *tl = tails;
    scan_exclusive(
        d, m1, [=] __device__(u64 j) { return F[j]; },
        [=] __device__(u64 j, u32 ex, u32 v) {
            if (v) {
                F[j] = ex;
                if (ex) tl[ex - 1] = (u32)j;
            }
        },
        cnt);
    const u32 kk = d.read_u32(cnt);
    for_each(d, 1, [=] __device__(u64) { tl[kk - 1] = mm1; });
    for_each(d, m1, [=] __device__(u64 r) { reduced[r] = F[L[r]]; });
    *k = kk;
    d.a->reset(mk);
    return true;
}

}  // namespace psk
