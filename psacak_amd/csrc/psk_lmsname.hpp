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
// Round 1 puts the substrings into groups by their first D digits, either by
// sorting all m1 keys (name_lms_direct) or, when the keys take few distinct
// values, through a hash dictionary of the distinct keys (name_lms_dict):
// only the distinct keys are sorted, and the substrings are counted per key.
// Later rounds sort the tied, unfinished substrings by (dense id of their tie
// group, next D digits, finished bit) with the LSD radix sort and split the
// groups; a substring is finished once its terminator has been compared or
// it is alone in its group.  Finished substrings record the slot of their
// group in the final order; a scan over the group-start marks turns slots
// into names.  Long repeated substrings need many rounds: after max_rounds
// the caller falls back to induced sorting.
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

// Texts of a batch (concatenated, each with its own sentinel): EB marks every
// text's last position, EBR[w] counts the marks in EB words before w.  The
// LMS substring that reaches its text's last position runs into that text's
// sentinel, a symbol below every other; the sentinels of the batch rank in
// text order.  Null EB: one text.
struct LmsEnds {
    const u32 *EB = nullptr;
    const u32 *EBR = nullptr;
    u32 *flag = nullptr;  // set when a key cannot hold its text's index
};

// Length of the r-th LMS substring (symbols, terminal included).  isend: it
// runs into a sentinel instead of ending at the next LMS position (the last
// substring of a text); tid: then the index of its text.
__device__ __forceinline__ u64 lms_span(const u32 *lms_text, u32 m1, u64 m, u32 r,
                                        const LmsEnds &E, bool &isend, u32 &tid) {
    const u64 p = lms_text[r];
    const bool last = r + 1 == m1;
    const u64 q = last ? m : (u64)lms_text[r + 1];
    tid = 0;
    if (E.EB) {
        // first text end at or after p, before q (a text end is L-type, so it
        // is never q itself); the last text ends at m - 1
        u64 w = p >> 5;
        const u64 wl = (q - 1) >> 5;
        u32 x = E.EB[w] & (0xFFFFFFFFu << (p & 31));
        while (!x && w < wl) x = E.EB[++w];
        if (x) {
            const u64 e = w * 32 + (u64)(__ffs(x) - 1);
            if (e < q) {
                isend = true;
                tid = E.EBR[w] + (u32)__popc(E.EB[w] & ((1u << (e & 31)) - 1u));
                return e - p + 1;
            }
        }
        isend = false;
        return q - p + 1;
    }
    isend = last;
    return last ? m - p : q - p + 1;
}

// A substring that runs into a sentinel is unique: among substrings with the
// same symbols the sentinels' order (text order) decides.  The digits after
// the sentinel digit are zeros in every key, so the text index goes there,
// in the round whose digits [o, o+D) hold the sentinel.
__device__ __forceinline__ u64 lms_end_index(u64 acc, u64 len, u64 o, int D, int Db, u32 tid,
                                             const LmsEnds &E) {
    if (tid == 0 || len < o || len >= o + (u64)D) return acc;
    const int bits = Db * (int)(o + (u64)D - 1 - len);
    if (bits < 32 && (tid >> bits) != 0) {
        if (E.flag) atomicOr(E.flag, 1u);
        return acc;
    }
    return acc | (u64)tid;
}

// Key of the r-th LMS substring for a round that compares digits [o, o+D):
// the digits packed Db bits each (first digit highest), then the finished
// bit (the terminator, or the sentinel digit of a text's last substring, lies
// within these digits).  D*Db + 1 <= 63, so a key never has bit 63 set.
template <class RF>
__device__ __forceinline__ u64 lms_key(RF rk, const u32 *lms_text, u32 m1, u64 m, u32 r, u64 o,
                                       int D, int Db, u64 term, const LmsEnds &E) {
    bool isend;
    u32 tid;
    const u64 len = lms_span(lms_text, m1, m, r, E, isend, tid);
    const u64 p = lms_text[r];
    const u64 td = isend ? 0ull : term;
    u64 acc = 0;
    for (int j = 0; j < D; j++) {
        u64 idx = o + (u64)j;
        u64 dg = idx < len ? (u64)rk(p + idx) + 1 : (idx == len ? td : 0ull);
        acc = (acc << Db) | dg;
    }
    if (isend) acc = lms_end_index(acc, len, o, D, Db, tid, E);
    return (acc << 1) | (len < o + (u64)D ? 1ull : 0ull);
}

// Ranks of a byte text packed as key digits: the digit of position i
// (rank(T[i]) + 1) takes Db bits at bit offset i*Db of the stream P, most
// significant bit first, so the digits of consecutive positions read out of
// one 64-bit window in key order.
struct PackedRanks {
    const u64 *P;
    int Db;
    // the 64 stream bits starting at the digit of position pos
    __device__ __forceinline__ u64 window(u64 pos) const {
        const u64 b = pos * (u64)Db;
        const u64 w = b >> 6;
        const u32 sh = (u32)(b & 63u);
        u64 x = P[w] << sh;
        if (sh) x |= P[w + 1] >> (64 - sh);
        return x;
    }
    __device__ __forceinline__ u32 operator()(u64 pos) const {
        return (u32)(window(pos) >> (64 - Db)) - 1u;
    }
};

// lms_key from the packed stream: the up to D symbol digits of the window
// come out of one 64-bit read (D*Db <= 62), instead of a lookup per symbol.
__device__ __forceinline__ u64 lms_key(PackedRanks rk, const u32 *lms_text, u32 m1, u64 m, u32 r,
                                       u64 o, int D, int Db, u64 term, const LmsEnds &E) {
    bool isend;
    u32 tid;
    const u64 len = lms_span(lms_text, m1, m, r, E, isend, tid);
    const u64 p = lms_text[r];
    const u64 td = isend ? 0ull : term;
    const u64 avail = len > o ? (len - o < (u64)D ? len - o : (u64)D) : 0ull;
    u64 acc = 0;
    if (avail)
        acc = (rk.window(p + o) >> (64 - (int)avail * Db)) << (Db * (D - (int)avail));
    // the terminator (or sentinel digit) at index len, if inside, then zeros
    if (len >= o && len < o + (u64)D) acc |= td << (Db * (int)(o + (u64)D - 1 - len));
    if (isend) acc = lms_end_index(acc, len, o, D, Db, tid, E);
    return (acc << 1) | (len < o + (u64)D ? 1ull : 0ull);
}

// Packs positions [64t, 64t + 64) into the DB words P[DB*t ..]; positions at
// or past n get digit 0.  The two words after the last thread's are zeroed
// (a window read may touch them).  VEC: T is 16-byte aligned.
template <int DB, bool VEC>
__global__ void __launch_bounds__(BLOCK) k_pack_ranks(const u8 *T, u64 n, const u8 *crank, u64 *P,
                                                      u64 nt) {
    __shared__ u8 lr[256];
    for (int i = threadIdx.x; i < 256; i += BLOCK) lr[i] = crank[i];
    __syncthreads();
    const u64 t = (u64)blockIdx.x * BLOCK + threadIdx.x;
    if (t == 0) {
        P[nt * DB] = 0;
        P[nt * DB + 1] = 0;
    }
    if (t >= nt) return;
    const u64 base = t * 64;
    u32 wv[16];
    if (VEC && base + 64 <= n) {
        const uint4 *q = (const uint4 *)(T + base);
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            uint4 v = q[k];
            wv[4 * k] = v.x;
            wv[4 * k + 1] = v.y;
            wv[4 * k + 2] = v.z;
            wv[4 * k + 3] = v.w;
        }
    } else {
        #pragma unroll
        for (int k = 0; k < 16; k++) {
            u32 x = 0;
            #pragma unroll
            for (int b = 0; b < 4; b++) {
                const u64 i = base + 4 * k + b;
                if (i < n) x |= (u32)T[i] << (8 * b);
            }
            wv[k] = x;
        }
    }
    u64 out[DB];
    #pragma unroll
    for (int k = 0; k < DB; k++) out[k] = 0;
    #pragma unroll
    for (int j = 0; j < 64; j++) {
        const u64 dg = base + j < n ? (u64)lr[(wv[j >> 2] >> (8 * (j & 3))) & 255u] + 1 : 0ull;
        const int b = j * DB, wd = b >> 6, off = b & 63;
        if (off + DB <= 64) {
            out[wd] |= dg << (64 - off - DB);
        } else {
            out[wd] |= dg >> (off + DB - 64);
            out[wd + 1] |= dg << (128 - off - DB);
        }
    }
    #pragma unroll
    for (int k = 0; k < DB; k++) P[t * DB + k] = out[k];
}

// The packed digit stream of a byte text of n symbols (sigma distinct
// bytes), in the arena: ceil(n/64)*Db + 2 words.
inline PackedRanks pack_ranks(Dev &d, ByteRanks rk, u64 n, u64 sigma) {
    const int Db = bit_width(sigma + 1);
    const u64 nt = ceil_div(n, 64);
    u64 *P = d.a->get<u64>(nt * (u64)Db + 2);
    const bool vec = (((uintptr_t)rk.T) & 15u) == 0;
    const u32 g = (u32)ceil_div(nt, BLOCK);
#define PSK_PACK(DB)                                                                      \
    case DB:                                                                              \
        if (vec) k_pack_ranks<DB, true><<<g, BLOCK, 0, d.s>>>(rk.T, n, rk.crank, P, nt);  \
        else k_pack_ranks<DB, false><<<g, BLOCK, 0, d.s>>>(rk.T, n, rk.crank, P, nt);     \
        break;
    switch (Db) {
        PSK_PACK(2)
        PSK_PACK(3)
        PSK_PACK(4)
        PSK_PACK(5)
        PSK_PACK(6)
        PSK_PACK(7)
        PSK_PACK(8)
        PSK_PACK(9)
        default: fail(PSK_EDEVICE, "byte alphabet out of range");
    }
#undef PSK_PACK
    PSK_LAUNCH_CHECK();
    return PackedRanks{P, Db};
}

// Working state of the refinement rounds.  act[0..A): the substrings (their
// text-order index r) still tied; gid[i]: dense id of item i's tie group,
// whose label (first slot) is glab[gid]; L[r]: slot of a finished substring;
// F[slot] = 1 at the first slot of every finished group.
struct LmsRounds {
    u32 *act, *vt, *gid, *gid2, *glab, *glab2;
    u64 *keys, *kt;
    u32 *L, *F, *cnt;
    u32 A = 0;
    u64 off = 0;  // digits compared so far by every tied substring
    int Bg = 0;   // bits of the group ids
    u32 G = 0;    // tie groups (dense ids 0..G-1 in gid)
    u32 NT = 0, LI = 0;  // tiles and items of the large ones (group_sort_kv)
    bool grouped = false;  // act holds the tie groups contiguously, in id order
    int rounds = 0;

    void alloc(Dev &d, u32 m1) {
        act = d.a->get<u32>(m1);
        vt = d.a->get<u32>(m1);
        gid = d.a->get<u32>(m1);
        gid2 = d.a->get<u32>(m1);
        glab = d.a->get<u32>((u64)m1 / 2 + 1);
        glab2 = d.a->get<u32>((u64)m1 / 2 + 1);
        keys = d.a->get<u64>(m1);
        kt = d.a->get<u64>(m1);
        L = d.a->get<u32>(m1);
        F = d.a->get<u32>(m1);
        cnt = d.a->get<u32>(6);  // the relabel's four counts, then two flags
    }
};

// Rounds until no substring is tied.  Round 1 (rounds == 0 on entry) sorts
// every substring by its first D1 digits; later rounds use the group ids.
// Returns false when the round limits are reached with ties left.
template <class RF>
inline bool lms_refine(Dev &d, LmsRounds &st, RF rk, u64 sigma, u64 m, const u32 *lms_text,
                       u32 m1, int D1, int max_rounds, LmsEnds E) {
    const int Db = bit_width(sigma + 1);
    const u64 term = sigma + 1, mm = m;
    const u32 mm1 = m1;
    u32 *err = d.err;
    while (st.A > 0) {
        if (st.rounds >= max_rounds || (st.rounds >= LMS_TIED_CHECK_ROUND && st.A > m1 / 4))
            return false;
        st.rounds++;
        const bool first = st.rounds == 1;
        const int D = first ? D1 : (62 - st.Bg) / Db;
        const int LB = D * Db + 1;  // digits + finished bit; keys keep <= 63 bits
        const u64 o = st.off;
        const int DD = D, DB = Db;
        const u32 *ac = st.act, *gi = st.gid;
        u64 *keys = st.keys;
        for_each(d, st.A, [=] __device__(u64 i) {
            u32 r = ac[i];
            u64 hi = first ? 0ull : (u64)gi[i];
            keys[i] = (hi << LB) | lms_key(rk, lms_text, mm1, mm, r, o, DD, DB, term, E);
        });
        // after round 1 the tie groups stay in place: each is sorted by its
        // new digits alone
        bool sw = !st.grouped
                      ? radix_sort_kv(d, st.A, st.keys, st.act, st.kt, st.vt, st.Bg + LB)
                      : group_sort_kv(d, st.A, st.G, st.gid, ~0u, st.keys, st.act, st.kt, st.vt,
                                      LB, st.NT, st.LI);
        const u64 *K = sw ? st.kt : st.keys;
        const u32 *V = sw ? st.vt : st.act;
        u32 *nxt = sw ? st.act : st.vt;
        {
            ArenaScope m3(*d.a);
            const u32 nt = ceil_div(st.A, REL_TILE);
            u32 *tg = d.a->get<u32>(nt), *ts = d.a->get<u32>(nt);
            u32 *tk = d.a->get<u32>(nt), *te = d.a->get<u32>(nt);
            u32 *tq = d.a->get<u32>(nt);
            RelIn in{K, st.A, LB, true};
            k_rel_reduce<<<nt, BLOCK, 0, d.s>>>(in, tg, ts, tk, te, tq);
            PSK_LAUNCH_CHECK();
            k_rel_tiles<<<4, BLOCK, 0, d.s>>>(tg, ts, tk, tq, te, nt, st.cnt);
            PSK_LAUNCH_CHECK();
            int pr = d.pbegin("lms_name_relabel", 28.0 * st.A);
            k_rel_apply<<<nt, BLOCK, 0, d.s>>>(in, V, FinishSink{st.L, st.F}, nxt, tg, ts, tk, te,
                                               tq, false, (u64)m1, first ? nullptr : st.glab,
                                               st.gid2, st.glab2, err, st.cnt);
            PSK_LAUNCH_CHECK();
            d.pend(pr);
            m3.reset();
        }
        u32 ag[4];
        d.read(ag, st.cnt, sizeof(ag));
        const u64 before = st.A;
        st.A = ag[0];
        if (nxt != st.act) std::swap(st.act, st.vt);
        std::swap(st.gid, st.gid2);
        std::swap(st.glab, st.glab2);
        trace_round(d, "lms names", st.rounds, before, st.Bg + LB, st.gid, st.A, ~0u, ag[1]);
        st.Bg = bit_width(ag[1] ? ag[1] - 1 : 0);
        st.G = ag[1];
        st.NT = ag[2];
        st.LI = ag[3];
        st.grouped = true;  // the relabel emits the kept items by group
        st.off += (u64)D;
    }
    return true;
}

// Slots -> names once nothing is tied: the marked group starts, counted in
// slot order.  tails[c] = number of substrings with name <= c; returns k.
inline u32 lms_slot_names(Dev &d, LmsRounds &st, u32 m1, u32 *tails) {
    u32 *tl = tails, *F = st.F;
    const u32 mm1 = m1;
    compact_exclusive(
        d, m1, [=] __device__(u64 j) { return F[j]; },
        [=] __device__(u64 j, u32 ex, u32 v) {
            if (v) {
                F[j] = ex;
                if (ex) tl[ex - 1] = (u32)j;
            }
        },
        st.cnt);
    const u32 kk = d.read_u32(st.cnt);
    for_each(d, 1, [=] __device__(u64) { tl[kk - 1] = mm1; });
    return kk;
}

// Names of the m1 LMS substrings of a text of length m (LMS positions
// lms_text, ascending): reduced[r] = name of the r-th, tails[c] = number of
// substrings with name <= c, *k = number of names.  Returns false (outputs
// undefined) if ties remain after max_rounds, or if more than a quarter of
// the substrings are still tied after LMS_TIED_CHECK_ROUND rounds.
template <class RF>
inline bool name_lms_direct(Dev &d, RF rk, u64 sigma, u64 m, const u32 *lms_text, u32 m1,
                            u32 *reduced, u32 *tails, u32 *k, int max_rounds,
                            LmsEnds E = LmsEnds{}) {
    if (m1 == 0) {
        *k = 0;
        return true;
    }
    ArenaScope mk(*d.a);
    const int Db = bit_width(sigma + 1);
    u32 *hc = d.a->get<u32>(LEN_BINS);
    d.zero(hc, LEN_BINS * sizeof(u32));
    {
        u64 blocks = std::min<u64>(ceil_div(m1, BLOCK), 4096);
        k_lms_lens<<<(u32)blocks, BLOCK, 0, d.s>>>(lms_text, m1, m, hc);
        PSK_LAUNCH_CHECK();
    }
    u32 hist[LEN_BINS];
    d.read(hist, hc, sizeof(hist));
    const int D = lms_first_digits(hist, m1, Db);

    LmsRounds st;
    st.alloc(d, m1);
    d.zero(st.F, (u64)m1 * sizeof(u32));
    u32 *act = st.act;
    for_each(d, m1, [=] __device__(u64 i) { act[i] = (u32)i; });
    st.A = m1;
    if (E.EB) {
        E.flag = st.cnt + 4;  // cnt[0..3] belong to the rounds
        d.zero(E.flag, sizeof(u32));
    }
    if (!lms_refine(d, st, rk, sigma, m, lms_text, m1, D, max_rounds, E) ||
        (E.EB && d.read_u32(E.flag))) {
        mk.reset();
        return false;
    }
    const u32 kk = lms_slot_names(d, st, m1, tails);
    const u32 *F = st.F, *L = st.L;
    for_each(d, m1, [=] __device__(u64 r) { reduced[r] = F[L[r]]; });
    *k = kk;
    mk.reset();
    return true;
}

// ---------------------------------------------------------------------------
// Round 1 by dictionary.  A level-0 text has few distinct LMS substrings (a
// DNA text of 2^30 bytes has under 10^4 among 3*10^8), so round 1 does not
// need to sort all m1 keys: the distinct keys are collected in a hash table,
// sorted, and each substring is counted under its key's dense rank.  A group
// (equal keys) is finished when its key is terminated or it holds one
// substring; the substrings of the other groups go on to the sorting rounds
// with their group's dense id and first slot.
// ---------------------------------------------------------------------------
constexpr u64 DICT_EMPTY = ~0ull;  // keys never have bit 63 set
#ifndef PSK_DICT_TBITS
#define PSK_DICT_TBITS 19
#endif
// insert table of 2^DICT_TBITS slots (4 MB), at most a quarter of them
// filled: a text with more distinct keys names faster by sorting (its keys
// do not fit the workgroups' LDS caches, so every substring probes the table)
constexpr int DICT_TBITS = PSK_DICT_TBITS;
constexpr u32 DICT_MAX_KEYS = 1u << (DICT_TBITS - 2);
constexpr u32 DICT_MAX_PROBES = 128;
constexpr int DICT_LDS = 8192;     // per-workgroup cache of inserted keys (two places each)
constexpr int DICT_BLOCK = 1024;   // insert and ranking workgroups (64 KB of LDS each)
constexpr u32 DICT_HIST = 16384;   // ranks per LDS histogram window (64 KB)
constexpr u32 DICT_WINDOWS = 8;    // windows before the rest is counted globally

__device__ __forceinline__ u32 dict_hash(u64 key, int bits) {
    return (u32)((key * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

// Inserts every substring's key and records the table slot that holds it in
// SLOT[r].  A workgroup remembers keys that sit in their home slot in LDS
// (each in one of two places), so a frequent key goes to the table once per
// workgroup and its slot is its hash: a wave waits for the table whenever
// one of its lanes misses, so the cache holds a whole byte text's
// dictionary.  ctr[0] counts distinct keys, ctr[1] flags a full table or
// more than DICT_MAX_KEYS distinct keys (the insert then stops early).
__global__ void __launch_bounds__(DICT_BLOCK) k_dict_insert(PackedRanks rk, const u32 *lms_text,
                                                            u32 m1, u64 m, int D, int Db, u64 term,
                                                            LmsEnds E, u64 *tkey, u32 *ctr,
                                                            u32 *SLOT) {
    __shared__ u64 seen[DICT_LDS];
    for (int i = threadIdx.x; i < DICT_LDS; i += DICT_BLOCK) seen[i] = DICT_EMPTY;
    __syncthreads();
    const u32 mask = (1u << DICT_TBITS) - 1u;
    bool abandon = false;
    u32 probes = 0;
    for (u64 r = (u64)blockIdx.x * DICT_BLOCK + threadIdx.x; r < m1;
         r += (u64)gridDim.x * DICT_BLOCK) {
        u64 key = lms_key(rk, lms_text, m1, m, (u32)r, 0, D, Db, term, E);
        const u32 home = dict_hash(key, DICT_TBITS);
        const u32 c1 = home & (DICT_LDS - 1);
        const u32 c2 = (u32)((key * 0xC2B2AE3D27D4EB4Full) >> 51) & (DICT_LDS - 1);
        if (seen[c1] == key || seen[c2] == key) {
            SLOT[r] = home;
            continue;
        }
        // an abandoned dictionary's results are discarded: every 16th trip
        // to the table reads the flag word first (a coherent load bypasses
        // the XCD's L2, so it is not made per item) and stops once it is set
        if ((probes++ & 15u) == 0 &&
            __hip_atomic_load(&ctr[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            break;
        u32 s = home;
        bool ok = false, ins = false;
        for (u32 pr = 0; pr < DICT_MAX_PROBES; pr++) {
            // a slot changes once, from empty to its key: a plain load that
            // still sees it empty is corrected by the CAS
            u64 cur = tkey[s];
            if (cur == key) { ok = true; break; }
            if (cur == DICT_EMPTY) {
                // a dictionary past its capacity is abandoned (the caller
                // sorts instead): the lane leaves (no further inserts, no
                // flag atomics), so the table stays about half full at most
                if (__hip_atomic_load(&ctr[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    abandon = true;
                    break;
                }
                u64 old = atomicCAS((unsigned long long *)&tkey[s], (unsigned long long)DICT_EMPTY,
                                    (unsigned long long)key);
                if (old == DICT_EMPTY) {
                    ins = true;
                    ok = true;
                    break;
                }
                if (old == key) { ok = true; break; }
            }
            s = (s + 1) & mask;
        }
        {
            // new keys counted once per wave (the lanes that probed): a
            // text with millions of distinct keys would otherwise send one
            // atomic per key to the same counter
            const u64 b = ballot(ins);
            if (ins && lane_id() == __ffsll((long long)b) - 1) {
                const u32 c = (u32)__popcll(b);
                if (atomicAdd(&ctr[0], c) + c > DICT_MAX_KEYS) atomicOr(&ctr[1], 4u);
            }
        }
        if (abandon) break;
        if (ok) {
            SLOT[r] = s;
            if (s == home) seen[seen[c1] == DICT_EMPTY ? c1 : c2] = key;
        } else {
            atomicOr(&ctr[1], 1u);
        }
    }
}

// Dense rank of every substring's key: SL[r] holds its key's table slot on
// entry and receives ros[slot] (the rank of the key in that slot); the number
// of substrings per rank below DICT_HIST goes to an LDS histogram per
// workgroup, written out as row blockIdx.x of part (hn = min(K1, DICT_HIST)
// words per row) and summed by k_dict_hist_sum.  A rank out of range sets
// ctr[1].
__global__ void __launch_bounds__(DICT_BLOCK) k_dict_rank(u32 *SL, u32 m1, const u32 *ros, u32 K1,
                                                          u32 *part, u32 *ctr) {
    __shared__ u32 h[DICT_HIST];
    const u32 hn = K1 < DICT_HIST ? K1 : DICT_HIST;
    for (u32 i = threadIdx.x; i < hn; i += DICT_BLOCK) h[i] = 0;
    __syncthreads();
    for (u64 r = (u64)blockIdx.x * DICT_BLOCK + threadIdx.x; r < m1;
         r += (u64)gridDim.x * DICT_BLOCK) {
        u32 dr = ros[SL[r] & ((1u << DICT_TBITS) - 1u)];
        if (dr >= K1) {
            atomicOr(&ctr[1], 2u);
            dr = 0;
        }
        SL[r] = dr;
        if (dr < DICT_HIST) atomicAdd(&h[dr], 1u);
    }
    __syncthreads();
    u32 *row = part + (u64)blockIdx.x * hn;
    for (u32 i = threadIdx.x; i < hn; i += DICT_BLOCK) row[i] = h[i];
}

// count[j] = sum of the nb rows of part at column j, for j < hn.
__global__ void k_dict_hist_sum(const u32 *part, u32 nb, u32 hn, u32 *count) {
    const u32 j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= hn) return;
    u32 acc = 0;
    for (u32 b = 0; b < nb; b++) acc += part[(u64)b * hn + j];
    count[j] = acc;
}

// Counts of the ranks in [lo, lo + DICT_HIST) from SL (LDS histogram); with
// TAIL also every rank above that window (wave-combined global atomics, for
// dictionaries too large for a few windows).
template <bool TAIL>
__global__ void __launch_bounds__(DICT_BLOCK) k_dict_count(const u32 *SL, u32 m1, u32 lo, u32 K1,
                                                           u32 *count) {
    __shared__ u32 h[DICT_HIST];
    const u32 hn = K1 - lo < DICT_HIST ? K1 - lo : DICT_HIST;
    for (u32 i = threadIdx.x; i < hn; i += DICT_BLOCK) h[i] = 0;
    __syncthreads();
    // wave-uniform trip count: every lane reaches the ballots below
    const u64 stride = (u64)gridDim.x * DICT_BLOCK;
    for (u64 b = (u64)blockIdx.x * DICT_BLOCK; b < m1; b += stride) {
        const u64 r = b + threadIdx.x;
        const u32 dr = r < m1 ? SL[r] : 0u;
        const bool in = r < m1 && dr >= lo && dr - lo < DICT_HIST;
        if (in) atomicAdd(&h[dr - lo], 1u);
        if (TAIL) {
            bool pending = r < m1 && dr >= lo && dr - lo >= DICT_HIST;
            while (true) {
                u64 act = ballot(pending);
                if (!act) break;
                int leader = __ffsll((long long)act) - 1;
                u32 dl = __shfl(dr, leader, WAVE);
                bool mine = pending && dr == dl;
                u64 peers = ballot(mine);
                if (lane_id() == leader) atomicAdd(&count[dl], (u32)__popcll(peers));
                if (mine) pending = false;
            }
        }
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < hn; i += DICT_BLOCK)
        if (h[i]) atomicAdd(&count[lo + i], h[i]);
}

// Names of a byte text's LMS substrings through the dictionary; same
// contract as name_lms_direct.  Returns false, with nothing written, if the
// keys take more than DICT_MAX_KEYS values or the rounds that follow hit
// their limits (the caller then uses another strategy).
inline bool name_lms_dict(Dev &d, PackedRanks rk, u64 sigma, u64 m, const u32 *lms_text, u32 m1,
                          u32 *reduced, u32 *tails, u32 *k, int max_rounds,
                          LmsEnds E = LmsEnds{}) {
    if (m1 == 0) {
        *k = 0;
        return true;
    }
    const int Db = bit_width(sigma + 1);
    const int D = 62 / Db;
    if (D < 1 || Db != rk.Db) return false;
    const u64 term = sigma + 1;
    ArenaScope mk(*d.a);
    const u64 tsize = 1ull << DICT_TBITS;
    u64 *tkey = d.a->get<u64>(tsize);
    u32 *ctr = d.a->get<u32>(4);
    d.zero(ctr, 4 * sizeof(u32));
    E.flag = ctr + 1;  // a text index that does not fit counts as a failure
    for_each(d, tsize, [=] __device__(u64 i) { tkey[i] = DICT_EMPTY; });
    u32 *SL = d.a->get<u32>(m1);  // table slot, then dense rank, of every substring
    {
        u64 blocks = std::min<u64>(ceil_div(m1, DICT_BLOCK), 512);
        int pr = d.pbegin("dict_insert", 16.0 * m1);
        k_dict_insert<<<(u32)blocks, DICT_BLOCK, 0, d.s>>>(rk, lms_text, m1, m, D, Db, term, E,
                                                           tkey, ctr, SL);
        PSK_LAUNCH_CHECK();
        d.pend(pr);
    }
    u32 hc[2];
    d.read(hc, ctr, sizeof(hc));
    const u32 K1 = hc[0];
    if (hc[1] || K1 > DICT_MAX_KEYS || K1 == 0) {
        mk.reset();
        return false;
    }
    // distinct keys with their slots, sorted, then the rank of each used slot
    u64 *dk = d.a->get<u64>(K1), *dk2 = d.a->get<u64>(K1);
    u32 *ds = d.a->get<u32>(K1), *ds2 = d.a->get<u32>(K1);
    scan_exclusive(
        d, tsize, [=] __device__(u64 i) { return tkey[i] != DICT_EMPTY ? 1u : 0u; },
        [=] __device__(u64 i, u32 ex, u32 v) {
            if (v) {
                dk[ex] = tkey[i];
                ds[ex] = (u32)i;
            }
        },
        nullptr);
    bool sw = radix_sort_kv(d, K1, dk, ds, dk2, ds2, 63);
    const u64 *DK = sw ? dk2 : dk;
    const u32 *DS = sw ? ds2 : ds;
    // the insert table's keys are extracted: its memory holds ros[slot] now
    u32 *ros = (u32 *)tkey;
    for_each(d, K1, [=] __device__(u64 j) { ros[DS[j]] = (u32)j; });
    // rank and count every substring
    u32 *count = d.a->get<u32>(K1 + 1);
    d.zero(count, ((u64)K1 + 1) * sizeof(u32));
    {
        int pr = d.pbegin("dict_rank", 8.0 * m1);
        u64 blocks = std::min<u64>(ceil_div(m1, DICT_BLOCK), 512);
        const u32 hn = std::min<u32>(K1, DICT_HIST);
        u32 *part = d.a->get<u32>(blocks * hn);
        k_dict_rank<<<(u32)blocks, DICT_BLOCK, 0, d.s>>>(SL, m1, ros, K1, part, ctr);
        PSK_LAUNCH_CHECK();
        k_dict_hist_sum<<<ceil_div(hn, BLOCK), BLOCK, 0, d.s>>>(part, (u32)blocks, hn, count);
        PSK_LAUNCH_CHECK();
        d.pend(pr);
        // the ranks above the first window, window by window from SL
        for (u32 w = 1; (u64)w * DICT_HIST < K1; w++) {
            const bool tail = w == DICT_WINDOWS - 1;
            if (tail)
                k_dict_count<true><<<(u32)blocks, DICT_BLOCK, 0, d.s>>>(SL, m1, w * DICT_HIST, K1,
                                                                       count);
            else
                k_dict_count<false><<<(u32)blocks, DICT_BLOCK, 0, d.s>>>(SL, m1, w * DICT_HIST, K1,
                                                                        count);
            PSK_LAUNCH_CHECK();
            if (tail) break;
        }
    }
    // per rank: first slot (exclusive scan of counts), finished?, and the
    // dense id among the unfinished groups
    u32 *start = d.a->get<u32>(K1 + 1);
    u32 *ugid = d.a->get<u32>(K1 + 1);
    u32 *ucnt = ctr + 2;  // unfinished groups
    scan_exclusive(
        d, K1, [=] __device__(u64 j) { return count[j]; },
        [=] __device__(u64 j, u32 ex, u32) { start[j] = ex; }, nullptr);
    auto unfinished = [=] __device__(u64 j) -> bool { return !(DK[j] & 1ull) && count[j] > 1; };
    scan_exclusive(
        d, K1, [=] __device__(u64 j) { return unfinished(j) ? 1u : 0u; },
        [=] __device__(u64 j, u32 ex, u32 v) { ugid[j] = v ? ex : ~0u; }, ucnt);
    u32 hc2[4];
    d.read(hc2, ctr, sizeof(hc2));
    if (hc2[1]) fail(PSK_EDEVICE, "dictionary lookup missed an inserted key");
    const u32 G = hc2[2];

    LmsRounds st;
    st.alloc(d, m1);
    d.zero(st.F, (u64)m1 * sizeof(u32));
    u32 *F = st.F;
    // finished groups start at their first slot
    for_each(d, K1, [=] __device__(u64 j) {
        if (!unfinished(j)) F[start[j]] = 1;
    });
    st.rounds = 1;
    st.off = (u64)D;
    if (G) {
        // the tied substrings, in text order, with their group's dense id;
        // a group's label is its first slot
        u32 *act = st.act, *gid = st.gid, *glab = st.glab;
        for_each(d, K1, [=] __device__(u64 j) {
            if (ugid[j] != ~0u) glab[ugid[j]] = start[j];
        });
        compact_exclusive(
            d, m1, [=] __device__(u64 r) { return ugid[SL[r]] != ~0u ? 1u : 0u; },
            [=] __device__(u64 r, u32 ex, u32 v) {
                if (v) {
                    act[ex] = (u32)r;
                    gid[ex] = ugid[SL[r]];
                }
            },
            ctr + 3);
        st.A = d.read_u32(ctr + 3);
        st.Bg = bit_width(G - 1);
        st.G = G;
        E.flag = ctr + 3;  // free again after the count of tied substrings
        d.zero(E.flag, sizeof(u32));
        if (!lms_refine(d, st, rk, sigma, m, lms_text, m1, D, max_rounds, E) ||
            (E.EB && d.read_u32(E.flag))) {
            mk.reset();
            return false;
        }
    }
    const u32 kk = lms_slot_names(d, st, m1, tails);
    // names: finished groups from their first slot's mark (a small table),
    // the refined ones from their own slot
    u32 *dname = ugid;  // reused: name per rank, ~0u for refined groups
    for_each(d, K1, [=] __device__(u64 j) { dname[j] = unfinished(j) ? ~0u : F[start[j]]; });
    const u32 *L = st.L;
    for_each(d, m1, [=] __device__(u64 r) {
        u32 v = dname[SL[r]];
        reduced[r] = v != ~0u ? v : F[L[r]];
    });
    *k = kk;
    mk.reset();
    return true;
}

}  // namespace psk
