/*
 * psk_par.c -- multi-threaded CPU restatement of pSACAK: the timed CPU
 * baseline of bench.py.
 *
 * TEST AND BENCH INFRASTRUCTURE ONLY, like psk_oracle.c: only tests/ and the
 * cpu_baseline leg of bench.py load it; the product path never calls it.
 *
 * The levels are the serial oracle's (psacak8.rs:19-53, psacak32.rs:24-56);
 * what this file adds is the crate's parallel structure, restated:
 *   par_induce_sort (u8)  psacak8.rs:201-344, RBuf :408-462, WBuf :464-498
 *   par_induce_sort (u32) psacak32.rs:289-484
 *   Pipeline              pipeline.rs:20-358
 *   parallel naming       naming.rs:97-259 (compare adjacent LMS substrings
 *                         in parallel, then names by a prefix count)
 * The SA is scanned in blocks of 128Ki entries (psacak8.rs:12 BLOCK_SIZE),
 * in parallel from 2 blocks up (psacak8.rs:15).  While the main thread
 * induces block b, the other threads fetch, for every entry x of block b+1,
 * the symbols it needs (the read buffer), and write out the entries that
 * block b-1 induced beyond block b (the write buffer).  The induce uses a
 * fetched entry only if the slot still holds the x that was fetched (an
 * entry written into the block after the fetch is read directly: what is
 * fetched is a function of x alone), writes destinations in the current or
 * the next block at once and defers the others (psacak8.rs:255-273).  The
 * integer levels keep one bucket pointer per symbol value in a side array
 * (as the serial oracle does) instead of the crate's in-SA counters: same
 * SA.  Typing, bucket counts, compaction and the reduced-string translation
 * run on all threads too (chunked, with run fix-ups at chunk ends).
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "psk_oracle.h"

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

#define PBLK (1u << 17)
#define EMPTY32 0xFFFFFFFFu

static inline u32 ld(const u32 *p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }
static inline void st(u32 *p, u32 v) { __atomic_store_n(p, v, __ATOMIC_RELAXED); }

/* ---- chunked helpers ----------------------------------------------------- */
static inline void chunk_of(u32 n, int k, int nk, u32 *lo, u32 *hi) {
    *lo = (u32)((u64)n * (u64)k / (u64)nk);
    *hi = (u32)((u64)n * (u64)(k + 1) / (u64)nk);
}

/* exclusive prefix over per-chunk counts: cnt[k] becomes chunk k's offset */
static u32 prefix_counts(u32 *cnt, int nk) {
    u32 s = 0;
    for (int k = 0; k < nk; k++) {
        u32 c = cnt[k];
        cnt[k] = s;
        s += c;
    }
    return s;
}

static void iota_par(u32 *p, u32 m) {
#pragma omp parallel for schedule(static)
    for (u32 v = 0; v < m; v++) p[v] = v;
}

static void fill_par(u32 *p, u32 m, u32 v) {
#pragma omp parallel for schedule(static)
    for (u32 i = 0; i < m; i++) p[i] = v;
}

/* L/S/LMS types (common.rs:43-68) in parallel: the type of i is fixed by the
 * first j > i with T[j] != T[i] (S iff T[j] > T[i], L if none), so a chunk
 * resolves its last run by looking ahead in the text, then scans right to
 * left on its own.  The LMS bit of a chunk's first entry reads the type of
 * the previous chunk's last one, taken before any LMS bit is written. */
#define PAR_TYPES_BODY(T)                                                    \
    if (n == 0) return;                                                      \
    _Pragma("omp parallel")                                                  \
    {                                                                        \
        const int k = omp_get_thread_num(), nk = omp_get_num_threads();      \
        u32 lo, hi;                                                          \
        chunk_of(n, k, nk, &lo, &hi);                                        \
        if (lo < hi) {                                                       \
            u32 j = hi;                                                      \
            while (j < n && T[j] == T[hi - 1]) j++;                          \
            int s = j < n && T[j] > T[hi - 1];                               \
            out[hi - 1] = (u8)s;                                             \
            for (u32 i = hi - 1; i-- > lo;) {                                \
                s = (T[i] < T[i + 1]) || (T[i] == T[i + 1] && s);            \
                out[i] = (u8)s;                                              \
            }                                                                \
        }                                                                    \
        _Pragma("omp barrier")                                               \
        int prev = lo > 0 && lo < hi ? out[lo - 1] & 1 : 1;                  \
        _Pragma("omp barrier")                                               \
        for (u32 i = lo; i < hi; i++) {                                      \
            int s = out[i] & 1;                                              \
            if (s && !prev) out[i] |= 2;                                     \
            prev = s;                                                        \
        }                                                                    \
    }

static void par_types8(const u8 *t, u32 n, u8 *out) { PAR_TYPES_BODY(t) }
static void par_types32(const u32 *t, u32 n, u8 *out) { PAR_TYPES_BODY(t) }

/* LMS positions in text order (position 0 is never LMS) */
static u32 par_lms_positions(const u8 *ty, u32 n, u32 *out) {
    const int P = omp_get_max_threads();
    u32 *cnt = (u32 *)calloc((size_t)P, sizeof(u32));
#pragma omp parallel for schedule(static, 1)
    for (int k = 0; k < P; k++) {
        u32 lo, hi, c = 0;
        chunk_of(n, k, P, &lo, &hi);
        for (u32 i = lo; i < hi; i++) c += (ty[i] >> 1) & 1;
        cnt[k] = c;
    }
    u32 total = prefix_counts(cnt, P);
#pragma omp parallel for schedule(static, 1)
    for (int k = 0; k < P; k++) {
        u32 lo, hi, o = cnt[k];
        chunk_of(n, k, P, &lo, &hi);
        for (u32 i = lo; i < hi; i++)
            if (ty[i] & 2) out[o++] = i;
    }
    free(cnt);
    return total;
}

/* the entries of sa[0..n) that KEEP, in order, to out; returns their count */
#define PAR_COMPACT(KEEP)                                                    \
    const int P = omp_get_max_threads();                                     \
    u32 *cnt = (u32 *)calloc((size_t)P, sizeof(u32));                        \
    _Pragma("omp parallel for schedule(static, 1)")                          \
    for (int k = 0; k < P; k++) {                                            \
        u32 lo, hi, c = 0;                                                   \
        chunk_of(n, k, P, &lo, &hi);                                         \
        for (u32 i = lo; i < hi; i++) { u32 x = sa[i]; c += (KEEP); }        \
        cnt[k] = c;                                                          \
    }                                                                        \
    u32 total = prefix_counts(cnt, P);                                       \
    _Pragma("omp parallel for schedule(static, 1)")                          \
    for (int k = 0; k < P; k++) {                                            \
        u32 lo, hi, o = cnt[k];                                              \
        chunk_of(n, k, P, &lo, &hi);                                         \
        for (u32 i = lo; i < hi; i++) { u32 x = sa[i]; if (KEEP) out[o++] = x; } \
    }                                                                        \
    free(cnt);                                                               \
    return total;

/* compact_left (common.rs:123-133): nonzero entries */
static u32 par_compact_nonzero(const u32 *sa, u32 n, u32 *out) { PAR_COMPACT(x != 0) }
/* LMS entries of an integer level's stage-1 SA */
static u32 par_compact_lms(const u32 *sa, u32 n, const u8 *ty, u32 *out) {
    PAR_COMPACT(x != EMPTY32 && (ty[x] & 2))
}

/* ---- the pipelined induce ------------------------------------------------ */
typedef struct {
    u32 x;    /* the slot value the entry was fetched for */
    u32 c1;   /* T[x-1] */
    u32 aux;  /* u8 levels: T[x]; u32 levels: type bits of x-1 */
} Fetch;

typedef struct {
    u32 pos, val;
} Defer;

typedef struct {
    Fetch *cache[2];
    Defer *wb[2];
} ScanBufs;

static void bufs_init(ScanBufs *b) {
    b->cache[0] = (Fetch *)malloc(sizeof(Fetch) * PBLK);
    b->cache[1] = (Fetch *)malloc(sizeof(Fetch) * PBLK);
    b->wb[0] = (Defer *)malloc(sizeof(Defer) * PBLK);
    b->wb[1] = (Defer *)malloc(sizeof(Defer) * PBLK);
}

static void bufs_free(ScanBufs *b) {
    free(b->cache[0]); free(b->cache[1]); free(b->wb[0]); free(b->wb[1]);
}

static inline void blk_range(u32 n, int forward, u32 b, u32 *lo, u32 *hi) {
    u32 a = b * PBLK, e = a + PBLK < n ? a + PBLK : n;
    if (forward) { *lo = a; *hi = e; }
    else { *lo = n - e; *hi = n - a; }
}

/* One scan over the SA in blocks (pipeline.rs:33-72 induce_outer): FETCH(c,
 * i, f) fills the read buffer entry of slot i, STEP(c, i, f, pos, val, hit)
 * induces from slot i.  Instantiated per symbol width and direction so the
 * main thread's loop is inline code.  Induced positions always lie ahead of
 * the scan (L: above i, S: below i), so a deferred write lands beyond the
 * next block and is flushed while that block is induced. */
#define DEFINE_SCAN(NAME, FORWARD, CTX_T, FETCH, STEP)                               \
    static void NAME(CTX_T *c, u32 n, ScanBufs *B) {                                 \
        const u32 nb = (n + PBLK - 1) / PBLK;                                        \
        u32 wn[2] = {0, 0};                                                          \
        const int P = n >= 2 * PBLK ? omp_get_max_threads() : 1;                     \
        {                                                                            \
            u32 lo, hi;                                                              \
            blk_range(n, FORWARD, 0, &lo, &hi);                                      \
            for (u32 i = lo; i < hi; i++) FETCH(c, i, &B->cache[0][i - lo]);         \
        }                                                                            \
        for (u32 b = 0; b < nb; b++) {                                               \
            u32 lo, hi, nlo = 0, nhi = 0;                                            \
            blk_range(n, FORWARD, b, &lo, &hi);                                      \
            if (b + 1 < nb) blk_range(n, FORWARD, b + 1, &nlo, &nhi);                \
            u32 dlo = lo, dhi = hi;                                                  \
            if (b + 1 < nb) { if (FORWARD) dhi = nhi; else dlo = nlo; }              \
            Fetch *cur = B->cache[b & 1], *nxt = B->cache[(b + 1) & 1];              \
            Defer *out = B->wb[b & 1], *prev = B->wb[(b + 1) & 1];                   \
            const u32 nprev = wn[(b + 1) & 1];                                       \
            u32 nout = 0;                                                            \
            _Pragma("omp parallel num_threads(P)")                                   \
            {                                                                        \
                const int tid = omp_get_thread_num(), nt = omp_get_num_threads();    \
                if (tid == 0) {                                                      \
                    for (u32 q = 0; q < hi - lo; q++) {                              \
                        const u32 i = FORWARD ? lo + q : hi - 1 - q;                 \
                        u32 pos = 0, val = 0;                                        \
                        int hit;                                                     \
                        STEP(c, i, &cur[i - lo], pos, val, hit);                     \
                        if (hit) {                                                   \
                            if (pos >= dlo && pos < dhi) st(&c->sa[pos], val);       \
                            else out[nout++] = (Defer){pos, val};                    \
                        }                                                            \
                    }                                                                \
                }                                                                    \
                if (nt == 1 || tid > 0) {                                            \
                    const u32 h = nt == 1 ? 0 : (u32)tid - 1;                        \
                    const u32 nh = nt == 1 ? 1 : (u32)nt - 1;                        \
                    for (u32 j = h; j < nprev; j += nh) st(&c->sa[prev[j].pos], prev[j].val); \
                    if (b + 1 < nb) {                                                \
                        const u32 len = nhi - nlo;                                   \
                        const u32 a = (u32)((u64)len * h / nh);                      \
                        const u32 e = (u32)((u64)len * (h + 1) / nh);                \
                        for (u32 i = nlo + a; i < nlo + e; i++) FETCH(c, i, &nxt[i - nlo]); \
                    }                                                                \
                }                                                                    \
            }                                                                        \
            wn[b & 1] = nout;                                                        \
            wn[(b + 1) & 1] = 0;                                                     \
        }                                                                            \
        {                                                                            \
            const u32 last = (nb - 1) & 1;                                           \
            for (u32 j = 0; j < wn[last]; j++) st(&c->sa[B->wb[last][j].pos], B->wb[last][j].val); \
        }                                                                            \
    }

/* byte levels: nonpar_induce_sort psacak8.rs:127-190 semantics (0 = empty) */
typedef struct {
    const u8 *t;
    u32 *sa;
    u32 ptr[256];
    int left_most;
} Ctx8;

#define FETCH8(c, i, f)                                                      \
    do {                                                                     \
        u32 x_ = ld(&(c)->sa[i]);                                            \
        (f)->x = x_;                                                         \
        if (x_ > 0) { (f)->c1 = (c)->t[x_ - 1]; (f)->aux = (c)->t[x_]; }     \
    } while (0)

/* L scan: the predecessor of x is L-type iff T[x-1] >= T[x] */
#define STEPL8(c, i, f, pos, val, hit)                                       \
    do {                                                                     \
        u32 x_ = ld(&(c)->sa[i]);                                            \
        hit = 0;                                                             \
        if (x_ > 0) {                                                        \
            u32 c1_, c0_;                                                    \
            if ((f)->x == x_) { c1_ = (f)->c1; c0_ = (f)->aux; }             \
            else { c1_ = (c)->t[x_ - 1]; c0_ = (c)->t[x_]; }                 \
            if (c1_ >= c0_) {                                                \
                pos = (c)->ptr[c1_]++;                                       \
                val = x_ - 1;                                                \
                hit = 1;                                                     \
                if ((c)->left_most) st(&(c)->sa[i], 0);                      \
            }                                                                \
        }                                                                    \
    } while (0)

/* S scan: S iff T[x-1] < T[x], or equal and x lies in the S part (p <= i) */
#define STEPS8(c, i, f, pos, val, hit)                                       \
    do {                                                                     \
        u32 x_ = ld(&(c)->sa[i]);                                            \
        hit = 0;                                                             \
        if (x_ > 0) {                                                        \
            u32 c1_, c0_;                                                    \
            if ((f)->x == x_) { c1_ = (f)->c1; c0_ = (f)->aux; }             \
            else { c1_ = (c)->t[x_ - 1]; c0_ = (c)->t[x_]; }                 \
            if (c1_ <= c0_ && (c)->ptr[c1_] <= (i)) {                        \
                pos = --(c)->ptr[c1_];                                       \
                val = x_ - 1;                                                \
                hit = 1;                                                     \
                if ((c)->left_most) st(&(c)->sa[i], 0);                      \
            }                                                                \
        }                                                                    \
    } while (0)

DEFINE_SCAN(scan_l8, 1, Ctx8, FETCH8, STEPL8)
DEFINE_SCAN(scan_s8, 0, Ctx8, FETCH8, STEPS8)

static void par_induce8(const u8 *t, u32 n, u32 *sa, const u32 *bounds, int left_most,
                        ScanBufs *B) {
    Ctx8 c;
    c.t = t;
    c.sa = sa;
    c.left_most = left_most;
    memcpy(c.ptr, bounds, sizeof(c.ptr)); /* heads */
    sa[c.ptr[t[n - 1]]++] = n - 1;        /* the sentinel induces n-1 first */
    scan_l8(&c, n, B);
    memcpy(c.ptr, bounds + 1, sizeof(c.ptr)); /* tails */
    scan_s8(&c, n, B);
}

/* integer levels on bucket-pointer texts (psacak32.rs:22) */
typedef struct {
    const u32 *t;
    const u8 *ty;
    u32 *sa;
    u32 *ptr;
} Ctx32;

#define FETCH32(c, i, f)                                                     \
    do {                                                                     \
        u32 x_ = ld(&(c)->sa[i]);                                            \
        (f)->x = x_;                                                         \
        if (x_ != EMPTY32 && x_ > 0) {                                       \
            (f)->c1 = (c)->t[x_ - 1];                                        \
            (f)->aux = (c)->ty[x_ - 1];                                      \
        }                                                                    \
    } while (0)

#define STEP32(c, i, f, pos, val, hit, WANT_S)                               \
    do {                                                                     \
        u32 x_ = ld(&(c)->sa[i]);                                            \
        hit = 0;                                                             \
        if (x_ != EMPTY32 && x_ > 0) {                                       \
            u32 c1_, ty_;                                                    \
            if ((f)->x == x_) { c1_ = (f)->c1; ty_ = (f)->aux; }             \
            else { c1_ = (c)->t[x_ - 1]; ty_ = (c)->ty[x_ - 1]; }            \
            if ((ty_ & 1u) == (WANT_S)) {                                    \
                pos = (WANT_S) ? (c)->ptr[c1_]-- : (c)->ptr[c1_]++;          \
                val = x_ - 1;                                                \
                hit = 1;                                                     \
            }                                                                \
        }                                                                    \
    } while (0)

#define STEPL32(c, i, f, pos, val, hit) STEP32(c, i, f, pos, val, hit, 0u)
#define STEPS32(c, i, f, pos, val, hit) STEP32(c, i, f, pos, val, hit, 1u)

DEFINE_SCAN(scan_l32, 1, Ctx32, FETCH32, STEPL32)
DEFINE_SCAN(scan_s32, 0, Ctx32, FETCH32, STEPS32)

static void par_induce32(const u32 *t, u32 m, u32 *sa, const u8 *ty, u32 *ptr, ScanBufs *B) {
    Ctx32 c = {t, ty, sa, ptr};
    iota_par(ptr, m);
    sa[ptr[t[m - 1]]++] = m - 1;
    scan_l32(&c, m, B);
    iota_par(ptr, m);
    scan_s32(&c, m, B);
}

/* ---- naming (naming.rs:97-165, 169-259) and translate (:263-274) ---------- */
/* LMS-substring length as the serial oracle (naming.rs:277-306) */
#define LMSLEN_BODY(T)                                                       \
    u32 k = 1;                                                               \
    while (i + k < n && T[i + k] >= T[i + k - 1]) k++;                       \
    while (i + k < n && T[i + k] <= T[i + k - 1]) k++;                       \
    if (i + k == n) { *sent = 1; return k + 1; }                             \
    while (k > 0 && T[i + k - 1] == T[i + k - 2]) k--;                       \
    *sent = 0;                                                               \
    return k;

static inline u32 lmslen8(const u8 *t, u32 n, u32 i, int *sent) { LMSLEN_BODY(t) }
static inline u32 lmslen32(const u32 *t, u32 n, u32 i, int *sent) { LMSLEN_BODY(t) }

/* lms: the sorted LMS positions (m1 >= 1 of them); lmspos: LMS positions in
 * text order.  Writes the bucket-pointer reduced string to reduced[0..m1)
 * and returns the number of names.  name_at holds n/2+1 words, tails and
 * flag m1 words each. */
#define PAR_NAME_BODY(T, LMSLEN, ELT)                                                     \
    const int P = omp_get_max_threads();                                                  \
    u32 *cnt = (u32 *)calloc((size_t)P, sizeof(u32));                                     \
    u32 k_total = 0;                                                                      \
    _Pragma("omp parallel num_threads(P)")                                                \
    {                                                                                     \
        const int q = omp_get_thread_num(), nq = omp_get_num_threads();                   \
        u32 lo, hi;                                                                       \
        chunk_of(m1, q, nq, &lo, &hi);                                                    \
        /* compare adjacent substrings of the chunk (naming.rs:107-126) */                \
        u32 c = 0;                                                                        \
        if (lo < hi) {                                                                    \
            int s0 = 1;                                                                   \
            u32 a = 0, la = 0;                                                            \
            if (lo > 0) { a = lms[lo - 1]; la = LMSLEN(T, n, a, &s0); }                   \
            for (u32 j = lo; j < hi; j++) {                                               \
                int s1;                                                                   \
                u32 b = lms[j], lb = LMSLEN(T, n, b, &s1);                                \
                int eq = j > 0 && !s0 && !s1 && la == lb &&                               \
                         memcmp(T + a, T + b, (size_t)la * sizeof(ELT)) == 0;             \
                flag[j] = !eq;                                                            \
                c += !eq;                                                                 \
                a = b; la = lb; s0 = s1;                                                  \
            }                                                                             \
        }                                                                                 \
        cnt[q] = c;                                                                       \
        _Pragma("omp barrier")                                                            \
        _Pragma("omp single")                                                             \
        k_total = prefix_counts(cnt, nq);                                                 \
        /* name of j = differences in [1, j]; tails[name] = end of its group */           \
        u32 nm = cnt[q];                                                                  \
        for (u32 j = lo; j < hi; j++) {                                                   \
            if (flag[j]) {                                                                \
                if (nm > 0) tails[nm - 1] = j;                                            \
                nm++;                                                                     \
            }                                                                             \
            name_at[lms[j] / 2] = nm - 1;                                                 \
        }                                                                                 \
        _Pragma("omp barrier")                                                            \
        _Pragma("omp single")                                                             \
        tails[k_total - 1] = m1;                                                          \
        _Pragma("omp for schedule(static)")                                               \
        for (u32 r = 0; r < m1; r++) reduced[r] = name_at[lmspos[r] / 2];                 \
        /* types of the reduced string (as PAR_TYPES_BODY) into flag */                   \
        chunk_of(m1, q, nq, &lo, &hi);                                                    \
        if (lo < hi) {                                                                    \
            u32 j = hi;                                                                   \
            while (j < m1 && reduced[j] == reduced[hi - 1]) j++;                          \
            int s = j < m1 && reduced[j] > reduced[hi - 1];                               \
            u32 nxt_ = reduced[hi - 1];                                                   \
            flag[hi - 1] = (u32)s;                                                        \
            for (u32 i = hi - 1; i-- > lo;) {                                             \
                u32 c0 = reduced[i];                                                      \
                s = (c0 < nxt_) || (c0 == nxt_ && s);                                     \
                flag[i] = (u32)s;                                                         \
                nxt_ = c0;                                                                \
            }                                                                             \
        }                                                                                 \
        _Pragma("omp barrier")                                                            \
        /* translate: L -> bucket head, S -> bucket tail - 1 */                           \
        for (u32 i = lo; i < hi; i++) {                                                   \
            u32 c0 = reduced[i];                                                          \
            reduced[i] = flag[i] ? tails[c0] - 1 : (c0 == 0 ? 0 : tails[c0 - 1]);         \
        }                                                                                 \
    }                                                                                     \
    free(cnt);                                                                            \
    return k_total;

static u32 par_name8(const u8 *t, u32 n, const u32 *lms, const u32 *lmspos, u32 m1,
                     u32 *reduced, u32 *name_at, u32 *tails, u32 *flag) {
    PAR_NAME_BODY(t, lmslen8, u8)
}

static u32 par_name32(const u32 *t, u32 n, const u32 *lms, const u32 *lmspos, u32 m1,
                      u32 *reduced, u32 *name_at, u32 *tails, u32 *flag) {
    PAR_NAME_BODY(t, lmslen32, u32)
}

/* ---- levels ---------------------------------------------------------------- */
static int par_sacak32(const u32 *t, u32 n, u32 *sa, ScanBufs *B);

/* The reduced problem of a level (text t of n symbols, types ty, sorted LMS
 * positions lms[0..m1), freed here): names, recursion or the direct inverse
 * when all names are distinct (psacak8.rs:37-47), then the sorted LMS
 * suffixes as text positions (permutate_lmssuffixes naming.rs:413-430).
 * Returns them in a new array; *rc gets the recursion's status. */
#define DEFINE_SOLVE_REDUCED(NAME, ELT, NAMEFN)                                       \
    static u32 *NAME(const ELT *t, u32 n, const u8 *ty, u32 *lms, u32 m1,             \
                     ScanBufs *B, int *rc) {                                          \
        const size_t w = (size_t)(m1 ? m1 : 1) * sizeof(u32);                         \
        u32 *lmspos = (u32 *)malloc(w), *red = (u32 *)malloc(w);                      \
        u32 *sa1 = (u32 *)malloc(w);                                                  \
        par_lms_positions(ty, n, lmspos);                                             \
        u32 k = 0;                                                                    \
        if (m1) {                                                                     \
            u32 *name_at = (u32 *)malloc(((size_t)n / 2 + 1) * sizeof(u32));          \
            u32 *tails = (u32 *)malloc(w), *flag = (u32 *)malloc(w);                  \
            k = NAMEFN(t, n, lms, lmspos, m1, red, name_at, tails, flag);             \
            free(name_at); free(tails); free(flag);                                   \
        }                                                                             \
        free(lms);                                                                    \
        *rc = 0;                                                                      \
        if (k < m1) *rc = par_sacak32(red, m1, sa1, B);                               \
        else {                                                                        \
            _Pragma("omp parallel for schedule(static)")                              \
            for (u32 i = 0; i < m1; i++) sa1[red[i]] = i;                             \
        }                                                                             \
        _Pragma("omp parallel for schedule(static)")                                  \
        for (u32 i = 0; i < m1; i++) sa1[i] = lmspos[sa1[i]];                         \
        free(lmspos); free(red);                                                      \
        return sa1;                                                                   \
    }

DEFINE_SOLVE_REDUCED(solve_reduced8, u8, par_name8)
DEFINE_SOLVE_REDUCED(solve_reduced32, u32, par_name32)

/* psacak32 (psacak32.rs:24-56) on a bucket-pointer text */
static int par_sacak32(const u32 *t, u32 n, u32 *sa, ScanBufs *B) {
    if (n == 0) return 0;
    if (n <= 3) { or_saca_naive32(t, n, sa); return 0; }
    u8 *ty = (u8 *)malloc(n);
    par_types32(t, n, ty);
    u32 *ptr = (u32 *)malloc((size_t)n * sizeof(u32));
    /* stage 1: LMS characters to bucket tails, right to left, then induce */
    fill_par(sa, n, EMPTY32);
    iota_par(ptr, n);
    for (u32 i = n - 1; i >= 1; i--)
        if (ty[i] & 2) sa[ptr[t[i]]--] = i;
    par_induce32(t, n, sa, ty, ptr, B);
    u32 *lms = (u32 *)malloc((size_t)n / 2 * sizeof(u32) + 4);
    const u32 m1 = par_compact_lms(sa, n, ty, lms);
    int rc;
    u32 *sa1 = solve_reduced32(t, n, ty, lms, m1, B, &rc);
    /* put_lmssuffixes psacak32.rs:103-121: from the largest, to tail-1 down */
    fill_par(sa, n, EMPTY32);
    iota_par(ptr, n);
    for (u32 i = m1; i-- > 0;) {
        u32 x = sa1[i];
        sa[ptr[t[x]]--] = x;
    }
    par_induce32(t, n, sa, ty, ptr, B);
    free(ty); free(ptr); free(sa1);
    return rc;
}

/* psacak8 (psacak8.rs:19-53) on OpenMP threads (OMP_NUM_THREADS).  Same SA
 * as or_sacak8.  Returns 0 on success. */
int or_par_sacak8(const uint8_t *t, uint32_t n, uint32_t *sa) {
    if (n == 0) return 0;
    if (n == 0xFFFFFFFFu) return -1;
    if (n <= 3) { or_saca_naive8(t, n, sa); return 0; }
    const int P = omp_get_max_threads();
    ScanBufs B;
    bufs_init(&B);
    u8 *ty = (u8 *)malloc(n);
    par_types8(t, n, ty);
    /* bucket bounds and per-character LMS counts (psacak8.rs:355-374), per
     * text chunk, then summed */
    u32 bounds[257], lmscnt[256];
    u32 (*hist)[256] = (u32 (*)[256])calloc((size_t)P, sizeof(u32[256]));
    u32 (*lmsoff)[256] = (u32 (*)[256])calloc((size_t)P, sizeof(u32[256]));
#pragma omp parallel for schedule(static, 1)
    for (int k = 0; k < P; k++) {
        u32 lo, hi;
        chunk_of(n, k, P, &lo, &hi);
        for (u32 i = lo; i < hi; i++) {
            hist[k][t[i]]++;
            lmsoff[k][t[i]] += (ty[i] >> 1) & 1;
        }
    }
    memset(bounds, 0, sizeof(bounds));
    memset(lmscnt, 0, sizeof(lmscnt));
    for (int k = 0; k < P; k++)
        for (int c = 0; c < 256; c++) {
            bounds[c + 1] += hist[k][c];
            lmscnt[c] += lmsoff[k][c];
        }
    for (int c = 1; c < 257; c++) bounds[c] += bounds[c - 1];
    free(hist);
    /* put_lmscharacters psacak8.rs:57-72: LMS to bucket tails, right to left;
     * text chunk k takes the slots below those of the chunks after it */
    for (int c = 0; c < 256; c++) {
        u32 tail = bounds[c + 1];
        for (int k = P - 1; k >= 0; k--) {
            u32 cnt = lmsoff[k][c];
            lmsoff[k][c] = tail;
            tail -= cnt;
        }
    }
    fill_par(sa, n, 0);
#pragma omp parallel for schedule(static, 1)
    for (int k = 0; k < P; k++) {
        u32 lo, hi;
        chunk_of(n, k, P, &lo, &hi);
        for (u32 i = hi; i-- > lo;)
            if (ty[i] & 2) sa[--lmsoff[k][t[i]]] = i;
    }
    free(lmsoff);
    par_induce8(t, n, sa, bounds, 1, &B);
    u32 *lms = (u32 *)malloc((size_t)n / 2 * sizeof(u32) + 4);
    const u32 m1 = par_compact_nonzero(sa, n, lms);
    int rc;
    u32 *sa1 = solve_reduced8(t, n, ty, lms, m1, &B, &rc);
    /* put_lmssuffixes psacak8.rs:76-87: char c's LMS block to its bucket tail */
    fill_par(sa, n, 0);
    u32 end = m1;
    for (int c = 255; c >= 0; c--) {
        u32 cnt = lmscnt[c];
        memcpy(sa + bounds[c + 1] - cnt, sa1 + end - cnt, (size_t)cnt * sizeof(u32));
        end -= cnt;
    }
    par_induce8(t, n, sa, bounds, 0, &B);
    free(sa1); free(ty);
    bufs_free(&B);
    return rc;
}

/* psacak32 (psacak32.rs:24-56) on a bucket-pointer text, on OpenMP threads.
 * Same SA as or_sacak32. */
int or_par_sacak32(const uint32_t *t, uint32_t m, uint32_t *sa) {
    if (m == 0) return 0;
    if (m > 0x7FFFFFFFu) return -1; /* psacak32.rs:25 */
    for (u32 i = 0; i < m; i++)
        if (t[i] >= m) return -2;   /* not a bucket-pointer text */
    ScanBufs B;
    bufs_init(&B);
    int rc = par_sacak32(t, m, sa, &B);
    bufs_free(&B);
    return rc;
}

int or_par_threads(void) { return omp_get_max_threads(); }

void or_par_set_threads(int p) {
    if (p > 0) omp_set_num_threads(p);
}
