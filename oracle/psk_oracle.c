/*
 * psk_oracle.c -- CPU oracle for the pSACAK suffix-array path.
 *
 * TEST INFRASTRUCTURE ONLY (see psk_oracle.h).  Serial, single-threaded,
 * written for clarity.  Citations are to hucsmn/psacak 0.1.0 sources.
 */
#define _GNU_SOURCE
#include "psk_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

#define EMPTY32 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* per-level statistics hook (only used by or_level_stats8)                  */
/* ------------------------------------------------------------------------ */
static u64 *g_stats = NULL;
static int g_levels = 0, g_max_levels = 0;

static void record_level(u64 m, u64 m1, u64 mL, u64 k, u64 sumlen) {
    if (!g_stats || g_levels >= g_max_levels) return;
    u64 *s = g_stats + 5 * g_levels++;
    s[0] = m; s[1] = m1; s[2] = mL; s[3] = k; s[4] = sumlen;
}

/* ------------------------------------------------------------------------ */
/* brute force: common.rs:13-25 (slice comparison, shorter prefix first)     */
/* ------------------------------------------------------------------------ */
struct sctx8 { const u8 *t; u32 n; };
struct sctx32 { const u32 *t; u32 n; };

static int cmp_suffix8(const void *a, const void *b, void *c) {
    const struct sctx8 *x = (const struct sctx8 *)c;
    u32 i = *(const u32 *)a, j = *(const u32 *)b;
    u32 li = x->n - i, lj = x->n - j, l = li < lj ? li : lj;
    int r = memcmp(x->t + i, x->t + j, l);
    if (r) return r;
    return li < lj ? -1 : (li > lj ? 1 : 0);
}

static int cmp_suffix32(const void *a, const void *b, void *c) {
    const struct sctx32 *x = (const struct sctx32 *)c;
    u32 i = *(const u32 *)a, j = *(const u32 *)b;
    while (i < x->n && j < x->n) {
        if (x->t[i] != x->t[j]) return x->t[i] < x->t[j] ? -1 : 1;
        i++; j++;
    }
    return (i == x->n) ? (j == x->n ? 0 : -1) : 1;
}

void or_saca_naive8(const uint8_t *t, uint32_t n, uint32_t *sa) {
    struct sctx8 c = {t, n};
    for (u32 i = 0; i < n; i++) sa[i] = i;
    qsort_r(sa, n, sizeof(u32), cmp_suffix8, &c);
}

void or_saca_naive32(const uint32_t *t, uint32_t n, uint32_t *sa) {
    struct sctx32 c = {t, n};
    for (u32 i = 0; i < n; i++) sa[i] = i;
    qsort_r(sa, n, sizeof(u32), cmp_suffix32, &c);
}

/* ------------------------------------------------------------------------ */
/* typing: common.rs:43-68.  T[n-1] is L; S(i) = T[i]<T[i+1] or (== and     */
/* S(i+1)); LMS(i) = S(i) and i>0 and L(i-1).                                */
/* ------------------------------------------------------------------------ */
#define TYPES_BODY(T)                                                        \
    if (n == 0) return;                                                      \
    int s = 0;                                                               \
    out[n - 1] = 0;                                                          \
    for (u32 i = n - 1; i-- > 0;) {                                          \
        s = (T[i] < T[i + 1]) || (T[i] == T[i + 1] && s);                    \
        out[i] = (u8)s;                                                      \
    }                                                                        \
    for (u32 i = 1; i < n; i++)                                              \
        if ((out[i] & 1) && !(out[i - 1] & 1)) out[i] |= 2;

void or_types8(const uint8_t *t, uint32_t n, uint8_t *out) { TYPES_BODY(t) }
void or_types32(const uint32_t *t, uint32_t n, uint8_t *out) { TYPES_BODY(t) }

/* psacak8.rs:355-374: histogram + LMS counts, exclusive scan into bounds. */
void or_buckets8(const uint8_t *t, uint32_t n, uint32_t *bounds, uint32_t *lmscnt) {
    u8 *ty = (u8 *)malloc(n ? n : 1);
    or_types8(t, n, ty);
    memset(bounds, 0, 257 * sizeof(u32));
    memset(lmscnt, 0, 256 * sizeof(u32));
    for (u32 i = 0; i < n; i++) {
        bounds[t[i] + 1]++;
        if (ty[i] & 2) lmscnt[t[i]]++;
    }
    for (int c = 1; c < 257; c++) bounds[c] += bounds[c - 1];
    free(ty);
}

/* ------------------------------------------------------------------------ */
/* LMS-substring length: naming.rs:277-306 (and the u128 variant :339-384).  */
/* Returns the length including the next LMS character; *sent = 1 when the  */
/* substring runs into the virtual sentinel (then it equals nothing).       */
/* ------------------------------------------------------------------------ */
#define LMSLEN_BODY(T)                                                       \
    u32 k = 1;                                                               \
    while (i + k < n && T[i + k] >= T[i + k - 1]) k++;                       \
    while (i + k < n && T[i + k] <= T[i + k - 1]) k++;                       \
    if (i + k == n) { *sent = 1; return k + 1; }                             \
    while (k > 0 && T[i + k - 1] == T[i + k - 2]) k--;                       \
    *sent = 0;                                                               \
    return k;

static u32 lmslen8(const u8 *t, u32 n, u32 i, int *sent) { LMSLEN_BODY(t) }
static u32 lmslen32(const u32 *t, u32 n, u32 i, int *sent) { LMSLEN_BODY(t) }

/* naming.rs:51-93 (serial naming) + translate naming.rs:263-274. */
#define NAME_BODY(T, LMSLEN, ELT)                                            \
    if (m1 == 0) return 0;                                                   \
    u8 *ty = (u8 *)malloc(n);                                                \
    TYPES_OF(T, n, ty);                                                      \
    u32 *name_at = (u32 *)malloc(((size_t)n / 2 + 1) * sizeof(u32));         \
    u32 *tails = (u32 *)malloc((size_t)m1 * sizeof(u32));                    \
    u32 k = 0;                                                               \
    int s0; u32 a = lms_sorted[0], la = LMSLEN(T, n, a, &s0);                \
    name_at[a / 2] = 0;                                                      \
    for (u32 j = 1; j < m1; j++) {                                           \
        int s1; u32 b = lms_sorted[j], lb = LMSLEN(T, n, b, &s1);            \
        int eq = !s0 && !s1 && la == lb &&                                   \
                 memcmp(T + a, T + b, (size_t)la * sizeof(ELT)) == 0;        \
        if (!eq) { tails[k] = j; k++; }                                      \
        name_at[b / 2] = k;                                                  \
        a = b; la = lb; s0 = s1;                                             \
    }                                                                        \
    tails[k] = m1; k++;                                                      \
    u32 r = 0;                                                               \
    for (u32 i = 1; i < n; i++)                                              \
        if (ty[i] & 2) reduced[r++] = name_at[i / 2];                        \
    /* translate: types of the reduced string, L -> head, S -> tail-1 */     \
    {                                                                        \
        int s = 0;                                                           \
        for (u32 i = m1; i-- > 0;) {                                         \
            u32 c = reduced[i];                                              \
            if (i + 1 < m1) {                                                \
                u32 c1 = name_at_next;                                       \
                s = (c < c1) || (c == c1 && s);                              \
            } else s = 0;                                                    \
            name_at_next = c;                                                \
            reduced[i] = s ? tails[c] - 1 : (c == 0 ? 0 : tails[c - 1]);     \
        }                                                                    \
    }                                                                        \
    free(ty); free(name_at); free(tails);                                    \
    return k;

#define TYPES_OF(T, n, ty) _Generic((T), const u8 *: or_types8, const u32 *: or_types32)(T, n, ty)

uint32_t or_name8(const uint8_t *t, uint32_t n, const uint32_t *lms_sorted, uint32_t m1,
                  uint32_t *reduced) {
    u32 name_at_next = 0;
    NAME_BODY(t, lmslen8, u8)
}

uint32_t or_name32(const uint32_t *t, uint32_t n, const uint32_t *lms_sorted, uint32_t m1,
                   uint32_t *reduced) {
    u32 name_at_next = 0;
    NAME_BODY(t, lmslen32, u32)
}

static u64 sum_lmslen8(const u8 *t, u32 n, const u32 *lms, u32 m1) {
    u64 s = 0; int sent;
    for (u32 j = 0; j < m1; j++) s += lmslen8(t, n, lms[j], &sent);
    return s;
}
static u64 sum_lmslen32(const u32 *t, u32 n, const u32 *lms, u32 m1) {
    u64 s = 0; int sent;
    for (u32 j = 0; j < m1; j++) s += lmslen32(t, n, lms[j], &sent);
    return s;
}

/* ------------------------------------------------------------------------ */
/* psacak8: psacak8.rs:19-190                                                */
/* ------------------------------------------------------------------------ */

/* nonpar_induce_sort, psacak8.rs:127-190.  0 marks an empty slot. */
static void induce8(const u8 *t, u32 n, u32 *sa, const u32 *bounds, int left_most) {
    u32 ptr[256];
    memcpy(ptr, bounds, sizeof(ptr)); /* heads */
    sa[ptr[t[n - 1]]++] = n - 1;      /* the sentinel induces n-1 first */
    for (u32 i = 0; i < n; i++) {
        u32 x = sa[i];
        if (x > 0) {
            u32 j = x - 1;
            if (t[j] >= t[x]) {          /* predecessor is L-type */
                sa[ptr[t[j]]++] = j;
                if (left_most) sa[i] = 0; /* keep only LML */
            }
        }
    }
    memcpy(ptr, bounds + 1, sizeof(ptr)); /* tails */
    for (u32 i = n; i-- > 0;) {
        u32 x = sa[i];
        if (x > 0) {
            u32 j = x - 1;
            /* S iff c0 < c1, or c0 == c1 and x lies in the S part (p <= i) */
            if (t[j] <= t[x] && ptr[t[j]] <= i) {
                sa[--ptr[t[j]]] = j;
                if (left_most) sa[i] = 0; /* keep only LMS */
            }
        }
    }
}

static u32 stage1_8(const u8 *t, u32 n, const u8 *ty, const u32 *bounds, u32 *sa) {
    /* put_lmscharacters psacak8.rs:57-72: LMS, right to left, to bucket tails */
    u32 ptr[256];
    memset(sa, 0, (size_t)n * sizeof(u32));
    memcpy(ptr, bounds + 1, sizeof(ptr));
    for (u32 i = n - 1; i >= 1; i--)
        if (ty[i] & 2) sa[--ptr[t[i]]] = i;
    induce8(t, n, sa, bounds, 1);
    /* compact_left(suf, 0): common.rs:123-133 */
    u32 m1 = 0;
    for (u32 i = 0; i < n; i++)
        if (sa[i]) sa[m1++] = sa[i];
    return m1;
}

uint32_t or_stage1_8(const uint8_t *t, uint32_t n, uint32_t *lms_sorted) {
    if (n <= 3) return 0;
    u8 *ty = (u8 *)malloc(n);
    u32 bounds[257], lmscnt[256];
    or_types8(t, n, ty);
    or_buckets8(t, n, bounds, lmscnt);
    u32 *sa = (u32 *)malloc((size_t)n * sizeof(u32));
    u32 m1 = stage1_8(t, n, ty, bounds, sa);
    memcpy(lms_sorted, sa, (size_t)m1 * sizeof(u32));
    free(sa); free(ty);
    return m1;
}

int or_sacak8(const uint8_t *t, uint32_t n, uint32_t *sa) {
    if (n == 0) return 0;
    if (n == 0xFFFFFFFFu) return -1; /* lib.rs:65,79: n < u32::MAX */
    if (n <= 3) { or_saca_naive8(t, n, sa); return 0; }

    u8 *ty = (u8 *)malloc(n);
    u32 bounds[257], lmscnt[256];
    or_types8(t, n, ty);
    or_buckets8(t, n, bounds, lmscnt);

    u32 m1 = stage1_8(t, n, ty, bounds, sa);
    u32 *red = (u32 *)malloc((size_t)(m1 ? m1 : 1) * sizeof(u32));
    u32 *sa1 = (u32 *)malloc((size_t)(m1 ? m1 : 1) * sizeof(u32));
    u32 k = or_name8(t, n, sa, m1, red);
    {
        u64 mL = 0;
        for (u32 i = 0; i < n; i++) mL += !(ty[i] & 1);
        record_level(n, m1, mL, k, sum_lmslen8(t, n, sa, m1));
    }
    int rc = 0;
    if (k < m1) rc = or_sacak32(red, m1, sa1);      /* psacak8.rs:37-40 */
    else for (u32 i = 0; i < m1; i++) sa1[red[i]] = i; /* :41-47 */

    /* permutate_lmssuffixes naming.rs:413-430 */
    u32 *lmspos = red; /* reuse: LMS positions in text order */
    {
        u32 r = 0;
        for (u32 i = 1; i < n; i++)
            if (ty[i] & 2) lmspos[r++] = i;
    }
    for (u32 i = 0; i < m1; i++) sa1[i] = lmspos[sa1[i]];

    /* put_lmssuffixes psacak8.rs:76-87: char c's LMS block to its bucket tail */
    memset(sa, 0, (size_t)n * sizeof(u32));
    u32 end = m1;
    for (int c = 255; c >= 0; c--) {
        u32 cnt = lmscnt[c];
        memcpy(sa + bounds[c + 1] - cnt, sa1 + end - cnt, (size_t)cnt * sizeof(u32));
        end -= cnt;
    }
    induce8(t, n, sa, bounds, 0); /* psacak8.rs:52 */
    free(red); free(sa1); free(ty);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* psacak32: psacak32.rs:24-56, on bucket-pointer texts (psacak32.rs:22):    */
/* an L-type character is its bucket's head, an S-type one its tail - 1.    */
/* The crate keeps bucket counters inside the SA (psacak32.rs:60-99,        */
/* 487-496); this restatement keeps one pointer per character value in a    */
/* side array instead, which yields the same SA.                            */
/* ------------------------------------------------------------------------ */
static void induce32(const u32 *t, u32 m, u32 *sa, const u8 *ty, u32 *ptr) {
    for (u32 v = 0; v < m; v++) ptr[v] = v;
    sa[ptr[t[m - 1]]++] = m - 1;
    for (u32 i = 0; i < m; i++) {
        u32 x = sa[i];
        if (x != EMPTY32 && x > 0 && !(ty[x - 1] & 1)) sa[ptr[t[x - 1]]++] = x - 1;
    }
    for (u32 v = 0; v < m; v++) ptr[v] = v;
    for (u32 i = m; i-- > 0;) {
        u32 x = sa[i];
        if (x != EMPTY32 && x > 0 && (ty[x - 1] & 1)) sa[ptr[t[x - 1]]--] = x - 1;
    }
}

static u32 stage1_32(const u32 *t, u32 m, const u8 *ty, u32 *sa, u32 *ptr, u32 *lms) {
    for (u32 i = 0; i < m; i++) sa[i] = EMPTY32;
    for (u32 v = 0; v < m; v++) ptr[v] = v;
    for (u32 i = m - 1; i >= 1; i--)
        if (ty[i] & 2) sa[ptr[t[i]]--] = i;
    induce32(t, m, sa, ty, ptr);
    u32 m1 = 0;
    for (u32 i = 0; i < m; i++) {
        u32 x = sa[i];
        if (x != EMPTY32 && (ty[x] & 2)) lms[m1++] = x;
    }
    return m1;
}

uint32_t or_stage1_32(const uint32_t *t, uint32_t m, uint32_t *lms_sorted) {
    if (m <= 3) return 0;
    u8 *ty = (u8 *)malloc(m);
    or_types32(t, m, ty);
    u32 *sa = (u32 *)malloc((size_t)m * sizeof(u32));
    u32 *ptr = (u32 *)malloc((size_t)m * sizeof(u32));
    u32 m1 = stage1_32(t, m, ty, sa, ptr, lms_sorted);
    free(sa); free(ptr); free(ty);
    return m1;
}

int or_sacak32(const uint32_t *t, uint32_t m, uint32_t *sa) {
    if (m == 0) return 0;
    if (m > 0x7FFFFFFFu) return -1; /* psacak32.rs:25 */
    for (u32 i = 0; i < m; i++)
        if (t[i] >= m) return -2;   /* not a bucket-pointer text */
    if (m <= 3) { or_saca_naive32(t, m, sa); return 0; }

    u8 *ty = (u8 *)malloc(m);
    or_types32(t, m, ty);
    u32 *ptr = (u32 *)malloc((size_t)m * sizeof(u32));
    u32 *lms = (u32 *)malloc((size_t)m / 2 * sizeof(u32) + 4);
    u32 m1 = stage1_32(t, m, ty, sa, ptr, lms);

    u32 *red = (u32 *)malloc((size_t)(m1 ? m1 : 1) * sizeof(u32));
    u32 *sa1 = (u32 *)malloc((size_t)(m1 ? m1 : 1) * sizeof(u32));
    u32 k = or_name32(t, m, lms, m1, red);
    {
        u64 mL = 0;
        for (u32 i = 0; i < m; i++) mL += !(ty[i] & 1);
        record_level(m, m1, mL, k, sum_lmslen32(t, m, lms, m1));
    }
    int rc = 0;
    if (k < m1) rc = or_sacak32(red, m1, sa1);        /* psacak32.rs:40-43 */
    else for (u32 i = 0; i < m1; i++) sa1[red[i]] = i; /* :44-50 */

    /* permutate_lmssuffixes naming.rs:413-430 */
    {
        u32 r = 0;
        for (u32 i = 1; i < m; i++)
            if (ty[i] & 2) red[r++] = i;
    }
    for (u32 i = 0; i < m1; i++) sa1[i] = red[sa1[i]];

    /* put_lmssuffixes psacak32.rs:103-121: from the largest, to tail-1 down */
    for (u32 i = 0; i < m; i++) sa[i] = EMPTY32;
    for (u32 v = 0; v < m; v++) ptr[v] = v;
    for (u32 i = m1; i-- > 0;) {
        u32 x = sa1[i];
        sa[ptr[t[x]]--] = x;
    }
    induce32(t, m, sa, ty, ptr);
    free(ty); free(ptr); free(lms); free(red); free(sa1);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* translate_sample psacak32.rs:912-943                                      */
/* ------------------------------------------------------------------------ */
static int cmp_u32(const void *a, const void *b) {
    u32 x = *(const u32 *)a, y = *(const u32 *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int or_translate32(const uint32_t *t, uint32_t n, uint32_t *out) {
    if (n == 0) return 0;
    u32 *v = (u32 *)malloc((size_t)n * sizeof(u32));
    memcpy(v, t, (size_t)n * sizeof(u32));
    qsort(v, n, sizeof(u32), cmp_u32);
    u32 d = 0;
    for (u32 i = 0; i < n; i++)
        if (i == 0 || v[i] != v[i - 1]) d++;
    if (d >= n) { free(v); return 0; }
    u8 *ty = (u8 *)malloc(n);
    or_types32(t, n, ty);
    for (u32 i = 0; i < n; i++) {
        /* head = number of symbols < t[i]; tail = number of symbols <= t[i] */
        u32 lo = 0, hi = n;
        while (lo < hi) { u32 mid = lo + (hi - lo) / 2; if (v[mid] < t[i]) lo = mid + 1; else hi = mid; }
        u32 head = lo;
        hi = n;
        while (lo < hi) { u32 mid = lo + (hi - lo) / 2; if (v[mid] <= t[i]) lo = mid + 1; else hi = mid; }
        u32 tail = lo;
        out[i] = (ty[i] & 1) ? tail - 1 : head;
    }
    free(ty); free(v);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* O(n) SA checker (Burkhardt & Kaerkkaeinen 2003): a permutation whose      */
/* adjacent entries satisfy (T[a], rank[a+1]) < (T[b], rank[b+1]), with the  */
/* empty suffix ranked -1, is the suffix array.                              */
/* ------------------------------------------------------------------------ */
#define CHECK_BODY(T)                                                        \
    if (n == 0) return 0;                                                    \
    u32 *rank = (u32 *)malloc((size_t)n * sizeof(u32));                      \
    memset(rank, 0xFF, (size_t)n * sizeof(u32));                             \
    for (u32 i = 0; i < n; i++) {                                            \
        u32 x = sa[i];                                                       \
        if (x >= n || rank[x] != EMPTY32) { free(rank); return 1; }          \
        rank[x] = i;                                                         \
    }                                                                        \
    for (u32 i = 1; i < n; i++) {                                            \
        u32 a = sa[i - 1], b = sa[i];                                        \
        if (T[a] > T[b]) { free(rank); return 2; }                           \
        if (T[a] == T[b]) {                                                  \
            int64_t ra = a + 1 < n ? (int64_t)rank[a + 1] : -1;              \
            int64_t rb = b + 1 < n ? (int64_t)rank[b + 1] : -1;              \
            if (ra >= rb) { free(rank); return 2; }                          \
        }                                                                    \
    }                                                                        \
    free(rank);                                                              \
    return 0;

int or_check_sa8(const uint8_t *t, uint32_t n, const uint32_t *sa) { CHECK_BODY(t) }
int or_check_sa32(const uint32_t *t, uint32_t n, const uint32_t *sa) { CHECK_BODY(t) }

int or_level_stats8(const uint8_t *t, uint32_t n, uint64_t *stats, int max_levels) {
    u32 *sa = (u32 *)malloc((size_t)(n ? n : 1) * sizeof(u32));
    g_stats = stats; g_levels = 0; g_max_levels = max_levels;
    or_sacak8(t, n, sa);
    int lv = g_levels;
    g_stats = NULL;
    free(sa);
    return lv;
}
