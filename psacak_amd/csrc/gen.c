/*
 * gen.c -- seeded synthetic workload generators for psacak_amd.
 *
 * One definition, used by tests and by bench.py through ctypes.  Every
 * generator is a pure function of (seed, sizes); multithreaded parts are
 * split into fixed 1 MiB chunks, each with its own seeded stream, so the
 * output does not depend on the thread count.
 *
 * PRNG: xoshiro256** (Blackman & Vigna 2018) seeded by SplitMix64.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

#define CHUNK ((u64)1 << 20)
#define MAX_THREADS 16

static const u8 DNA[4] = {65, 67, 71, 84}; /* A C G T */

typedef struct { u64 s[4]; } rng_t;

static u64 splitmix64(u64 *x) {
    u64 z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, u64 seed, u64 stream) {
    u64 x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline u64 rotl(u64 x, int k) { return (x << k) | (x >> (64 - k)); }

static inline u64 rng_next(rng_t *r) {
    u64 *s = r->s;
    u64 res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return res;
}

/* uniform double in [0, 1) with 53 bits */
static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* uniform integer in [0, bound) (Lemire's multiply-shift, bound > 0) */
static inline u64 rng_below(rng_t *r, u64 bound) {
    return (u64)(((unsigned __int128)rng_next(r) * bound) >> 64);
}

/* ---- chunked parallel driver ------------------------------------------- */
typedef void (*chunk_fn)(void *arg, u64 chunk, u64 lo, u64 hi);
typedef struct {
    chunk_fn fn; void *arg; u64 n, nchunks; int tid, nthreads;
} job_t;

static void *job_run(void *p) {
    job_t *j = (job_t *)p;
    for (u64 c = (u64)j->tid; c < j->nchunks; c += (u64)j->nthreads) {
        u64 lo = c * CHUNK, hi = lo + CHUNK < j->n ? lo + CHUNK : j->n;
        j->fn(j->arg, c, lo, hi);
    }
    return NULL;
}

static void run_chunks(chunk_fn fn, void *arg, u64 n, int nthreads) {
    u64 nchunks = (n + CHUNK - 1) / CHUNK;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > MAX_THREADS) nthreads = MAX_THREADS;
    if ((u64)nthreads > nchunks) nthreads = (int)(nchunks ? nchunks : 1);
    pthread_t th[MAX_THREADS];
    job_t jobs[MAX_THREADS];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = (job_t){fn, arg, n, nchunks, t, nthreads};
        if (t) pthread_create(&th[t], NULL, job_run, &jobs[t]);
    }
    job_run(&jobs[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* ---- i.i.d. uniform over an alphabet ----------------------------------- */
typedef struct { u64 seed; u8 *out; u32 alphabet; int dna; } iid_arg;

static void iid_chunk(void *p, u64 c, u64 lo, u64 hi) {
    iid_arg *a = (iid_arg *)p;
    rng_t r;
    rng_seed(&r, a->seed, c + 1);
    if (a->dna) {
        for (u64 i = lo; i < hi; i++) a->out[i] = DNA[rng_next(&r) >> 62];
    } else {
        for (u64 i = lo; i < hi; i++) a->out[i] = (u8)rng_below(&r, a->alphabet);
    }
}

/* i.i.d. DNA: b_i = "ACGT"[x_i >> 62] */
void psk_gen_iid_dna(u64 seed, u64 n, u8 *out, int nthreads) {
    iid_arg a = {seed, out, 4, 1};
    run_chunks(iid_chunk, &a, n, nthreads);
}

/* i.i.d. bytes in [0, alphabet) (the shape of bench/generate.sh:7-13). */
void psk_gen_uniform8(u64 seed, u64 n, u32 alphabet, u8 *out, int nthreads) {
    iid_arg a = {seed, out, alphabet ? alphabet : 256, 0};
    run_chunks(iid_chunk, &a, n, nthreads);
}

/* ---- order-3 Markov DNA with Dirichlet(0.5) rows + repeats -------------- */
static double rng_normal(rng_t *r) {
    double u1 = rng_unit(r), u2 = rng_unit(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Marsaglia & Tsang (2000) gamma(a) for a >= 1; a < 1 via the U^(1/a) boost */
static double rng_gamma(rng_t *r, double a) {
    if (a < 1.0) {
        double u = rng_unit(r);
        return rng_gamma(r, a + 1.0) * pow(u > 1e-300 ? u : 1e-300, 1.0 / a);
    }
    double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (;;) {
        double x = rng_normal(r), v = 1.0 + c * x;
        if (v <= 0) continue;
        v = v * v * v;
        double u = rng_unit(r);
        if (log(u > 1e-300 ? u : 1e-300) < 0.5 * x * x + d - d * v + d * log(v)) return d * v;
    }
}

typedef struct { u64 seed; u8 *out; double cdf[64][4]; } markov_arg;

static void markov_chunk(void *p, u64 c, u64 lo, u64 hi) {
    markov_arg *a = (markov_arg *)p;
    rng_t r;
    rng_seed(&r, a->seed, 0x1000 + c);
    u32 ctx = (u32)rng_below(&r, 64); /* uniform initial context per chunk */
    for (u64 i = lo; i < hi; i++) {
        double u = rng_unit(&r);
        const double *row = a->cdf[ctx];
        u32 s = u < row[0] ? 0 : u < row[1] ? 1 : u < row[2] ? 2 : 3;
        a->out[i] = (u8)s;
        ctx = ((ctx << 2) | s) & 63;
    }
}

typedef struct { u64 seed; u8 *out; } map_arg;
static void map_dna_chunk(void *p, u64 c, u64 lo, u64 hi) {
    map_arg *a = (map_arg *)p;
    (void)c;
    for (u64 i = lo; i < hi; i++) a->out[i] = DNA[a->out[i] & 3];
}

/* Order-3 Markov chain over ACGT with Dirichlet(0.5) transition rows, then
 * repeats: while the copied total is below repeat_frac*n, copy a segment of
 * length floor(exp(U(ln lo, ln hi))) from a uniform earlier source to a
 * uniform destination, substituting each byte with probability 0.01 by a
 * uniformly chosen different base.  lo/hi = 1 KiB / 100 KiB, shrunk for
 * small n. */
void psk_gen_markov_dna(u64 seed, u64 n, double repeat_frac, u8 *out, int nthreads) {
    markov_arg *a = (markov_arg *)malloc(sizeof(markov_arg));
    a->seed = seed;
    a->out = out;
    rng_t r;
    rng_seed(&r, seed, 0);
    for (int ctx = 0; ctx < 64; ctx++) {
        double g[4], s = 0;
        for (int k = 0; k < 4; k++) { g[k] = rng_gamma(&r, 0.5); s += g[k]; }
        double acc = 0;
        for (int k = 0; k < 4; k++) { acc += g[k] / s; a->cdf[ctx][k] = acc; }
        a->cdf[ctx][3] = 2.0; /* guard */
    }
    run_chunks(markov_chunk, a, n, nthreads);
    /* repeats (sequential: later copies may read earlier ones) */
    double lo = 1024.0, hi = 102400.0;
    if (hi > (double)n / 8) hi = (double)n / 8;
    if (lo > hi / 4) lo = hi / 4;
    if (lo >= 2 && repeat_frac > 0) {
        rng_t q;
        rng_seed(&q, seed, 0xABCDEF);
        u64 copied = 0, target = (u64)(repeat_frac * (double)n);
        double llo = log(lo), lhi = log(hi);
        while (copied < target) {
            u64 L = (u64)floor(exp(llo + (lhi - llo) * rng_unit(&q)));
            if (L < 1) L = 1;
            if (2 * L + 1 > n) break;
            u64 dst = L + rng_below(&q, n - 2 * L + 1); /* dst in [L, n-L] */
            u64 src = rng_below(&q, dst - L + 1);       /* src in [0, dst-L] */
            for (u64 k = 0; k < L; k++) {
                u8 b = out[src + k];
                if (rng_below(&q, 100) == 0) b = (u8)((b + 1 + rng_below(&q, 3)) & 3);
                out[dst + k] = b;
            }
            copied += L;
        }
    }
    map_arg m = {seed, out};
    run_chunks(map_dna_chunk, &m, n, nthreads);
    free(a);
}

/* ---- Zipf u32 text with exact repeats ------------------------------------ */
typedef struct { u64 seed; u32 *out; const double *cdf; const u32 *perm; u32 sigma; } zipf_arg;

static void zipf_chunk(void *p, u64 c, u64 lo, u64 hi) {
    zipf_arg *a = (zipf_arg *)p;
    rng_t r;
    rng_seed(&r, a->seed, 0x2000 + c);
    for (u64 i = lo; i < hi; i++) {
        double u = rng_unit(&r);
        u32 l = 0, h = a->sigma - 1;
        while (l < h) { u32 m = l + (h - l) / 2; if (a->cdf[m] <= u) l = m + 1; else h = m; }
        a->out[i] = a->perm[l];
    }
}

/* Symbols r ~ Zipf(s) over ranks [0, sigma), mapped through a seeded
 * permutation of [0, sigma); then, while copied < repeat_frac*n, an exact
 * copy of length U[16, 4096] from a uniform earlier source. */
void psk_gen_zipf32(u64 seed, u64 n, u32 sigma, double s, double repeat_frac, u32 *out,
                    int nthreads) {
    if (sigma < 1) sigma = 1;
    double *cdf = (double *)malloc(sizeof(double) * sigma);
    u32 *perm = (u32 *)malloc(sizeof(u32) * sigma);
    double acc = 0;
    for (u32 k = 0; k < sigma; k++) { acc += pow((double)k + 1.0, -s); cdf[k] = acc; }
    for (u32 k = 0; k < sigma; k++) cdf[k] /= acc;
    cdf[sigma - 1] = 2.0;
    rng_t r;
    rng_seed(&r, seed, 0x3000);
    for (u32 k = 0; k < sigma; k++) perm[k] = k;
    for (u32 k = sigma - 1; k > 0; k--) {
        u32 j = (u32)rng_below(&r, (u64)k + 1);
        u32 t = perm[k]; perm[k] = perm[j]; perm[j] = t;
    }
    zipf_arg a = {seed, out, cdf, perm, sigma};
    run_chunks(zipf_chunk, &a, n, nthreads);
    u64 copied = 0, target = (u64)(repeat_frac * (double)n);
    while (copied < target) {
        u64 L = 16 + rng_below(&r, 4096 - 16 + 1);
        if (2 * L + 1 > n) break;
        u64 dst = L + rng_below(&r, n - 2 * L + 1);
        u64 src = rng_below(&r, dst - L + 1);
        memcpy(out + dst, out + src, L * sizeof(u32));
        copied += L;
    }
    free(cdf);
    free(perm);
}

/* count lengths, log-uniform in [lo, hi]: floor(exp(U(ln lo, ln hi))) */
void psk_gen_lengths(u64 seed, u64 count, u64 lo, u64 hi, u64 *out) {
    rng_t r;
    rng_seed(&r, seed, 0x4000);
    double a = log((double)lo), b = log((double)hi);
    for (u64 i = 0; i < count; i++) {
        u64 L = (u64)floor(exp(a + (b - a) * rng_unit(&r)));
        out[i] = L < lo ? lo : (L > hi ? hi : L);
    }
}

/* Poisson arrivals: count exponential inter-arrival gaps (mean_ms) and
 * request lengths (log-uniform [lo, hi], with probability big_p uniform in
 * [big_lo, big_hi]). */
void psk_gen_requests(u64 seed, u64 count, double mean_ms, u64 lo, u64 hi, double big_p,
                      u64 big_lo, u64 big_hi, double *gaps_ms, u64 *lens) {
    rng_t r;
    rng_seed(&r, seed, 0x5000);
    double a = log((double)lo), b = log((double)hi);
    for (u64 i = 0; i < count; i++) {
        double u = rng_unit(&r);
        gaps_ms[i] = -mean_ms * log(1.0 - u);
        if (rng_unit(&r) < big_p)
            lens[i] = big_lo + rng_below(&r, big_hi - big_lo + 1);
        else {
            u64 L = (u64)floor(exp(a + (b - a) * rng_unit(&r)));
            lens[i] = L < lo ? lo : (L > hi ? hi : L);
        }
    }
}
