/*
 * psacak.h -- C ABI of psacak_amd, an MI355X (gfx950) implementation of the
 * pSACAK suffix-array construction of hucsmn/psacak 0.1.0.
 *
 * The suffix array (SA) of a text is unique, so every entry point below
 * returns exactly the crate's SA for the same input.  The virtual sentinel is
 * not part of the SA (lib.rs:30-43).  Every call returns a status code; none
 * aborts the process.  Device work runs on one GPU (the current HIP device
 * when the context is created).
 *
 * Drop-in map (reference file:line -> entry point):
 *   lib.rs:64-69     pub fn psacak(&[u8]) -> Vec<u32>        psk_psacak (caller allocates)
 *   lib.rs:77-82     pub fn psacak_inplace(&[u8], &mut [u32]) psk_psacak_inplace
 *   psacak8.rs:19    fn psacak8(&[u8], &mut [u32])            psk_psacak8_dev (device buffers)
 *   psacak32.rs:24   fn psacak32(&[u32], &mut [u32], ..)      psk_psacak32 / psk_psacak32_dev
 *   psacak32.rs:912  translate_sample (test helper)           psk_psacak32_raw (translate + psacak32)
 */
#ifndef PSACAK_H
#define PSACAK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PSK_OK 0
#define PSK_EINVAL -1   /* null pointer, sa_len < n (lib.rs:78 assert) */
#define PSK_ETOOLONG -2 /* u8: n >= 2^32-1 (lib.rs:65,79); u32: n > 2^31-1 (psacak32.rs:25) */
#define PSK_EPRECOND -3 /* psacak32: a symbol >= n (not a bucket-pointer text, psacak32.rs:22) */
#define PSK_ENOMEM -4   /* device or host allocation failed */
#define PSK_EBUDGET -5  /* request exceeds the context's HBM budget */
#define PSK_EDEVICE -6  /* HIP error, or no GPU present */
#define PSK_EBUSY -7    /* server: queue full */
#define PSK_ENOTFOUND -8 /* server: unknown request id */

typedef struct psk_ctx psk_ctx;
typedef struct psk_server psk_server;

/* Human-readable text of a status code (static storage). */
const char *psk_strerror(int status);

/* Library version string, e.g. "psacak_amd 0.1.0 (gfx950)". */
const char *psk_version(void);

/* ---- contexts ---------------------------------------------------------- */
/* A context owns a HIP stream and cached device workspace on the current
 * device.  hbm_budget_bytes = 0 means "no budget".  A context is externally
 * synchronised; distinct contexts may be used from different threads.
 * Creation does not touch the GPU; the first call that needs it does. */
int psk_ctx_create(psk_ctx **out, uint64_t hbm_budget_bytes);
void psk_ctx_destroy(psk_ctx *ctx);
/* Message of the last failed call on this context ("" if none). */
const char *psk_last_error(const psk_ctx *ctx);
/* Release cached device workspace (keeps the stream).  ctx NULL: the calling
 * thread's default context (the one the entry points use for ctx NULL). */
int psk_ctx_trim(psk_ctx *ctx);
/* Device workspace the context holds now and the most it has held (bytes);
 * either pointer may be NULL; ctx NULL: the calling thread's default context.
 * A failed call leaves no workspace in use: the next call on the context
 * starts from the same state.  No counterpart in the crate, which uses the
 * caller's SA plus O(block) buffers: for sizing HBM budgets. */
int psk_ctx_workspace(const psk_ctx *ctx, uint64_t *reserved, uint64_t *peak);
/* Naming strategy of this implementation (no counterpart in the crate, whose
 * naming always follows induced sorting, psacak8.rs:57-72, naming.rs:20-259):
 * LMS substrings are named by sorting them directly for at most `rounds`
 * refinement rounds, and by induced sorting when ties remain after that.
 * (or earlier when a quarter of them are still tied after 4 rounds).
 * 0 = always induced sorting.  Default 32.  Results are identical either way.
 * PSK_EINVAL if rounds is outside [0, 64]. */
int psk_ctx_set_direct_rounds(psk_ctx *ctx, int rounds);
/* Round 1 of that direct naming on a byte text (level 0): 1 (default) groups
 * the LMS substrings through a hash dictionary of their distinct leading
 * symbols and sorts only the distinct keys; 0 sorts every substring's key.
 * A text whose substrings take too many distinct values uses sorting either
 * way.  No counterpart in the crate; results are identical.  PSK_EINVAL
 * unless on is 0 or 1. */
int psk_ctx_set_dict_naming(psk_ctx *ctx, int on);
/* A recursion level whose text has at least distinct_per_1024/1024 distinct
 * symbols is sorted by prefix doubling over all its suffixes instead of
 * naming and recursing further (no counterpart in the crate, which always
 * recurses, psacak32.rs:24-56; results are identical).  A level whose
 * suffixes keep tying (long repeats: ties that stop shrinking after 2 rounds,
 * or more than an eighth tied after 4) recurses after all, and so do the
 * levels below it.  0 = always recurse.  The default is the autotuner's
 * choice (psacak_amd/tuning.json).  PSK_EINVAL outside [0, 1024]. */
int psk_ctx_set_doubling(psk_ctx *ctx, int distinct_per_1024);
/* Final induce of a byte text (level 0) with at least min_distinct_bytes
 * distinct bytes: as two chain sorts over its L and S positions, the way the
 * integer levels induce (cost independent of the alphabet), instead of one
 * multi-split step per bucket and scan (the crate's bucket-by-bucket scans,
 * psacak8.rs:127-190, 201-344).  Taken only when its workspace stays within
 * psk_workspace_bytes8.  0 = always bucket steps.  The defaults (one for
 * texts below 64 MiB, one above) are the crossovers the autotuner measured
 * (python -m psacak_amd.tune, psacak_amd/tuning.json).  No
 * counterpart in the crate; results are identical.  PSK_EINVAL outside
 * [0, 256]. */
int psk_ctx_set_chain_induce(psk_ctx *ctx, int min_distinct_bytes);
/* Device bytes a psacak8 call on a text of n bytes needs beyond its
 * input and output buffers (the admission footprint). */
uint64_t psk_workspace_bytes8(uint64_t n);
uint64_t psk_workspace_bytes32(uint64_t n);

/* ---- host-buffer entry points (blocking) -------------------------------- */
/* = psacak_inplace (lib.rs:77-82): SA of text[0..n) into sa[0..n);
 * sa[n..sa_len) is left untouched.  Requires sa_len >= n, n < 2^32-1.
 * ctx may be NULL (a per-thread default context is used). */
int psk_psacak_inplace(psk_ctx *ctx, const uint8_t *text, uint64_t n, uint32_t *sa,
                       uint64_t sa_len);
/* = psacak (lib.rs:64-69) with a caller-owned output of exactly n entries. */
int psk_psacak(psk_ctx *ctx, const uint8_t *text, uint64_t n, uint32_t *sa);
/* = psacak32 (psacak32.rs:24): text must be in bucket-pointer form
 * (an L-type symbol is its bucket's head, an S-type one its tail-1);
 * every symbol < n, n <= 2^31-1. */
int psk_psacak32(psk_ctx *ctx, const uint32_t *text, uint64_t n, uint32_t *sa, uint64_t sa_len);
/* SA of an arbitrary u32 text: translate (psacak32.rs:912-943) then psacak32.
 * All-distinct texts are sorted by symbol (psacak32.rs:873-881). */
int psk_psacak32_raw(psk_ctx *ctx, const uint32_t *text, uint64_t n, uint32_t *sa,
                     uint64_t sa_len);

/* ---- device-buffer entry points ---------------------------------------- */
/* Pointers are device memory on the context's device; stream is a hipStream_t
 * (NULL = the context's own stream).  The call enqueues all work on that stream
 * and returns when the SA is complete on the device (it synchronises the
 * stream); d_sa must hold n u32 and must not overlap d_text (as the crate's
 * borrows exclude it). */
int psk_psacak8_dev(psk_ctx *ctx, const uint8_t *d_text, uint64_t n, uint32_t *d_sa,
                    void *stream);
int psk_psacak32_dev(psk_ctx *ctx, const uint32_t *d_text, uint64_t n, uint32_t *d_sa,
                     void *stream);
int psk_psacak32_raw_dev(psk_ctx *ctx, const uint32_t *d_text, uint64_t n, uint32_t *d_sa,
                         void *stream);

/* ---- zero-allocation device entry point ---------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): psk_psacak8_dev on the current device without a context,
 * taking all of its device workspace from a fixed array of
 * PSK_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * (until the process ends).  The call makes no hipMalloc/hipFree or other
 * allocating call; it enqueues on `stream` (NULL: the null stream) and
 * returns when the SA is complete.  Texts longer than PSK_NOALLOC_MAX_N
 * bytes return PSK_ETOOLONG.  Calls on one GPU share its array: they are
 * serialised by a lock, so calls that would overlap on one GPU run one
 * after another. */
#define PSK_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define PSK_NOALLOC_MAX_N (2ull << 20)
int psk_psacak8_dev_noalloc(const uint8_t *d_text, uint64_t n, uint32_t *d_sa, void *stream);

/* Per-level statistics of the last psacak8/psacak32 call on ctx:
 * rows of {m, m1, mL, k, sum of LMS-substring lengths, symbol bytes}.
 * Returns the number of levels written (<= max_levels). */
int psk_last_levels(const psk_ctx *ctx, uint64_t *rows, int max_levels);

/* Capture of an intermediate, for parity tests against a CPU restatement (no
 * counterpart in the crate): with level >= 0, the following calls on ctx copy
 * the translated reduced text of that recursion level (bucket-pointer form of
 * the names of its LMS substrings, LMS positions in text order: the next
 * level's text, naming.rs:263-274) to host memory of the context; -1 (default)
 * turns it off.  PSK_EINVAL outside [-1, 63]. */
int psk_ctx_set_capture(psk_ctx *ctx, int level);
/* The text captured by the last call: *count entries, copied to out if out
 * is non-NULL (cap entries available, PSK_EINVAL if too few).  PSK_ENOTFOUND
 * if the last call did not name that level (it was sorted otherwise, or the
 * recursion ended above it). */
int psk_last_capture(const psk_ctx *ctx, uint32_t *out, uint64_t cap, uint64_t *count);

/* The last call on ctx: [0] time of its device work (ms, HIP events around
 * the call on its stream), [1] host synchronisations with the stream inside
 * the call (readbacks of small counters that size later launches), [2] of a
 * psk_psacak32_raw call, the time of its translation into bucket-pointer form
 * (translate_sample, psacak32.rs:912-943) within [0], else 0, [3] reserved
 * (0). */
int psk_last_timing(const psk_ctx *ctx, double *ms, int max_entries);

/* ---- kernel profiling ---------------------------------------------------- */
/* No counterpart in the reference (its CLI times whole calls, bin/psacak.rs:
 * 40-48).  on != 0 starts recording HIP-event durations and algorithmic bytes
 * of the tagged kernels launched by this context (and clears old totals);
 * on == 0 stops. */
int psk_profile(psk_ctx *ctx, int on);
/* Records only the tags in names (comma-separated; NULL or "" = every tag),
 * so a timed run can bracket its dominant kernel alone with events. */
int psk_profile_only(psk_ctx *ctx, const char *names);
/* Totals per tagged kernel since psk_profile(ctx, 1): names are written to
 * names ('\n'-separated, NUL-terminated, names_len bytes), and launches[i],
 * ms[i], bytes[i] for i < max.  Returns the number of kernels. */
int psk_profile_stats(psk_ctx *ctx, char *names, uint64_t names_len, uint64_t *launches,
                      double *ms, double *bytes, int max);

/* ---- batched ------------------------------------------------------------ */
/* SA of count independent texts.  texts[i] (lens[i] bytes) -> sas[i].
 * Host buffers.  Small texts are packed into shared device batches. */
int psk_batch8(psk_ctx *ctx, const uint8_t *const *texts, const uint64_t *lens,
               uint32_t *const *sas, uint64_t count);
/* Same on device buffers: one concatenated text buffer d_texts with offsets
 * (count+1 entries, host array), SA written to d_sas at the same offsets. */
int psk_batch8_dev(psk_ctx *ctx, const uint8_t *d_texts, const uint64_t *offsets,
                   uint64_t count, uint32_t *d_sas, void *stream);

/* ---- device-side checker ------------------------------------------------ */
/* Verifies on the GPU that d_sa is the SA of d_text (permutation + the
 * Burkhardt-Kaerkkaeinen adjacent-pair condition).  *ok = 1 if valid. */
int psk_check8_dev(psk_ctx *ctx, const uint8_t *d_text, uint64_t n, const uint32_t *d_sa,
                   int *ok, void *stream);
/* The same check for an integer text (u32 symbols; n <= 2^31-1). */
int psk_check32_dev(psk_ctx *ctx, const uint32_t *d_text, uint64_t n, const uint32_t *d_sa,
                    int *ok, void *stream);
/* ---- serving ------------------------------------------------------------ */
/* No counterpart in the reference (a library call only).  A server runs
 * max_workers worker threads, each with its own context and stream.  Requests
 * are admitted in arrival order while the estimated device footprint of the
 * work in flight (psk_server_footprint) stays within hbm_budget_bytes
 * (0 = no budget); consecutive small requests are coalesced into one batch;
 * jobs of different workers overlap on the device. */
int psk_server_create(psk_server **out, uint64_t hbm_budget_bytes, int max_workers);
/* Device bytes the server reserves for a job of `bytes` text bytes. */
uint64_t psk_server_footprint(uint64_t bytes, int batched);
/* Submit a text (copied).  *id receives the request id.  PSK_EBUDGET if the
 * request alone exceeds the budget. */
int psk_server_submit(psk_server *s, const uint8_t *text, uint64_t n, uint64_t *id);
/* As psk_server_submit, but the text is not copied and the SA is written
 * straight into sa (n entries): both buffers must stay valid and untouched
 * until poll/wait reports the request done (poll/wait then copy nothing).
 * No counterpart in the crate (whose psacak_inplace, lib.rs:77-82, is the
 * blocking form of the same contract). */
int psk_server_submit_into(psk_server *s, const uint8_t *text, uint64_t n, uint32_t *sa,
                           uint64_t *id);
/* *done = 1 and, when finished, the SA (n entries) is copied to sa and
 * *latency_ms = submit-to-completion time; the request is then released.
 * Returns the request's status (or PSK_ENOTFOUND). */
int psk_server_poll(psk_server *s, uint64_t id, int *done, uint32_t *sa, double *latency_ms);
/* Block until request id completes (timeout_ms < 0: no limit), then as poll.
 * Several callers may poll or wait on one id at once: exactly one of them
 * receives the result; the others return PSK_ENOTFOUND. */
int psk_server_wait(psk_server *s, uint64_t id, uint32_t *sa, double *latency_ms,
                    double timeout_ms);
/* Peak device bytes reserved by the server's contexts since creation. */
uint64_t psk_server_peak_bytes(psk_server *s);
/* Number of requests submitted but not yet finished. */
uint64_t psk_server_pending(psk_server *s);
/* Counters since creation, out[0..min(n,7)): jobs admitted, requests in
 * them, admission checks that found the head request blocked by the budget,
 * workspace trims, the most estimated footprint in flight at once (bytes),
 * requests admitted out of arrival order (0 by design), and the peak device
 * bytes reserved.  Returns the number written. */
int psk_server_stats(psk_server *s, uint64_t *out, int n);
void psk_server_destroy(psk_server *s);

#ifdef __cplusplus
}
#endif
#endif /* PSACAK_H */
