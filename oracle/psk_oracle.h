/*
 * psk_oracle.h -- CPU oracle for the pSACAK suffix-array path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as a checker
 * or as the timed CPU baseline.  The product path (psacak_amd) never calls
 * into it.
 *
 * Every function restates the reference crate hucsmn/psacak 0.1.0; the
 * file:line citations point into that crate's source tree.
 *
 * Pinning (see tests/test_oracle.py):
 *   or_saca_naive8/32   brute-force suffix sort (definition, common.rs:13-25);
 *                       pinned by the crate's doc-test vectors (lib.rs:15,27,42)
 *                       and by the BK linear-time checker below.
 *   or_sacak8           serial psacak8 restatement (psacak8.rs:19-190,
 *                       naming.rs:20-93,263-430); pinned against naive on the
 *                       crate's table texts (psacak8.rs:508-524), random and
 *                       adversarial texts.
 *   or_sacak32          serial psacak32 restatement on bucket-pointer texts
 *                       (psacak32.rs:24-56); pinned against naive on the crate's
 *                       table texts (psacak32.rs:826-842) after translation.
 *   or_translate32      translate_sample (psacak32.rs:912-943); pinned by the
 *                       closed form (heads/tails of rank buckets) in tests.
 *   or_check_sa8/32     Burkhardt-Kaerkkaeinen O(n) checker; pinned by
 *                       planted mistakes (swapped / duplicated entries).
 *   or_types8/32        L/S/LMS typing (common.rs:43-117); pinned against the
 *                       brute-force definitions is_stype/is_left_most
 *                       (common.rs:370-396) restated in Python.
 *   or_name8/32         LMS-substring naming + translate (naming.rs:51-93,
 *                       263-274); pinned by brute-force substring ranks.
 */
#ifndef PSK_ORACLE_H
#define PSK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* brute-force suffix sort, sentinel excluded (lib.rs:30-43). */
void or_saca_naive8(const uint8_t *t, uint32_t n, uint32_t *sa);
void or_saca_naive32(const uint32_t *t, uint32_t n, uint32_t *sa);

/* serial SACA-K restatements. Return 0 on success, <0 on bad input. */
int or_sacak8(const uint8_t *t, uint32_t n, uint32_t *sa);
int or_sacak32(const uint32_t *t, uint32_t n, uint32_t *sa);

/* translate an arbitrary u32 text into bucket-pointer form.
 * Returns 1 if translated, 0 if every symbol is distinct (no translation,
 * psacak32.rs:921-923). */
int or_translate32(const uint32_t *t, uint32_t n, uint32_t *out);

/* SA checker: 0 = valid SA of t; 1 = not a permutation; 2 = order violated. */
int or_check_sa8(const uint8_t *t, uint32_t n, const uint32_t *sa);
int or_check_sa32(const uint32_t *t, uint32_t n, const uint32_t *sa);

/* typing: out[i] bit0 = S-type, bit1 = LMS (common.rs:43-68). */
void or_types8(const uint8_t *t, uint32_t n, uint8_t *out);
void or_types32(const uint32_t *t, uint32_t n, uint8_t *out);

/* bucket bounds (257 entries: exclusive prefix of counts) and per-char LMS
 * counts (256), psacak8.rs:355-374. */
void or_buckets8(const uint8_t *t, uint32_t n, uint32_t *bounds, uint32_t *lmscnt);

/* Name LMS substrings given the sorted LMS positions (lms_sorted, m1
 * entries).  Writes the translated reduced string (bucket-pointer form,
 * text order of the LMS positions) to reduced[0..m1) and returns the number
 * of distinct names k (naming.rs:51-93 + translate :263-274). */
uint32_t or_name8(const uint8_t *t, uint32_t n, const uint32_t *lms_sorted, uint32_t m1,
                  uint32_t *reduced);
uint32_t or_name32(const uint32_t *t, uint32_t n, const uint32_t *lms_sorted, uint32_t m1,
                   uint32_t *reduced);

/* Stage-1 of one level: sorted LMS substrings (ties in induced order).
 * Returns m1, writes lms_sorted[0..m1). */
uint32_t or_stage1_8(const uint8_t *t, uint32_t n, uint32_t *lms_sorted);
uint32_t or_stage1_32(const uint32_t *t, uint32_t n, uint32_t *lms_sorted);

/* per-level statistics of the recursion of or_sacak8 (levels written to
 * stats[5*l .. 5*l+5) = {m, m1, mL, k, sum of LMS-substring lengths});
 * returns the number of levels (<= max_levels). */
int or_level_stats8(const uint8_t *t, uint32_t n, uint64_t *stats, int max_levels);

/* Multi-threaded psacak8 (psk_par.c): the levels of or_sacak8 with the
 * crate's pipelined parallel induce (psacak8.rs:201-344, pipeline.rs:20-358)
 * on OpenMP threads.  Same SA as or_sacak8.  Timed CPU baseline only. */
int or_par_sacak8(const uint8_t *t, uint32_t n, uint32_t *sa);
/* Multi-threaded psacak32 on a bucket-pointer text; same SA as or_sacak32. */
int or_par_sacak32(const uint32_t *t, uint32_t m, uint32_t *sa);
/* threads or_par_sacak8 uses (omp_get_max_threads()), and a setter. */
int or_par_threads(void);
void or_par_set_threads(int p);

#ifdef __cplusplus
}
#endif
#endif
