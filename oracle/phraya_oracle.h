/*
 * phraya_oracle.h -- CPU restatement of phraya-align's extension engines.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path and the CPU baseline that bench.py times beside it.  Nothing in the
 * product (phraya_amd/, include/) links, loads or calls it.
 *
 * Every function restates a function of phraya-align (Rust) at the cited
 * file:line of the upstream repository.  The upstream crate cannot be built
 * in this environment (no Rust toolchain), so the restatement is pinned by the
 * upstream's own known-answer tests, independent DP / Gotoh oracles, CIGAR
 * replay, closed forms and brute force (tests/test_oracle_*.py).
 *
 * Conventions (phraya-align/src/wfa_simd.rs:502-506, :852, :867):
 *   M = exact match, X = mismatch, I = target-only base, D = query-only base.
 *   CIGAR text is "<count><op>" runs with no separators (:1200-1222).
 */
#ifndef PHRAYA_ORACLE_H
#define PHRAYA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_INVALID_INPUT = 1,    /* WfaError::InvalidInput   (lib.rs:107-111) */
  ORC_ALIGNMENT_FAILED = 2, /* WfaError::AlignmentFailed (lib.rs:163-165) */
  ORC_BAD_COSTS = 3         /* costs the upstream code cannot run (x<1 or e<1) */
};

enum {
  ORC_ENGINE_LINEAR = 0,     /* wfa_extend            lib.rs:74-76   */
  ORC_ENGINE_AFFINE = 1,     /* wfa_extend_affine     lib.rs:139-167 */
  ORC_ENGINE_MYERS = 2,      /* myers_extend          lib.rs:106-127 */
  ORC_ENGINE_MYERS_GLOBAL = 3,  /* myers_edit_distance lib.rs:94-96  */
  ORC_ENGINE_MYERS_FWD = 4,  /* myers_extend distance + end column only (no CIGAR) */
  ORC_ENGINE_LINEAR_GLOBAL = 5, /* fill_wfa, global whatever the lengths  wfa_simd.rs:231-345 */
  ORC_ENGINE_AFFINE_GLOBAL = 6  /* fill_wfa_affine_generic(fitting=false) wfa_simd.rs:581-779 */
};

typedef struct orc_result {
  int32_t status;
  int32_t reserved;
  int64_t score;         /* search score: linear s, affine s, Myers distance */
  int64_t edit_distance; /* reported edit distance */
  int64_t query_start, query_end, target_start, target_end;
  int64_t cigar_len;     /* ASCII length of the CIGAR (may exceed cigar_cap) */
} orc_result;

/* count_matching_prefix tiers (wfa_simd.rs:52-155). */
size_t orc_lcp(const uint8_t* a, size_t an, const uint8_t* b, size_t bn);
size_t orc_lcp_scalar(const uint8_t* a, size_t an, const uint8_t* b, size_t bn);
size_t orc_lcp_u64(const uint8_t* a, size_t an, const uint8_t* b, size_t bn);

/* Per-pair wrappers.  cap < 0 means "no cap" (Option::None).  The CIGAR is
 * written NUL-terminated into cigar[0..cigar_cap); r->cigar_len reports the
 * full length so a caller can retry with a bigger buffer. */
int orc_wfa_extend(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                   size_t seed_q, size_t seed_t, int64_t cap,
                   orc_result* r, char* cigar, size_t cigar_cap);
int orc_wfa_extend_affine(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                          size_t seed_q, size_t seed_t,
                          int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                          int64_t cap, orc_result* r, char* cigar, size_t cigar_cap);
int orc_myers_extend(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                     size_t seed_q, size_t seed_t,
                     orc_result* r, char* cigar, size_t cigar_cap);
int orc_myers_edit_distance(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                            orc_result* r, char* cigar, size_t cigar_cap);
/* The global branches alone (no fitting switch), for gaps pinned at both
 * ends: fill_wfa (wfa_simd.rs:231-345) and fill_wfa_affine_generic with
 * fitting = false (:581-779; distance from the CIGAR, :932-934). */
int orc_wfa_global(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn, orc_result* r,
                   char* cigar, size_t cigar_cap);
int orc_wfa_affine_global(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                          int32_t mismatch, int32_t gap_open, int32_t gap_extend, orc_result* r,
                          char* cigar, size_t cigar_cap);
/* myers_extend without the backtrace: distance + target_end only. */
int orc_myers_extend_distance(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                              size_t seed_q, size_t seed_t, orc_result* r);

/* Batch driver used as the CPU baseline: `threads` workers pull one pair at a
 * time from a shared counter (the granularity of rayon par_iter over reads).
 * q_off / t_off hold n+1 offsets into q_seq / t_seq.  seed arrays may be NULL.
 * Results land in `res[n]`; CIGAR strings stay in the returned handle. */
typedef struct orc_batch orc_batch;
orc_batch* orc_batch_run(int engine, uint32_t n,
                         const uint8_t* q_seq, const uint64_t* q_off,
                         const uint8_t* t_seq, const uint64_t* t_off,
                         const uint64_t* seed_q, const uint64_t* seed_t,
                         int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                         int64_t cap, int threads, orc_result* res);
uint64_t orc_batch_cigar_bytes(const orc_batch* b);
/* Copies all CIGARs back to back into arena; off gets n+1 offsets. */
void orc_batch_cigar_copy(const orc_batch* b, char* arena, uint64_t* off);
void orc_batch_free(orc_batch* b);

int orc_hardware_threads(void);

/* Minimizer sketch of s[0, n) (sketch_alphabet, phraya-core/src/types.rs:
 * 1846-1867): (value, position) of each selected k-mer in window order,
 * canonical (DNA) or forward; vals / pos hold at least n - k - w + 2 entries.
 * Returns the count, or -1 for k outside 1..32, w = 0, or a canonical sketch
 * with k + w - 1 even.  The scheme is stated in phraya_oracle.cpp. */
int64_t orc_sketch(const uint8_t* s, uint64_t n, uint32_t k, uint32_t w, int canonical, uint64_t* vals,
                   uint32_t* pos);
/* The same sketch by rolling values and a running T/G count (the CPU
 * baseline's form), and the timed batch of it. */
int64_t orc_sketch_rolling(const uint8_t* s, uint64_t n, uint32_t k, uint32_t w, int canonical, uint64_t* vals,
                           uint32_t* pos);
double orc_sketch_timed(const uint8_t* seq, const uint64_t* off, uint32_t nseq, uint32_t k, uint32_t w, int canonical,
                        int threads, int passes, uint64_t* checksum);

#ifdef __cplusplus
}
#endif
#endif
