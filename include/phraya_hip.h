/*
 * phraya_hip.h -- C ABI of libphraya_hip.so, the MI355X (gfx950) extension
 * engines of phraya-align.
 *
 * The per-pair entry points replace phraya-align's extension API one for one
 * (paths relative to the upstream repository root):
 *   phx_wfa_extend           <- wfa_extend / wfa_extend_naive    phraya-align/src/lib.rs:63-76
 *   phx_wfa_extend_affine    <- wfa_extend_affine                phraya-align/src/lib.rs:139-167
 *   phx_myers_extend         <- myers_extend                     phraya-align/src/lib.rs:106-127
 *   phx_myers_edit_distance  <- myers_edit_distance              phraya-align/src/lib.rs:94-96
 * and phx_align_batch is the batched form of all four that a batched
 * extend_chains (phraya-align/src/executor.rs:528-567) calls once per read
 * batch.  Results are bit-identical to the upstream engines: score, edit
 * distance, query/target end and CIGAR, including tie-breaking.
 *
 * CIGAR convention (phraya-align/src/wfa_simd.rs:502-506): M = exact match,
 * X = mismatch, I = target-only base, D = query-only base.  Runs are returned
 * as u32 words (length << 4 | op) with op 0 M, 1 X, 2 I, 3 D;
 * phx_format_cigar renders them as phraya's "<count><op>" text.
 *
 * Threading: a context serialises its own submissions (internal lock), so one
 * context may be shared by many host threads; results never depend on what
 * else runs on the device.  No C++ exception crosses this boundary.
 */
#ifndef PHRAYA_HIP_H
#define PHRAYA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, per pair and per call. */
enum {
  PHX_OK = 0,
  PHX_INVALID_INPUT = 1,    /* WfaError::InvalidInput: seed beyond a sequence (lib.rs:107-111, :145-149) */
  PHX_ALIGNMENT_FAILED = 2, /* WfaError::AlignmentFailed: cap / search bound exhausted (lib.rs:163-165) */
  PHX_E_ARG = 32,           /* bad argument to the call itself */
  PHX_E_DEVICE = 33,        /* HIP runtime error */
  PHX_E_NOMEM = 34,         /* device memory exhausted */
  PHX_E_CAPACITY = 35,      /* caller's CIGAR arena too small; run_off[n] holds the size needed */
  PHX_E_INTERNAL = 36       /* an alignment could not be completed by any kernel tier */
};

/* Engines (what phraya-align's extend_anchor dispatches to, executor.rs:379-404). */
enum {
  PHX_ENGINE_LINEAR = 0,       /* wfa_extend: unit-cost WFA, fitting iff tn > qn + qn/2 + 10 */
  PHX_ENGINE_AFFINE = 1,       /* wfa_extend_affine: gap-affine WFA */
  PHX_ENGINE_MYERS = 2,        /* myers_extend: Myers, fitting iff tn > qn + qn/2 + 10 */
  PHX_ENGINE_MYERS_GLOBAL = 3, /* myers_edit_distance: global Myers (seeds ignored) */
  PHX_ENGINE_MYERS_DIST = 4    /* myers_extend without the backtrace: distance + end column */
};

/* AffineCosts (phraya-align/src/wfa_simd.rs:520-525): an L-base gap costs
 * gap_open + L * gap_extend.  The device path requires mismatch >= 1,
 * gap_extend >= 1, gap_open >= 0 (the upstream loop cannot run otherwise). */
typedef struct phx_costs {
  int32_t mismatch;
  int32_t gap_open;
  int32_t gap_extend;
} phx_costs;

/* Alignment (phraya-align/src/lib.rs:31-39) plus the engine's search score. */
typedef struct phx_alignment {
  int32_t status;
  int32_t reserved;
  int64_t score;          /* linear s, affine s, or Myers distance */
  int64_t edit_distance;  /* what upstream reports as edit_distance */
  int64_t query_start, query_end, target_start, target_end;
} phx_alignment;

/* A batch of pairs in host memory: sequence i is seq[off[i] .. off[i+1]). */
typedef struct phx_batch {
  uint32_t n;
  int32_t engine;
  phx_costs costs;        /* PHX_ENGINE_AFFINE only */
  int64_t cap;            /* max search score (upstream max_s_cap); < 0 = none */
  const uint8_t* q_seq;
  const uint64_t* q_off;  /* n + 1 entries */
  const uint8_t* t_seq;
  const uint64_t* t_off;  /* n + 1 entries */
  const uint64_t* seed_q; /* may be NULL (all zero) */
  const uint64_t* seed_t; /* may be NULL (all zero) */
} phx_batch;

/* Caller-owned result arrays (n entries each; `runs` holds runs_cap words).
 * Pair i's CIGAR runs are runs[run_off[i] .. run_off[i+1]).  For
 * PHX_ENGINE_MYERS_DIST no runs are produced and runs may be NULL. */
typedef struct phx_results {
  phx_alignment* aln;
  uint32_t* runs;
  uint64_t runs_cap;
  uint64_t* run_off;      /* n + 1 entries */
} phx_results;

typedef struct phx_ctx phx_ctx;
typedef struct phx_plan phx_plan;

/* Tuning knobs; NULL (or a zero field) selects the built-in default. */
typedef struct phx_tuning {
  uint32_t wave_lds_kib;     /* LDS tier: LDS bytes (KiB) per wave (ring + sequences + history) */
  uint32_t wave_slots;       /* LDS tier: pairs in flight */
  uint32_t hbm_slots;        /* HBM wave tier: pairs in flight */
  uint32_t hbm_slot_mib;     /* HBM wave tier: history slot per pair when no estimate (MiB) */
  uint64_t hbm_budget_mib;   /* device scratch budget of the HBM tiers (MiB) */
  uint32_t block256_min_len; /* qn + tn above which a pair runs a 256-thread workgroup (0 = 2048 for
                                affine costs of 4/6/2 or 2/4/1 shape, which have a register
                                wavefront on the workgroup tiers; 8192 otherwise) */
  uint32_t block1024_min_len;/* ... a 1024-thread workgroup (0 = 16384) */
  uint32_t heavy_permille;   /* per mille of each length bucket (largest estimates) run one tier wider (0 = 50) */
  uint32_t block_lds_ring;   /* workgroup tiers: LDS ring width in diagonals for pairs whose wavefront fits
                                (power of two; 0 = 1024; 1 = ring always in HBM) */
  uint32_t block_flight;     /* workgroup tiers: pairs in flight per CU, as a multiple of 4 (256-thread)
                                and 1 (1024-thread); 0 = 1, at most 4 */
  uint32_t block_regs;       /* workgroup tiers, affine 4/6/2 and 2/4/1 in score steps: keep the
                                wavefront in registers (0 = pairs whose estimate fits the window,
                                1 = every pair whose sequences fit LDS, a pair that outgrows the
                                window being re-run on the ring path, 2 = never) */
  uint32_t lds_rows;         /* LDS tier, global-branch affine pairs with mismatch 2 and extend 1 in score
                                steps (4/6/2, 2/4/1) or 1/0/1 and |qn - tn| <= 48: banded rows, four pairs
                                per wave in 16-lane rows when an alignment within the row's band exists,
                                else two per wave in 32-lane rows (0 or 1 = on, 2 = off: one pair per
                                wave, 3 = every such pair in 32-lane rows) */
  uint32_t band_permille;    /* workgroup register tiers: the band's bound on a pair's optimum is its step
                                estimate x band_permille / 1000 (0 = 800; at most 1000); a pair whose
                                optimum exceeds it is re-run with the whole estimate as bound, then
                                without one.  Results do not depend on it. */
  uint32_t regs_narrow;      /* workgroup register tiers: a pair runs on the narrowest register tier
                                whose window holds its band: one wave (1,024 diagonals, 16 pairs per
                                CU), 256 threads (4,096, four per CU) or 512 (8,192, two per CU)
                                (0 or 1 = on, 2 = off, 3 = on without the one-wave tier) */
} phx_tuning;

/* Context: owns the stream, device buffers and the submission lock.
 * device < 0 selects the current device. Returns NULL on failure. */
phx_ctx* phx_ctx_create(int device, uint32_t flags, const phx_tuning* tuning);
void phx_ctx_destroy(phx_ctx* ctx);
/* Text of the last failure of a call the calling thread made on ctx ("" when
 * its last failure was on another context or there was none).  Errors are
 * kept per thread, so threads sharing a context never overwrite each other's;
 * the text stays valid until that thread's next failing call.  One-shot entry
 * points (phx_align_batch, phx_seed_and_chain_batch, phx_extract_variants)
 * return the failing step's code: PHX_E_NOMEM for device memory,
 * PHX_E_DEVICE for a HIP runtime error, PHX_E_ARG for the input. */
const char* phx_ctx_last_error(const phx_ctx* ctx);

/* Upper bound of CIGAR run words a batch can need (sum of qn + tn + 1). */
uint64_t phx_cigar_bound(const phx_batch* batch);

/* Align a batch end to end (pack, upload, run, download). Returns PHX_OK when
 * every pair got a final status (PHX_OK / PHX_INVALID_INPUT /
 * PHX_ALIGNMENT_FAILED in aln[i].status), else a PHX_E_* code. */
int phx_align_batch(phx_ctx* ctx, const phx_batch* batch, phx_results* out);

/* Allocation-free form of phx_align_batch (no counterpart in the reference;
 * for callers that cannot allocate device memory: tight serving loops and
 * harnesses that forbid hipMalloc).  Same arguments, results and status
 * codes as phx_align_batch, bit for bit.  It calls no hipMalloc, hipFree,
 * hipHostMalloc or other allocating routine and creates no events: every
 * device buffer (packed sequences, descriptors, results, CIGAR run slots,
 * work lists, overflow flags, wavefront history, Myers scratch) is carved
 * from a fixed per-GPU scratch array of PHX_NOALLOC_SCRATCH_BYTES that the
 * library holds from its load until the process ends.  A batch whose
 * buffers do not fit returns PHX_E_NOMEM (the reason in
 * phx_ctx_last_error) before anything runs; pairs whose wavefront history
 * outgrows what is left fail the same way.  Calls on one GPU share its
 * array and are serialised by a process-wide lock (they would otherwise race
 * on it); each call clears what it needs zeroed and leaves the rest. */
#define PHX_NOALLOC_SCRATCH_BYTES (64ull << 20)
int phx_align_batch_noalloc(phx_ctx* ctx, const phx_batch* batch, phx_results* out);

/* Split form of phx_align_batch for callers that keep inputs resident in HBM
 * and launch repeatedly: create packs and uploads, run launches the kernels
 * on the resident inputs (results stay in HBM), results downloads. */
phx_plan* phx_plan_create(phx_ctx* ctx, const phx_batch* batch);
int phx_plan_run(phx_plan* plan);
int phx_plan_results(phx_plan* plan, phx_results* out);
/* Device time of the last phx_plan_run, in milliseconds: from the run's
 * start (its first kernel's start event when the run is a single launch) to
 * its end; and the summed time of its alignment kernel launches alone (events
 * recorded on the launch stream right before and after each launch).  The
 * events belong to the plan; the times are read from them on the first call
 * after a run (not by phx_plan_run itself). */
double phx_plan_last_ms(const phx_plan* plan);
double phx_plan_last_kernel_ms(const phx_plan* plan);
uint32_t phx_plan_last_launches(const phx_plan* plan);
/* Launch i (0 <= i < launches) of the last run: which kernel tier ran it
 * (0 LDS-resident wave per pair, 1 wave per pair with HBM slots, 2 a
 * 256-thread workgroup per pair, 3 a 1024-thread workgroup per pair, 8 Myers
 * lane per pair), its workgroups and pairs, its device time (events on its
 * stream around it), and what its pairs account for: algorithmic HBM bytes
 * (packed inputs + 12 B of results + 4 B per CIGAR run; Myers distance: 16 B),
 * wavefront cells written and score steps.  Instrumentation; no counterpart
 * in the reference. */
typedef struct phx_launch_info {
  int32_t tier;     /* 0 LDS wave, 1 HBM wave, 2 256-thread, 3 1024-thread, 4 512-thread,
                       5 one-wave workgroup, 8 Myers */
  uint32_t pairs;
  uint32_t grid;
  uint32_t storage; /* workgroup tiers: 0 ring in HBM, 1 ring in LDS, 2 register wavefront;
                       LDS tier: 0 a pair per wave, 3 banded rows (2-bit pairs), 4 banded rows
                       (byte pairs) */
  double ms;
  uint64_t algorithmic_bytes;
  uint64_t cells;
  uint64_t steps;
} phx_launch_info;
int phx_plan_last_launch(const phx_plan* plan, uint32_t i, phx_launch_info* out);
/* Re-tune a plan (non-zero fields of *tuning replace the plan's): the pairs
 * are regrouped into tiers without re-uploading them.  Results never depend
 * on the tuning. */
int phx_plan_set_tuning(phx_plan* plan, const phx_tuning* tuning);
int phx_plan_get_tuning(const phx_plan* plan, phx_tuning* tuning);
/* Autotuner: times the plan under candidate tunings (tier length edges, i.e.
 * threads per pair, and heavy share; workgroup-tier and HBM wave-tier pairs
 * in flight; LDS per wave vs waves in flight of the LDS tier), stage by
 * stage, keeps the fastest and applies it; every trial runs the whole plan.
 * max_trials = 0: no limit (at most 16 trials).
 * *best / *best_ms (may be NULL) get the chosen tuning and its run time. */
int phx_plan_autotune(phx_plan* plan, uint32_t max_trials, phx_tuning* best, double* best_ms);
/* Graph mode (WFA engines; no counterpart in the reference, which has no
 * device launches): the run after enabling records its launch sequence
 * (counter reset, every group's kernel on its stream, join, counter
 * readback) into a HIP graph, and later runs replay it with one
 * hipGraphLaunch.  A new tuning or a grown scratch arena re-records it; the
 * rare overflow re-runs stay outside the graph.  Results are the same in
 * both modes; event nodes around each kernel node keep the per-launch
 * timings above.  enable: 0 or 1. */
int phx_plan_use_graph(phx_plan* plan, int enable);
/* Algorithmic HBM bytes of one run: packed inputs + results + CIGAR runs. */
uint64_t phx_plan_algorithmic_bytes(const phx_plan* plan);
void phx_plan_destroy(phx_plan* plan);

/* Render runs as phraya CIGAR text. Writes at most cap bytes including the
 * NUL; returns the full text length (excluding the NUL). */
size_t phx_format_cigar(const uint32_t* runs, uint64_t nruns, char* buf, size_t cap);
/* Renders the CIGARs of n pairs in one call: pair i's runs[run_off[i] ..
 * run_off[i + 1]) become buf[text_off[i] .. text_off[i + 1]) (no NULs).
 * Returns the total text length; writes no run whose text would pass cap
 * (text_off is complete either way, so a caller can size buf and retry).
 * Host only; text_off may be NULL. */
uint64_t phx_format_cigars(const uint32_t* runs, const uint64_t* run_off, uint64_t n, char* buf,
                           uint64_t cap, uint64_t* text_off);

/* Per-pair forms of the upstream API.  The CIGAR text is written to cigar
 * (NUL-terminated, truncated to cigar_cap); the return value is the pair
 * status or a PHX_E_* code.  cigar_len (may be NULL) gets the full length. */
/* Name of the implementation the calling thread's last per-pair call ran,
 * "hip-<arch>" (e.g. "hip-gfx950"), or "" before any: the device counterpart
 * of upstream's thread-local LAST_IMPL / get_active_dispatch_target
 * (phraya-align/src/wfa_simd.rs:25-29, :181-184), kept per thread like it.
 * The string lives as long as the context that ran the call. */
const char* phx_last_impl(void);
int phx_wfa_extend(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t, size_t tlen,
                   uint64_t seed_q, uint64_t seed_t, phx_alignment* out, char* cigar,
                   size_t cigar_cap, size_t* cigar_len);
int phx_wfa_extend_affine(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t,
                          size_t tlen, uint64_t seed_q, uint64_t seed_t, phx_costs costs,
                          phx_alignment* out, char* cigar, size_t cigar_cap, size_t* cigar_len);
int phx_myers_extend(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t, size_t tlen,
                     uint64_t seed_q, uint64_t seed_t, phx_alignment* out, char* cigar,
                     size_t cigar_cap, size_t* cigar_len);
int phx_myers_edit_distance(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t,
                            size_t tlen, int64_t* edit_distance, char* cigar, size_t cigar_cap,
                            size_t* cigar_len);

/* ---- Request server (continuous batching) -------------------------------
 *
 * Serves a stream of extension requests as they arrive, the per-anchor calls
 * of extend_anchor (phraya-align/src/executor.rs:379-404) made by many reads
 * at once.  A native scheduler thread admits each request against an HBM
 * budget for wavefront state, using a per-request estimate of its history
 * (O(s^2) cells for an s-step search).  It launches admitted requests in
 * length-tiered groups on their own streams: short reads, long reads and
 * the widest pairs each have a lane, so a short read never waits behind a
 * long one.  Results are delivered as soon as their group completes.
 * Targets are either explicit or windows of a reference kept resident in HBM.
 * Every result equals the one phx_align_batch / phx_wfa_extend_affine gives
 * for the same pair alone.  Requests are independent: one that fails (bad
 * seed, no memory) reports its own status and never affects another. */
typedef struct phx_server phx_server;

typedef struct phx_serve_config {
  int32_t engine;            /* PHX_ENGINE_AFFINE or PHX_ENGINE_LINEAR */
  phx_costs costs;           /* affine costs */
  int64_t cap;               /* max search score (upstream max_s_cap); < 0 = none */
  const uint8_t* reference;  /* copied to HBM at start; NULL = no reference */
  uint64_t reference_len;
  uint64_t hbm_budget_mib;   /* HBM for per-request wavefront state; 0 = default */
  uint32_t max_group;        /* most requests per launch; 0 = default (8192) */
  uint32_t lanes_per_tier;   /* launch groups in flight per tier; 0 = default (1) */
} phx_serve_config;

typedef struct phx_request {
  uint64_t tag;              /* echoed in the response */
  const uint8_t* query;
  uint64_t query_len;
  const uint8_t* target;     /* explicit target, or NULL for a reference window */
  uint64_t target_len;
  uint64_t ref_start;        /* window [ref_start, ref_start + ref_len) when target is NULL */
  uint64_t ref_len;
  uint64_t seed_q, seed_t;   /* SeedAnchor into the query and the target (window) */
} phx_request;

typedef struct phx_response {
  uint64_t tag;
  phx_alignment aln;         /* positions are relative to the query and target (window) */
  double latency_ms;         /* from submission to the result being ready on the host */
} phx_response;

typedef struct phx_serve_stats {
  uint64_t submitted, completed, waiting, in_flight;
  uint64_t launches, reruns;
  uint64_t arena_bytes, arena_used_bytes, arena_peak_bytes;
} phx_serve_stats;

/* Start a server on ctx's device (own streams; ctx stays usable).  NULL on
 * failure (phx_ctx_last_error says why). */
phx_server* phx_serve_start(phx_ctx* ctx, const phx_serve_config* cfg);
/* Queue n requests (sequences are copied; the call does not block on the
 * GPU).  PHX_E_ARG if a reference window lies outside the reference. */
int phx_serve_submit(phx_server* srv, const phx_request* reqs, uint32_t n);
/* Take up to max_out finished responses, waiting up to timeout_us for the
 * first one (0 = do not wait).  Response i's CIGAR runs go to
 * runs[run_off[i] .. run_off[i+1]); responses whose runs do not fit stay
 * queued (PHX_E_CAPACITY when not even the first one fits; run_off[0] then
 * holds the words it needs). */
int phx_serve_poll(phx_server* srv, phx_response* out, uint32_t max_out, uint32_t* runs,
                   uint64_t runs_cap, uint64_t* run_off, uint32_t* n_out, int32_t timeout_us);
int phx_serve_stats_get(phx_server* srv, phx_serve_stats* st);
/* Finish every submitted request, then stop the scheduler and free the
 * server (responses not yet polled are dropped). */
int phx_serve_stop(phx_server* srv);

/* ---- Contig mode -----------------------------------------------------------
 *
 * Aligns pairs of long contig segments (query i against target i, end to
 * end) by anchoring them on shared minimizers and aligning every gap between
 * anchors with the batched engine.  No counterpart in phraya-align's
 * extension API; its gap alignments are the global branches fill_wfa and
 * fill_wfa_affine_generic(fitting = false) (phraya-align/src/wfa_simd.rs:231-345,
 * :581-779) on the windows: a gap is pinned at both ends, so the wrappers'
 * fitting switch never applies.
 * Anchors: forward-strand (k, w) minimizers (leftmost smallest hash of a
 * mixing function of the 2-bit k-mer) that occur once among each
 * sequence's minimizers, paired by k-mer, reduced to the longest chain
 * increasing in both sequences, then spaced out: walking the chain, the next
 * anchor kept is the first at least `spacing` query bases past the previous
 * one, spacing drawn log-uniform in [gap_min, gap_max] from a SplitMix64
 * stream seeded by (seed, pair index).  The contig CIGAR is gap 0, anchor 1
 * as k M, gap 1, ..., with equal neighbouring runs merged; the contig score
 * is the sum of its gap scores, which equals the cost of the stitched CIGAR
 * (anchors are exact matches and cost nothing). */
typedef struct phx_contig_params {
  uint32_t k;        /* minimizer k (8..16); 0 = 15 */
  uint32_t w;        /* minimizer window in k-mers; 0 = 10 */
  uint32_t gap_min;  /* anchor spacing range in bases; 0 = 1000 */
  uint32_t gap_max;  /* 0 = 20000 */
  uint64_t seed;
} phx_contig_params;

typedef struct phx_contig_plan phx_contig_plan;

/* Anchors of one pair (host only): writes up to cap (query pos, target pos)
 * pairs to out[2i], out[2i+1] and returns how many there are (-1 on bad
 * arguments). */
int64_t phx_contig_anchors(const uint8_t* q, uint64_t qn, const uint8_t* t, uint64_t tn,
                           const phx_contig_params* params, uint64_t pair_index, uint64_t* out,
                           uint64_t cap);
/* Anchor every pair of the batch (seeds ignored), keep both sides resident in
 * HBM and prepare one batched plan over all gaps. */
phx_contig_plan* phx_contig_plan_create(phx_ctx* ctx, int32_t engine, phx_costs costs,
                                        const phx_batch* contigs, const phx_contig_params* params);
/* Align all gaps on the GPU, then stitch each contig's CIGAR (host threads). */
int phx_contig_plan_run(phx_contig_plan* plan);
/* Per contig: aln[i] (score = sum of gap scores, edit distance of the stitched
 * CIGAR) and the stitched runs, as phx_plan_results. */
int phx_contig_plan_results(phx_contig_plan* plan, phx_results* out);
/* Gap alignments per run, and contig i's gaps as (q0, q1, t0, t1) quadruples
 * (returns the gap count; writes up to cap quadruples). */
uint64_t phx_contig_plan_gaps(const phx_contig_plan* plan);
int64_t phx_contig_plan_layout(const phx_contig_plan* plan, uint32_t contig, uint64_t* gaps,
                               uint64_t cap);
/* The batched plan of all gaps (to read its kernel times / algorithmic bytes). */
phx_plan* phx_contig_plan_gap_plan(phx_contig_plan* plan);
/* Wall time of the last run (GPU alignment + result download + stitching),
 * the device time of the gap alignment, and the stitching time, in ms. */
double phx_contig_plan_last_ms(const phx_contig_plan* plan);
double phx_contig_plan_last_gpu_ms(const phx_contig_plan* plan);
double phx_contig_plan_last_stitch_ms(const phx_contig_plan* plan);
void phx_contig_plan_destroy(phx_contig_plan* plan);

/* Memory model of one WFA request; host only, no GPU needed.  *steps (may be
 * NULL) gets the score steps its search is estimated to take: the end gap
 * for the length difference plus the per-base event rate implied by how many
 * of the query's 12-mers occur in the target, with margin.  Returns the HBM
 * bytes (history cells of every step + score-frame ring) that the server
 * reserves for it in an HBM tier, or -1 for bad arguments.  No counterpart
 * in phraya-align; it is for callers that budget device memory themselves. */
int64_t phx_estimate_request(int32_t engine, phx_costs costs, const uint8_t* q, uint64_t qn,
                             const uint8_t* t, uint64_t tn, int32_t* steps);

/* Library version string. */
const char* phx_version(void);

/* ---------------------------------------------------------------------------
 * Batched read extension: phraya-align's extend_chains (executor.rs:528-567)
 * for many reads at once.  Every anchor of every read is one pair of one
 * device batch per engine; the host then picks each read's primary and
 * alternatives with score_alignments' rules (lib.rs:170-199).
 * ------------------------------------------------------------------------- */

/* Strategy and gap model of phraya's executor (executor.rs:330-336, :379-404). */
enum { PHX_STRATEGY_FAST = 0, PHX_STRATEGY_BALANCED = 1, PHX_STRATEGY_SENSITIVE = 2 };
enum { PHX_GAP_AFFINE = 0, PHX_GAP_LINEAR = 1 };

/* One read's extension input: its bases and the target_start of each chain
 * of chaining's output, in rank order (Chain::target_start). */
typedef struct phx_read_chains {
  const uint8_t* query;
  uint64_t query_len;
  const uint64_t* chain_target_start;
  uint32_t n_chains;
} phx_read_chains;

/* One read's extension output.  Anchors [first_anchor, first_anchor +
 * n_anchors) of the anchor arrays belong to it, in anchor order. */
typedef struct phx_read_extension {
  int32_t status;          /* PHX_OK: some anchor extended; PHX_ALIGNMENT_FAILED: none did
                              (upstream: extend_chains returns None) */
  int32_t primary;         /* anchor index (relative to first_anchor) of the primary, or -1 */
  uint32_t n_anchors;
  uint32_t n_alternatives; /* anchors with role 3 */
  uint64_t first_anchor;
} phx_read_extension;

/* A target kept resident in HBM for the executor entry points below (the
 * `_ref` forms): packed once, so a call uploads only its reads and each
 * anchor's window target[pos .. min(pos + 2 qlen, |target|)] is a descriptor
 * into it.  phraya's TargetContext (executor.rs) would hold one per target;
 * the reference API has no counterpart (it takes the target slice per call).
 * The bytes are copied; NULL on failure (phx_ctx_last_error says why). */
typedef struct phx_reference phx_reference;
phx_reference* phx_reference_create(phx_ctx* ctx, const uint8_t* target, uint64_t len);
void phx_reference_destroy(phx_reference* ref);

/* chain_cap (executor.rs:330-336): anchors extended per read. */
uint32_t phx_chain_cap(int32_t strategy);
/* Anchors of a batch and an upper bound of their CIGAR runs (host only). */
int phx_extend_chains_bounds(int32_t strategy, uint64_t target_len, const phx_read_chains* reads,
                             uint32_t n_reads, uint64_t* anchors, uint64_t* runs);
/* Extends every read's anchors against target (anchors_from_chains,
 * executor.rs:347-362; window target[..min(target_pos + 2 qlen, |target|)],
 * :551-553; engine by extend_anchor, :379-404) and scores them.  Per anchor:
 * anchor_aln (status per pair as in phx_align_batch; positions from the
 * window start), anchor_role (0 failed, 1 extended, 2 primary, 3
 * alternative) and its runs at run_off (anchors + 1 entries).  Capacities
 * below phx_extend_chains_bounds give PHX_E_CAPACITY before any work. */
int phx_extend_chains_batch(phx_ctx* ctx, int32_t strategy, int32_t gap_model,
                            const uint8_t* target, uint64_t target_len,
                            const phx_read_chains* reads, uint32_t n_reads,
                            phx_read_extension* out, phx_alignment* anchor_aln,
                            uint8_t* anchor_role, uint64_t anchors_cap, uint32_t* runs,
                            uint64_t runs_cap, uint64_t* run_off);
/* phx_extend_chains_batch against a resident reference (same results). */
int phx_extend_chains_batch_ref(phx_ctx* ctx, int32_t strategy, int32_t gap_model,
                                const phx_reference* ref, const phx_read_chains* reads,
                                uint32_t n_reads, phx_read_extension* out,
                                phx_alignment* anchor_aln, uint8_t* anchor_role,
                                uint64_t anchors_cap, uint32_t* runs, uint64_t runs_cap,
                                uint64_t* run_off);
/* score_alignments (lib.rs:170-199) over n >= 1 successful alignments with
 * their runs: *primary = the first with the smallest edit distance;
 * is_alternative[i] = 1 for every alignment not equal to the primary (all
 * fields and the CIGAR) whose (1 - ed/query_len) / (1 - ed_primary/query_len)
 * is >= 0.95.  Host only.  n = 0 (upstream panics) is PHX_E_ARG. */
int phx_score_alignments(const phx_alignment* aln, const uint32_t* runs, const uint64_t* run_off,
                         uint32_t n, uint64_t query_len, int32_t* primary, uint8_t* is_alternative);

/* --- Pre-extension: seeding and chaining -------------------------------------
 *
 * seed_and_chain (phraya-align/src/executor.rs:511-519) for a batch of query
 * minimizer sketches against one target sketch: find_seeds_indexed_capped
 * (seeding.rs:56-74; values occurring more than max_occ times in the target
 * are masked) and chain_seeds (chaining.rs:131-253), with upstream's orders
 * throughout.  Sketches are (value, position) lists as phraya_core's
 * MinimizerSketch holds them; sketching itself (simd-minimizers) is the
 * caller's. */
typedef struct phx_seed {
  uint32_t query_pos;
  uint32_t target_pos;
  uint64_t minimizer;
} phx_seed;

typedef struct phx_chain_params {  /* ChainParams (chaining.rs:41-85); defaults 21, 100, 100, 5000, 25, 0.21, 25 */
  int64_t seed_len;
  int64_t max_band;                /* at most 2^20 */
  int64_t max_dist_x;
  uint64_t max_iter;
  uint64_t max_skip;
  double gap_penalty;
  int64_t min_chain_score;
} phx_chain_params;

typedef struct phx_chain {
  int64_t score;
  uint64_t target_start;  /* Chain::target_start (chaining.rs:31-35) */
  uint64_t seed_off;      /* its seeds at seeds[seed_off .. + n_seeds], by query position */
  uint32_t n_seeds;
  uint32_t reserved;
} phx_chain;

typedef struct phx_chain_results {
  phx_chain* chains;      /* query i's at chain_off[i] .. chain_off[i + 1], score descending */
  uint64_t chains_cap;
  phx_seed* seeds;
  uint64_t seeds_cap;
  uint64_t* chain_off;    /* nq + 1 */
  uint64_t* n_seeds;      /* nq: seeds found per query (seed_and_chain's count) */
  uint64_t n_chains;      /* out: totals (set also on PHX_E_CAPACITY) */
  uint64_t n_chain_seeds;
} phx_chain_results;

typedef struct phx_seed_index phx_seed_index;
typedef struct phx_chain_plan phx_chain_plan;

/* build_minimizer_index (seeding.rs:29-35) of a target sketch, on the device. */
phx_seed_index* phx_seed_index_create(phx_ctx* ctx, const uint64_t* values, const uint32_t* positions,
                                      uint64_t n);
/* seed_occurrence_cap (seeding.rs:90-97): max(floor, 8 x median occurrence). */
uint64_t phx_seed_occurrence_cap(const phx_seed_index* index, uint64_t floor);
void phx_seed_index_destroy(phx_seed_index* index);
/* Resident form: create uploads the query sketches and lays out the working
 * arrays (a seed-count pass fixes each query's offset); run is one launch,
 * a wave per query (lookups, seed count, seeds, sorts, DP, chains); results
 * packs them per query. */
phx_chain_plan* phx_chain_plan_create(phx_ctx* ctx, const phx_seed_index* index, uint64_t max_occ,
                                      const phx_chain_params* params, const uint64_t* q_values,
                                      const uint32_t* q_positions, const uint64_t* q_off, uint32_t nq);
int phx_chain_plan_run(phx_chain_plan* plan);
int phx_chain_plan_results(phx_chain_plan* plan, phx_chain_results* out);
double phx_chain_plan_last_kernel_ms(const phx_chain_plan* plan);
uint64_t phx_chain_plan_seeds(const phx_chain_plan* plan);
void phx_chain_plan_destroy(phx_chain_plan* plan);
/* One call: create, run, results, destroy (PHX_E_ARG when creation fails). */
int phx_seed_and_chain_batch(phx_ctx* ctx, const phx_seed_index* index, uint64_t max_occ,
                             const phx_chain_params* params, const uint64_t* q_values,
                             const uint32_t* q_positions, const uint64_t* q_off, uint32_t nq,
                             phx_chain_results* out);

/* --- Minimizer sketches ------------------------------------------------------
 *
 * sketch_alphabet (phraya-core/src/types.rs:1846-1867) for a batch of
 * sequences: canonical (k, w) minimizers (canonical = 1: Alphabet::Dna, the
 * sketch / sketch_sequence entry points, types.rs:1873-1900) or forward ones
 * (canonical = 0: Alphabet::Protein), as MinimizerSketch's (value, position)
 * lists (types.rs:1801-1809).  Upstream takes them from simd-minimizers 2.3;
 * this computes the same scheme (2-bit ASCII code (b >> 1) & 3, k-mer value
 * first base most significant, canonical = min with the reverse complement's
 * value, 64-bit hash order, per-window leftmost or, for canonical windows
 * with at most half T/G, rightmost minimum, consecutive repeats reported
 * once) with its own hash, so values and positions follow that definition
 * and not the crate's bit for bit.  Sequences are concatenated in seqs,
 * sequence i at seq_off[i] .. seq_off[i + 1] (each under 2^32 bytes);
 * 1 <= k <= 32, 1 <= w <= 128, k + w - 1 odd for canonical sketches (the
 * crate's own requirement).  A sequence shorter than k + w - 1 has an empty
 * sketch. */
typedef struct phx_sketch_plan phx_sketch_plan;

/* Resident form: create uploads the sequences and lays out the tasks; run
 * is three launches (compute into per-task staging, scan of the counts,
 * copy to the final offsets); results downloads:
 * off[0 .. nseq] (sequence i's entries at off[i] .. off[i + 1]) and, when
 * cap >= the total, values / positions.  *total is set also on
 * PHX_E_CAPACITY.  capacity is an upper bound on the total (the number of
 * windows). */
phx_sketch_plan* phx_sketch_plan_create(phx_ctx* ctx, const uint8_t* seqs, const uint64_t* seq_off,
                                        uint32_t nseq, uint32_t k, uint32_t w, int32_t canonical);
int phx_sketch_plan_run(phx_sketch_plan* plan);
int phx_sketch_plan_results(phx_sketch_plan* plan, uint64_t* values, uint32_t* positions, uint64_t cap,
                            uint64_t* off, uint64_t* total);
uint64_t phx_sketch_plan_total(phx_sketch_plan* plan);
uint64_t phx_sketch_plan_capacity(const phx_sketch_plan* plan);
/* Kernel time of the last run (events around its launches), milliseconds. */
double phx_sketch_plan_last_kernel_ms(const phx_sketch_plan* plan);
/* One launch of the last run: 0 compute, 1 scan, 2 copy (milliseconds). */
double phx_sketch_plan_last_launch_ms(const phx_sketch_plan* plan, uint32_t i);
void phx_sketch_plan_destroy(phx_sketch_plan* plan);
/* One call: create, run, results, destroy. */
int phx_sketch_batch(phx_ctx* ctx, const uint8_t* seqs, const uint64_t* seq_off, uint32_t nseq, uint32_t k,
                     uint32_t w, int32_t canonical, uint64_t* values, uint32_t* positions, uint64_t cap,
                     uint64_t* off, uint64_t* total);

/* --- align_read's strand handling for a batch --------------------------------
 *
 * phraya-align/src/executor.rs:573-733 after seeding and chaining: each read
 * comes with its forward and reverse chains (the caller's seed_and_chain of
 * the read and of its reverse complement; ignored for protein); the
 * orientations worth extending (only the one with chains when the other has
 * none, otherwise both) go through one batched extend_chains; the reverse one
 * is kept only with a strictly smaller primary edit distance. */
enum { PHX_STRAND_FORWARD = 0, PHX_STRAND_REVERSE = 1 };
enum {
  PHX_OUTCOME_PLACED = 0,           /* primary score 1 - ed / max(qlen, 1) >= 0.95 */
  PHX_OUTCOME_NO_SEED = 1,          /* below 0.95 and the read had no seeds */
  PHX_OUTCOME_BELOW_THRESHOLD = 2,  /* below 0.95 with seeds */
  PHX_OUTCOME_FAST_CUTOFF = 3,      /* Fast: ed / max(qlen, 1) > 0.20 (upstream returns None) */
  PHX_OUTCOME_NO_ALIGNMENT = 4      /* no anchor extended in either orientation */
};

typedef struct phx_read_strands {
  const uint8_t* query;                   /* forward bytes as stored */
  uint64_t query_len;
  const uint64_t* fwd_chain_target_start; /* chains of the forward orientation, rank order */
  uint32_t n_fwd_chains;
  const uint64_t* rev_chain_target_start; /* chains of the reverse complement */
  uint32_t n_rev_chains;
  uint32_t had_seeds;                     /* fwd_seeds > 0 || rev_seeds > 0 */
} phx_read_strands;

typedef struct phx_read_placement {
  int32_t outcome;        /* PHX_OUTCOME_* */
  int32_t strand;         /* the kept orientation (PHX_STRAND_*), -1 when none */
  int32_t entry;          /* its extension entry in ext[], -1 when none */
  int32_t entry_fwd;      /* entries of the orientations extended, -1 when skipped */
  int32_t entry_rev;
  double primary_score;   /* 1 - ed / max(qlen, 1) of the kept primary */
} phx_read_placement;

/* reverse_complement (phraya-core/src/types.rs:91-107): A<->T, C<->G in both
 * cases, other bytes unchanged, order reversed.  Host only. */
void phx_reverse_complement(const uint8_t* seq, uint64_t n, uint8_t* out);
/* Sizes for phx_align_reads_batch: extension entries, anchors, runs. */
int phx_align_reads_bounds(int32_t strategy, int32_t is_dna, uint64_t target_len,
                           const phx_read_strands* reads, uint32_t n_reads, uint64_t* entries,
                           uint64_t* anchors, uint64_t* runs);
/* ext[]: one phx_read_extension per extended orientation (the layout of
 * phx_extend_chains_batch, anchors and runs likewise); placement[] per read.
 * PHX_E_CAPACITY when ext_cap, anchors_cap or runs_cap is short. */
int phx_align_reads_batch(phx_ctx* ctx, int32_t strategy, int32_t gap_model, int32_t is_dna,
                          const uint8_t* target, uint64_t target_len, const phx_read_strands* reads,
                          uint32_t n_reads, phx_read_placement* placement, phx_read_extension* ext,
                          uint64_t ext_cap, phx_alignment* anchor_aln, uint8_t* anchor_role,
                          uint64_t anchors_cap, uint32_t* runs, uint64_t runs_cap, uint64_t* run_off);

/* phx_align_reads_batch against a resident reference (same results). */
int phx_align_reads_batch_ref(phx_ctx* ctx, int32_t strategy, int32_t gap_model, int32_t is_dna,
                              const phx_reference* ref, const phx_read_strands* reads,
                              uint32_t n_reads, phx_read_placement* placement,
                              phx_read_extension* ext, uint64_t ext_cap, phx_alignment* anchor_aln,
                              uint8_t* anchor_role, uint64_t anchors_cap, uint32_t* runs,
                              uint64_t runs_cap, uint64_t* run_off);

/* --- Post-extension: variant sites from CIGAR runs ----------------------------
 *
 * extract_variants_from_cigar (phraya-align/src/executor.rs:814-1044) for a
 * batch of alignments of reads against one target, reading the CIGAR runs
 * directly (no ASCII): each X base is an SNV, each I run (target-only bases,
 * upstream's WFA convention) a deletion, each D run (query-only bases) an
 * insertion, under upstream's bounds checks.  Each variant carries its local
 * coverage window [p - radius (floored at 0), min(p + radius + 1, |target|))
 * with, per position, how many of the alignment's coverage windows
 * (compute_windowed_coverage, :1071-1090: one [start, end) per alignment of
 * the read, clipped to the target, empty ones dropped) cover it
 * (get_abs_multi, :218-220), and flags for membership of the sorted,
 * non-overlapping repeat and hotspot intervals (is_in_repeat_region /
 * is_in_hotspot, :795-811).  The per-read fields upstream attaches (mapq,
 * qualities, provenance, mate info, strand, confidence) pass through the
 * caller unchanged and are not part of this call. */
enum { PHX_VAR_SNV = 0, PHX_VAR_DELETION = 1, PHX_VAR_INSERTION = 2 };
enum { PHX_VAR_IN_REPEAT = 1, PHX_VAR_IN_HOTSPOT = 2 };

typedef struct phx_variant_batch {
  uint32_t n;                   /* alignments */
  const uint32_t* runs;         /* CIGAR runs (length << 4 | op), alignment i at run_off[i] .. run_off[i + 1] */
  const uint64_t* run_off;      /* n + 1 */
  const uint64_t* target_start; /* n: target position of the alignment's first base */
  const uint8_t* q_seq;         /* query bytes, alignment i at q_off[i] .. q_off[i + 1] */
  const uint64_t* q_off;        /* n + 1 */
  const uint8_t* target;        /* the target all positions index */
  uint64_t target_len;
  const uint64_t* spans;        /* coverage windows, (start, end) pairs; alignment i's at span_off[i] .. span_off[i + 1] */
  const uint64_t* span_off;     /* n + 1 (in pairs) */
  const uint64_t* repeats;      /* (start, end) pairs sorted by start, non-overlapping as upstream's */
  uint64_t n_repeats;
  const uint32_t* hotspots;     /* likewise; starts out of order are refused (PHX_E_ARG) */
  uint64_t n_hotspots;
  uint32_t radius;              /* coverage_window_radius (Fast 150, Balanced 50, Sensitive 25) */
} phx_variant_batch;

typedef struct phx_variant {
  uint32_t alignment;           /* index of its alignment in the batch */
  uint32_t position;            /* target position */
  uint8_t type;                 /* PHX_VAR_* */
  uint8_t ref_base;             /* SNV: target base; deletion: first deleted base; insertion: '.' */
  uint8_t alt_base;             /* SNV: query base; deletion: '.'; insertion: 0 (see ins_*) */
  uint8_t flags;                /* PHX_VAR_IN_REPEAT | PHX_VAR_IN_HOTSPOT */
  uint32_t ins_query_pos;       /* insertion: alleles are query[ins_query_pos .. + ins_len) */
  uint32_t ins_len;
  uint32_t window_start;        /* coverage window: values at coverage[coverage_off ..], */
  uint32_t window_len;          /*   one per target position window_start .. + window_len */
  uint32_t window_offset;       /* position - window_start */
  uint64_t coverage_off;
} phx_variant;

typedef struct phx_variant_results {
  phx_variant* variants;        /* in alignment order, then CIGAR order */
  uint64_t variants_cap;
  uint32_t* coverage;
  uint64_t coverage_cap;
  uint64_t* var_off;            /* n + 1: alignment i's variants at var_off[i] .. var_off[i + 1] */
  uint64_t n_variants;          /* out: totals (set also on PHX_E_CAPACITY) */
  uint64_t n_coverage;
} phx_variant_results;

typedef struct phx_variant_plan phx_variant_plan;

/* Resident form: create uploads the batch and lays out the outputs (a
 * counting pass and a scan), run extracts on the device (one pass: records,
 * flags, coverage windows), results downloads (checking the capacities;
 * PHX_E_CAPACITY sets the totals and copies nothing). */
phx_variant_plan* phx_variant_plan_create(phx_ctx* ctx, const phx_variant_batch* batch);
int phx_variant_plan_run(phx_variant_plan* plan);
int phx_variant_plan_results(phx_variant_plan* plan, phx_variant_results* out);
/* Kernel time of the last run (events around its launches), milliseconds. */
double phx_variant_plan_last_kernel_ms(const phx_variant_plan* plan);
void phx_variant_plan_destroy(phx_variant_plan* plan);
/* One call: create, run, results, destroy.  PHX_E_ARG for a batch that
 * plan creation would refuse (the reason in phx_ctx_last_error). */
int phx_extract_variants(phx_ctx* ctx, const phx_variant_batch* batch, phx_variant_results* out);

/* --- Output files: .phraya and .phraya.queries --------------------------------
 *
 * write_phraya (phraya-io/src/phraya.rs:67-80): the MessagePack form of
 * PhrayaFile::new(reference_length, sample_id, timestamp, observations,
 * CoverageTrack::new(coverage)) (phraya.rs:40-62, types.rs:626-651),
 * zstd-compressed at `level` (upstream: 3).  The observations are the device
 * variant records (phx_variant_results, in their order) made into
 * VariantObservations as extract_variants_from_cigar does
 * (phraya-align/src/executor.rs:814-1044): alleles {alt: 1} for an SNV,
 * {'.': 1} for a deletion, inserted-base counts for an insertion; confidence
 * 0 in a hotspot; the per-read fields from reads[record.alignment].
 * Host code; PHX_E_ARG for inconsistent inputs, PHX_E_CAPACITY (size set)
 * when cap is short, PHX_E_DEVICE when libzstd.so.1 cannot be loaded. */
typedef struct phx_phraya_header {
  uint32_t reference_length;
  const char* sample_id;          /* NUL-terminated UTF-8 */
  const char* timestamp;
} phx_phraya_header;

typedef struct phx_mate_info {    /* MateInfo (phraya-core/src/types.rs:2031-2050) */
  const char* mate_id;
  int32_t insert_size;
  uint8_t proper_pair;
  uint8_t is_first_in_pair;
  uint8_t is_second_in_pair;
  uint8_t mate_mapped;
} phx_mate_info;

typedef struct phx_obs_read {     /* extract_variants_from_cigar's per-read arguments (executor.rs:814-831) */
  const char* cigar;              /* the alignment's CIGAR string */
  const char* provenance;         /* e.g. "sample1:read42" */
  const uint8_t* query;           /* the aligned (oriented) query bytes: insertion alleles */
  uint64_t query_len;
  double avg_base_quality;
  double confidence;
  int64_t insert_size_sum;        /* insert_stats, used with has_insert_stats and a mate */
  uint32_t insert_size_count;
  uint32_t edit_distance;
  uint8_t mapq;
  uint8_t strand;                 /* PHX_STRAND_FORWARD "+", PHX_STRAND_REVERSE "-" */
  uint8_t has_insert_stats;
  uint8_t reserved;
  const phx_mate_info* mate;      /* NULL: mate_info None (the field is omitted) */
} phx_obs_read;

int phx_zstd_available(void);
/* The uncompressed MessagePack payload (what zstd decodes the file to). */
int phx_phraya_encode(const phx_phraya_header* header, const phx_variant* variants, uint64_t n_variants,
                      const uint32_t* window_coverage, uint64_t n_window_coverage, const phx_obs_read* reads,
                      uint32_t n_reads, const uint32_t* coverage, uint64_t coverage_len, uint8_t* out,
                      uint64_t cap, uint64_t* size);
int phx_phraya_write(const char* path, const phx_phraya_header* header, const phx_variant* variants,
                     uint64_t n_variants, const uint32_t* window_coverage, uint64_t n_window_coverage,
                     const phx_obs_read* reads, uint32_t n_reads, const uint32_t* coverage, uint64_t coverage_len,
                     int level);

/* write_queries (phraya-io/src/queries.rs:62-103): per read its
 * (position, score ratio) placements; kept are scores >= 0.95, reads left
 * with none are dropped, ids ascending, placements by position then score
 * descending.  Ids must be distinct (a map key each). */
typedef struct phx_query_hits {
  const char* query_id;
  const uint32_t* pos;
  const double* score;
  uint32_t n;
} phx_query_hits;
int phx_queries_encode(const phx_query_hits* queries, uint32_t n, uint8_t* out, uint64_t cap, uint64_t* size);
int phx_queries_write(const char* path, const phx_query_hits* queries, uint32_t n, int level);

/* write_cross_space_queries (queries.rs:132-167): CrossSpacePlacement
 * (space, pos, identity) lists, identity >= 0.95 kept, ordered by space,
 * position, identity descending. */
typedef struct phx_placement {
  const char* space;
  uint32_t pos;
  double identity;
} phx_placement;
typedef struct phx_cross_space_hits {
  const char* query_id;
  const phx_placement* placements;
  uint32_t n;
} phx_cross_space_hits;
int phx_cross_space_queries_encode(const phx_cross_space_hits* queries, uint32_t n, uint8_t* out, uint64_t cap,
                                   uint64_t* size);
int phx_cross_space_queries_write(const char* path, const phx_cross_space_hits* queries, uint32_t n, int level);

#ifdef __cplusplus
}
#endif
#endif
