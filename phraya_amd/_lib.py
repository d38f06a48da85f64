"""ctypes binding of libphraya_hip.so (include/phraya_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphraya_hip.so")

PHX_OK = 0
PHX_INVALID_INPUT = 1
PHX_ALIGNMENT_FAILED = 2
PHX_E_ARG = 32
PHX_E_DEVICE = 33
PHX_E_NOMEM = 34
PHX_E_CAPACITY = 35
PHX_E_INTERNAL = 36

ENGINE_LINEAR = 0
ENGINE_AFFINE = 1
ENGINE_MYERS = 2
ENGINE_MYERS_GLOBAL = 3
ENGINE_MYERS_DIST = 4

# Every symbol include/phraya_hip.h declares.
EXPORTS = (
    "phx_ctx_create", "phx_ctx_destroy", "phx_ctx_last_error", "phx_cigar_bound",
    "phx_align_batch", "phx_align_batch_noalloc", "phx_plan_create", "phx_plan_run", "phx_plan_results",
    "phx_plan_last_ms", "phx_plan_last_kernel_ms", "phx_plan_last_launches", "phx_plan_last_launch",
    "phx_plan_algorithmic_bytes", "phx_plan_destroy", "phx_format_cigar", "phx_format_cigars",
    "phx_plan_set_tuning", "phx_plan_get_tuning", "phx_plan_autotune", "phx_plan_use_graph",
    "phx_wfa_extend", "phx_wfa_extend_affine", "phx_myers_extend", "phx_myers_edit_distance",
    "phx_last_impl",
    "phx_serve_start", "phx_serve_submit", "phx_serve_poll", "phx_serve_stats_get",
    "phx_serve_stop", "phx_estimate_request", "phx_contig_anchors", "phx_contig_plan_create",
    "phx_contig_plan_run", "phx_contig_plan_results", "phx_contig_plan_gaps",
    "phx_contig_plan_layout", "phx_contig_plan_gap_plan", "phx_contig_plan_last_ms",
    "phx_contig_plan_last_gpu_ms", "phx_contig_plan_last_stitch_ms", "phx_contig_plan_destroy",
    "phx_version", "phx_chain_cap", "phx_extend_chains_bounds", "phx_extend_chains_batch",
    "phx_extend_chains_batch_ref", "phx_reference_create", "phx_reference_destroy",
    "phx_align_reads_batch_ref",
    "phx_seed_index_create", "phx_seed_occurrence_cap", "phx_seed_index_destroy", "phx_chain_plan_create",
    "phx_chain_plan_run", "phx_chain_plan_results", "phx_chain_plan_last_kernel_ms", "phx_chain_plan_seeds",
    "phx_chain_plan_destroy", "phx_seed_and_chain_batch",
    "phx_sketch_plan_create", "phx_sketch_plan_run", "phx_sketch_plan_results", "phx_sketch_plan_total",
    "phx_sketch_plan_capacity", "phx_sketch_plan_last_kernel_ms", "phx_sketch_plan_last_launch_ms",
    "phx_sketch_plan_destroy",
    "phx_sketch_batch",
    "phx_zstd_available", "phx_phraya_encode", "phx_phraya_write", "phx_queries_encode", "phx_queries_write",
    "phx_cross_space_queries_encode", "phx_cross_space_queries_write",
    "phx_score_alignments", "phx_reverse_complement", "phx_align_reads_bounds",
    "phx_align_reads_batch", "phx_variant_plan_create", "phx_variant_plan_run",
    "phx_variant_plan_results", "phx_variant_plan_last_kernel_ms", "phx_variant_plan_destroy",
    "phx_extract_variants",
)


class Costs(ctypes.Structure):
    _fields_ = [("mismatch", ctypes.c_int32), ("gap_open", ctypes.c_int32),
                ("gap_extend", ctypes.c_int32)]


class CAlignment(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("score", ctypes.c_int64), ("edit_distance", ctypes.c_int64),
                ("query_start", ctypes.c_int64), ("query_end", ctypes.c_int64),
                ("target_start", ctypes.c_int64), ("target_end", ctypes.c_int64)]


ALN_DTYPE = np.dtype([("status", np.int32), ("reserved", np.int32), ("score", np.int64),
                      ("edit_distance", np.int64), ("query_start", np.int64),
                      ("query_end", np.int64), ("target_start", np.int64),
                      ("target_end", np.int64)])
assert ALN_DTYPE.itemsize == ctypes.sizeof(CAlignment)


class CLaunchInfo(ctypes.Structure):
    _fields_ = [("tier", ctypes.c_int32), ("pairs", ctypes.c_uint32), ("grid", ctypes.c_uint32),
                ("storage", ctypes.c_uint32), ("ms", ctypes.c_double),
                ("algorithmic_bytes", ctypes.c_uint64), ("cells", ctypes.c_uint64),
                ("steps", ctypes.c_uint64)]


class CBatch(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("engine", ctypes.c_int32), ("costs", Costs),
                ("cap", ctypes.c_int64),
                ("q_seq", ctypes.c_void_p), ("q_off", ctypes.c_void_p),
                ("t_seq", ctypes.c_void_p), ("t_off", ctypes.c_void_p),
                ("seed_q", ctypes.c_void_p), ("seed_t", ctypes.c_void_p)]


class CResults(ctypes.Structure):
    _fields_ = [("aln", ctypes.c_void_p), ("runs", ctypes.c_void_p),
                ("runs_cap", ctypes.c_uint64), ("run_off", ctypes.c_void_p)]


class CReadChains(ctypes.Structure):
    _fields_ = [("query", ctypes.c_void_p), ("query_len", ctypes.c_uint64),
                ("chain_target_start", ctypes.c_void_p), ("n_chains", ctypes.c_uint32)]


class CReadExtension(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("primary", ctypes.c_int32),
                ("n_anchors", ctypes.c_uint32), ("n_alternatives", ctypes.c_uint32),
                ("first_anchor", ctypes.c_uint64)]


class CReadStrands(ctypes.Structure):
    _fields_ = [("query", ctypes.c_void_p), ("query_len", ctypes.c_uint64),
                ("fwd_chain_target_start", ctypes.c_void_p), ("n_fwd_chains", ctypes.c_uint32),
                ("rev_chain_target_start", ctypes.c_void_p), ("n_rev_chains", ctypes.c_uint32),
                ("had_seeds", ctypes.c_uint32)]


class CReadPlacement(ctypes.Structure):
    _fields_ = [("outcome", ctypes.c_int32), ("strand", ctypes.c_int32), ("entry", ctypes.c_int32),
                ("entry_fwd", ctypes.c_int32), ("entry_rev", ctypes.c_int32),
                ("primary_score", ctypes.c_double)]


class CChainParams(ctypes.Structure):
    _fields_ = [("seed_len", ctypes.c_int64), ("max_band", ctypes.c_int64),
                ("max_dist_x", ctypes.c_int64), ("max_iter", ctypes.c_uint64),
                ("max_skip", ctypes.c_uint64), ("gap_penalty", ctypes.c_double),
                ("min_chain_score", ctypes.c_int64)]


class CChainResults(ctypes.Structure):
    _fields_ = [("chains", ctypes.c_void_p), ("chains_cap", ctypes.c_uint64),
                ("seeds", ctypes.c_void_p), ("seeds_cap", ctypes.c_uint64),
                ("chain_off", ctypes.c_void_p), ("n_seeds", ctypes.c_void_p),
                ("n_chains", ctypes.c_uint64), ("n_chain_seeds", ctypes.c_uint64)]


# phx_seed / phx_chain (include/phraya_hip.h) as numpy records.
SEED_DTYPE = np.dtype([("query_pos", np.uint32), ("target_pos", np.uint32), ("minimizer", np.uint64)])
CHAIN_DTYPE = np.dtype([("score", np.int64), ("target_start", np.uint64), ("seed_off", np.uint64),
                        ("n_seeds", np.uint32), ("reserved", np.uint32)])


# phx_variant (include/phraya_hip.h) as a numpy record.
VARIANT_DTYPE = np.dtype([("alignment", np.uint32), ("position", np.uint32), ("type", np.uint8),
                          ("ref_base", np.uint8), ("alt_base", np.uint8), ("flags", np.uint8),
                          ("ins_query_pos", np.uint32), ("ins_len", np.uint32),
                          ("window_start", np.uint32), ("window_len", np.uint32),
                          ("window_offset", np.uint32), ("coverage_off", np.uint64)], align=True)


class CVariantBatch(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("runs", ctypes.c_void_p), ("run_off", ctypes.c_void_p),
                ("target_start", ctypes.c_void_p), ("q_seq", ctypes.c_void_p),
                ("q_off", ctypes.c_void_p), ("target", ctypes.c_void_p),
                ("target_len", ctypes.c_uint64), ("spans", ctypes.c_void_p),
                ("span_off", ctypes.c_void_p), ("repeats", ctypes.c_void_p),
                ("n_repeats", ctypes.c_uint64), ("hotspots", ctypes.c_void_p),
                ("n_hotspots", ctypes.c_uint64), ("radius", ctypes.c_uint32)]


class CPhrayaHeader(ctypes.Structure):
    _fields_ = [("reference_length", ctypes.c_uint32), ("sample_id", ctypes.c_char_p),
                ("timestamp", ctypes.c_char_p)]


class CMateInfo(ctypes.Structure):
    _fields_ = [("mate_id", ctypes.c_char_p), ("insert_size", ctypes.c_int32), ("proper_pair", ctypes.c_uint8),
                ("is_first_in_pair", ctypes.c_uint8), ("is_second_in_pair", ctypes.c_uint8),
                ("mate_mapped", ctypes.c_uint8)]


class CObsRead(ctypes.Structure):
    _fields_ = [("cigar", ctypes.c_char_p), ("provenance", ctypes.c_char_p), ("query", ctypes.c_void_p),
                ("query_len", ctypes.c_uint64), ("avg_base_quality", ctypes.c_double),
                ("confidence", ctypes.c_double), ("insert_size_sum", ctypes.c_int64),
                ("insert_size_count", ctypes.c_uint32), ("edit_distance", ctypes.c_uint32),
                ("mapq", ctypes.c_uint8), ("strand", ctypes.c_uint8), ("has_insert_stats", ctypes.c_uint8),
                ("reserved", ctypes.c_uint8), ("mate", ctypes.c_void_p)]


class CQueryHits(ctypes.Structure):
    _fields_ = [("query_id", ctypes.c_char_p), ("pos", ctypes.c_void_p), ("score", ctypes.c_void_p),
                ("n", ctypes.c_uint32)]


class CPlacement(ctypes.Structure):
    _fields_ = [("space", ctypes.c_char_p), ("pos", ctypes.c_uint32), ("identity", ctypes.c_double)]


class CCrossSpaceHits(ctypes.Structure):
    _fields_ = [("query_id", ctypes.c_char_p), ("placements", ctypes.c_void_p), ("n", ctypes.c_uint32)]


class CVariantResults(ctypes.Structure):
    _fields_ = [("variants", ctypes.c_void_p), ("variants_cap", ctypes.c_uint64),
                ("coverage", ctypes.c_void_p), ("coverage_cap", ctypes.c_uint64),
                ("var_off", ctypes.c_void_p), ("n_variants", ctypes.c_uint64),
                ("n_coverage", ctypes.c_uint64)]


class CTuning(ctypes.Structure):
    _fields_ = [("wave_lds_kib", ctypes.c_uint32), ("wave_slots", ctypes.c_uint32),
                ("hbm_slots", ctypes.c_uint32), ("hbm_slot_mib", ctypes.c_uint32),
                ("hbm_budget_mib", ctypes.c_uint64), ("block256_min_len", ctypes.c_uint32),
                ("block1024_min_len", ctypes.c_uint32), ("heavy_permille", ctypes.c_uint32),
                ("block_lds_ring", ctypes.c_uint32), ("block_flight", ctypes.c_uint32),
                ("block_regs", ctypes.c_uint32), ("lds_rows", ctypes.c_uint32),
                ("band_permille", ctypes.c_uint32), ("regs_narrow", ctypes.c_uint32)]


class CServeConfig(ctypes.Structure):
    _fields_ = [("engine", ctypes.c_int32), ("costs", Costs), ("cap", ctypes.c_int64),
                ("reference", ctypes.c_void_p), ("reference_len", ctypes.c_uint64),
                ("hbm_budget_mib", ctypes.c_uint64), ("max_group", ctypes.c_uint32),
                ("lanes_per_tier", ctypes.c_uint32)]


class CRequest(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("query", ctypes.c_void_p),
                ("query_len", ctypes.c_uint64), ("target", ctypes.c_void_p),
                ("target_len", ctypes.c_uint64), ("ref_start", ctypes.c_uint64),
                ("ref_len", ctypes.c_uint64), ("seed_q", ctypes.c_uint64),
                ("seed_t", ctypes.c_uint64)]


class CResponse(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("aln", CAlignment), ("latency_ms", ctypes.c_double)]


class CServeStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in
                ("submitted", "completed", "waiting", "in_flight", "launches", "reruns",
                 "arena_bytes", "arena_used_bytes", "arena_peak_bytes")]


class CContigParams(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("w", ctypes.c_uint32), ("gap_min", ctypes.c_uint32),
                ("gap_max", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


_lib = None


def lib() -> ctypes.CDLL:
    """Load libphraya_hip.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m phraya_amd.build` "
            "(the HIP path has no fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.phx_ctx_create.argtypes = [ctypes.c_int, ctypes.c_uint32, P]
    L.phx_ctx_create.restype = P
    L.phx_ctx_destroy.argtypes = [P]
    L.phx_ctx_last_error.argtypes = [P]
    L.phx_ctx_last_error.restype = ctypes.c_char_p
    L.phx_last_impl.argtypes = []
    L.phx_last_impl.restype = ctypes.c_char_p
    L.phx_cigar_bound.argtypes = [P]
    L.phx_cigar_bound.restype = ctypes.c_uint64
    L.phx_align_batch.argtypes = [P, P, P]
    L.phx_align_batch_noalloc.argtypes = [P, P, P]
    L.phx_plan_create.argtypes = [P, P]
    L.phx_plan_create.restype = P
    L.phx_plan_run.argtypes = [P]
    L.phx_plan_results.argtypes = [P, P]
    L.phx_plan_last_ms.argtypes = [P]
    L.phx_plan_last_ms.restype = ctypes.c_double
    L.phx_plan_last_kernel_ms.argtypes = [P]
    L.phx_plan_last_kernel_ms.restype = ctypes.c_double
    L.phx_plan_last_launches.argtypes = [P]
    L.phx_plan_last_launches.restype = ctypes.c_uint32
    L.phx_plan_last_launch.argtypes = [P, ctypes.c_uint32, P]
    L.phx_plan_algorithmic_bytes.argtypes = [P]
    L.phx_plan_algorithmic_bytes.restype = ctypes.c_uint64
    L.phx_plan_destroy.argtypes = [P]
    L.phx_plan_set_tuning.argtypes = [P, P]
    L.phx_plan_get_tuning.argtypes = [P, P]
    L.phx_plan_autotune.argtypes = [P, ctypes.c_uint32, P, P]
    L.phx_plan_use_graph.argtypes = [P, ctypes.c_int]
    L.phx_chain_cap.argtypes = [ctypes.c_int32]
    L.phx_chain_cap.restype = ctypes.c_uint32
    L.phx_extend_chains_bounds.argtypes = [ctypes.c_int32, ctypes.c_uint64, P, ctypes.c_uint32, P, P]
    L.phx_extend_chains_batch.argtypes = [P, ctypes.c_int32, ctypes.c_int32, P, ctypes.c_uint64, P,
                                          ctypes.c_uint32, P, P, P, ctypes.c_uint64, P,
                                          ctypes.c_uint64, P]
    L.phx_extend_chains_batch_ref.argtypes = [P, ctypes.c_int32, ctypes.c_int32, P, P,
                                              ctypes.c_uint32, P, P, P, ctypes.c_uint64, P,
                                              ctypes.c_uint64, P]
    L.phx_score_alignments.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_uint64, P, P]
    L.phx_seed_index_create.argtypes = [P, P, P, ctypes.c_uint64]
    L.phx_seed_index_create.restype = P
    L.phx_seed_occurrence_cap.argtypes = [P, ctypes.c_uint64]
    L.phx_seed_occurrence_cap.restype = ctypes.c_uint64
    L.phx_seed_index_destroy.argtypes = [P]
    L.phx_seed_index_destroy.restype = None
    L.phx_chain_plan_create.argtypes = [P, P, ctypes.c_uint64, P, P, P, P, ctypes.c_uint32]
    L.phx_chain_plan_create.restype = P
    L.phx_chain_plan_run.argtypes = [P]
    L.phx_chain_plan_results.argtypes = [P, P]
    L.phx_chain_plan_last_kernel_ms.argtypes = [P]
    L.phx_chain_plan_last_kernel_ms.restype = ctypes.c_double
    L.phx_chain_plan_seeds.argtypes = [P]
    L.phx_chain_plan_seeds.restype = ctypes.c_uint64
    L.phx_chain_plan_destroy.argtypes = [P]
    L.phx_chain_plan_destroy.restype = None
    L.phx_seed_and_chain_batch.argtypes = [P, P, ctypes.c_uint64, P, P, P, P, ctypes.c_uint32, P]
    L.phx_sketch_plan_create.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_int32]
    L.phx_sketch_plan_create.restype = P
    L.phx_sketch_plan_run.argtypes = [P]
    L.phx_sketch_plan_results.argtypes = [P, P, P, ctypes.c_uint64, P, P]
    L.phx_sketch_plan_total.argtypes = [P]
    L.phx_sketch_plan_total.restype = ctypes.c_uint64
    L.phx_sketch_plan_capacity.argtypes = [P]
    L.phx_sketch_plan_capacity.restype = ctypes.c_uint64
    L.phx_sketch_plan_last_kernel_ms.argtypes = [P]
    L.phx_sketch_plan_last_kernel_ms.restype = ctypes.c_double
    L.phx_sketch_plan_last_launch_ms.argtypes = [P, ctypes.c_uint32]
    L.phx_sketch_plan_last_launch_ms.restype = ctypes.c_double
    L.phx_sketch_plan_destroy.argtypes = [P]
    L.phx_sketch_plan_destroy.restype = None
    L.phx_sketch_batch.argtypes = [P, P, P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                   P, P, ctypes.c_uint64, P, P]
    L.phx_zstd_available.argtypes = []
    L.phx_phraya_encode.argtypes = [P, P, ctypes.c_uint64, P, ctypes.c_uint64, P, ctypes.c_uint32, P,
                                    ctypes.c_uint64, P, ctypes.c_uint64, P]
    L.phx_phraya_write.argtypes = [ctypes.c_char_p, P, P, ctypes.c_uint64, P, ctypes.c_uint64, P, ctypes.c_uint32,
                                   P, ctypes.c_uint64, ctypes.c_int]
    L.phx_queries_encode.argtypes = [P, ctypes.c_uint32, P, ctypes.c_uint64, P]
    L.phx_queries_write.argtypes = [ctypes.c_char_p, P, ctypes.c_uint32, ctypes.c_int]
    L.phx_cross_space_queries_encode.argtypes = [P, ctypes.c_uint32, P, ctypes.c_uint64, P]
    L.phx_cross_space_queries_write.argtypes = [ctypes.c_char_p, P, ctypes.c_uint32, ctypes.c_int]
    L.phx_reverse_complement.argtypes = [P, ctypes.c_uint64, P]
    L.phx_reverse_complement.restype = None
    L.phx_align_reads_bounds.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, P,
                                         ctypes.c_uint32, P, P, P]
    L.phx_align_reads_batch.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P,
                                        ctypes.c_uint64, P, ctypes.c_uint32, P, P, ctypes.c_uint64,
                                        P, P, ctypes.c_uint64, P, ctypes.c_uint64, P]
    L.phx_align_reads_batch_ref.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P, P,
                                            ctypes.c_uint32, P, P, ctypes.c_uint64, P, P,
                                            ctypes.c_uint64, P, ctypes.c_uint64, P]
    L.phx_reference_create.argtypes = [P, P, ctypes.c_uint64]
    L.phx_reference_create.restype = P
    L.phx_reference_destroy.argtypes = [P]
    L.phx_variant_plan_create.argtypes = [P, P]
    L.phx_variant_plan_create.restype = P
    L.phx_variant_plan_run.argtypes = [P]
    L.phx_variant_plan_results.argtypes = [P, P]
    L.phx_variant_plan_last_kernel_ms.argtypes = [P]
    L.phx_variant_plan_last_kernel_ms.restype = ctypes.c_double
    L.phx_variant_plan_destroy.argtypes = [P]
    L.phx_extract_variants.argtypes = [P, P, P]
    L.phx_format_cigar.argtypes = [P, ctypes.c_uint64, P, ctypes.c_size_t]
    L.phx_format_cigar.restype = ctypes.c_size_t
    L.phx_format_cigars.argtypes = [P, P, ctypes.c_uint64, P, ctypes.c_uint64, P]
    L.phx_format_cigars.restype = ctypes.c_uint64
    sz, u64 = ctypes.c_size_t, ctypes.c_uint64
    L.phx_wfa_extend.argtypes = [P, P, sz, P, sz, u64, u64, P, P, sz, P]
    L.phx_wfa_extend_affine.argtypes = [P, P, sz, P, sz, u64, u64, Costs, P, P, sz, P]
    L.phx_myers_extend.argtypes = [P, P, sz, P, sz, u64, u64, P, P, sz, P]
    L.phx_myers_edit_distance.argtypes = [P, P, sz, P, sz, P, P, sz, P]
    L.phx_serve_start.argtypes = [P, P]
    L.phx_serve_start.restype = P
    L.phx_serve_submit.argtypes = [P, P, ctypes.c_uint32]
    L.phx_serve_poll.argtypes = [P, P, ctypes.c_uint32, P, u64, P, P, ctypes.c_int32]
    L.phx_serve_stats_get.argtypes = [P, P]
    L.phx_serve_stop.argtypes = [P]
    L.phx_estimate_request.argtypes = [ctypes.c_int32, Costs, P, u64, P, u64, P]
    L.phx_estimate_request.restype = ctypes.c_int64
    L.phx_contig_anchors.argtypes = [P, u64, P, u64, P, u64, P, u64]
    L.phx_contig_anchors.restype = ctypes.c_int64
    L.phx_contig_plan_create.argtypes = [P, ctypes.c_int32, Costs, P, P]
    L.phx_contig_plan_create.restype = P
    L.phx_contig_plan_run.argtypes = [P]
    L.phx_contig_plan_results.argtypes = [P, P]
    L.phx_contig_plan_gaps.argtypes = [P]
    L.phx_contig_plan_gaps.restype = u64
    L.phx_contig_plan_layout.argtypes = [P, ctypes.c_uint32, P, u64]
    L.phx_contig_plan_layout.restype = ctypes.c_int64
    L.phx_contig_plan_gap_plan.argtypes = [P]
    L.phx_contig_plan_gap_plan.restype = P
    for f in ("phx_contig_plan_last_ms", "phx_contig_plan_last_gpu_ms",
              "phx_contig_plan_last_stitch_ms"):
        getattr(L, f).argtypes = [P]
        getattr(L, f).restype = ctypes.c_double
    L.phx_contig_plan_destroy.argtypes = [P]
    L.phx_version.restype = ctypes.c_char_p
    _lib = L
    return L
