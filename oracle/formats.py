"""Test infrastructure (not the product): the .phraya / .phraya.queries
payloads restated as plain Python values, encoded with msgpack-python (an
independent MessagePack encoder) and decoded with libzstd through ctypes.

Only tests import this module.  Each builder follows the Rust declaration
its output mirrors; rmp_serde::to_vec writes a struct as an array of its
fields in declaration order, so the builders return lists.
"""
from __future__ import annotations

import ctypes

import msgpack
import numpy as np

PHRAYA_VERSION = 1        # phraya-io/src/phraya.rs:6
SCORE_THRESHOLD = 0.95    # phraya-io/src/queries.rs:63, :137
SNV, DELETION, INSERTION = 0, 1, 2
IN_REPEAT, IN_HOTSPOT = 1, 2


def quantize(v: int) -> int:
    """CoverageTrack::quantize (phraya-core/src/types.rs:655-658)."""
    return min((v + 2) // 5 * 5, 255)


def coverage_track(cov) -> list:
    """CoverageTrack::new (types.rs:626-651) as [runs, total_length]."""
    runs, cur, n = [], 0, 0
    for c in cov:
        q = quantize(int(c))
        if q == cur:
            n += 1
        else:
            if n:
                runs.append([cur, n])
            cur, n = q, 1
    if n:
        runs.append([cur, n])
    return [runs, len(cov)]


def observation(rec, window, read) -> list:
    """One VariantObservation (types.rs:217-287) as extract_variants_from_cigar
    builds it (phraya-align/src/executor.rs:845-1036) from a device variant
    record, its coverage window and the read's fields (a dict)."""
    t = int(rec["type"])
    if t == INSERTION:
        ins = read["query"][int(rec["ins_query_pos"]):int(rec["ins_query_pos"]) + int(rec["ins_len"])]
        alleles = {}
        for b in ins:
            alleles[b] = alleles.get(b, 0) + 1
    elif t == DELETION:
        alleles = {ord("."): 1}
    else:
        alleles = {int(rec["alt_base"]): 1}
    alleles = dict(sorted(alleles.items()))                 # serialize_alleles_sorted (types.rs:17-23)
    ku = 0.0 if int(rec["flags"]) & IN_HOTSPOT else 1.0     # is_in_hotspot (executor.rs:868-872)
    mate = read.get("mate")
    stats = read.get("insert_stats")
    obs = [
        int(rec["position"]),
        int(rec["ref_base"]),
        alleles,
        ku * read.get("confidence", 1.0),
        read["cigar"],
        read.get("mapq", 60),
        read.get("edit_distance", 0),
        [int(x) for x in window],
        read.get("avg_base_quality", 0.0),
        read["provenance"],
        bool(int(rec["flags"]) & IN_REPEAT),
        ["snp", "deletion", "insertion"][t],
        ku,
        "-" if read.get("strand", 0) else "+",
        1 if mate else 0,                                    # with_pair_counts (executor.rs:897-900)
        1 if mate and mate.get("proper_pair") else 0,
        stats[0] if mate and stats else 0,                   # with_insert_stats (:901-903)
        stats[1] if mate and stats else 0,
        int(rec["window_offset"]),
    ]
    if mate:
        obs.append([mate["mate_id"], bool(mate.get("proper_pair", False)), mate.get("insert_size", 0),
                    bool(mate.get("is_first_in_pair", False)), bool(mate.get("is_second_in_pair", False)),
                    bool(mate.get("mate_mapped", False))])   # MateInfo (types.rs:2031-2050)
    return obs


def phraya_file(reference_length, sample_id, timestamp, variants, window_coverage, reads, coverage) -> list:
    """PhrayaFile::new (phraya-io/src/phraya.rs:40-62): [header, observations,
    coverage_track], header = [version, reference_length, sample_id,
    timestamp, observation_count] (:22-30)."""
    obs = []
    for rec in variants:
        a, n = int(rec["coverage_off"]), int(rec["window_len"])
        obs.append(observation(rec, window_coverage[a:a + n], reads[int(rec["alignment"])]))
    return [[PHRAYA_VERSION, reference_length, sample_id, timestamp, len(obs)], obs, coverage_track(coverage)]


def queries(index: dict) -> dict:
    """write_queries' filtered, ordered map (queries.rs:62-92)."""
    out = {}
    for qid in sorted(index, key=lambda s: s.encode()):
        kept = [(p, s) for p, s in index[qid] if s >= SCORE_THRESHOLD]
        if not kept:
            continue
        kept = sorted(kept, key=lambda x: x[1], reverse=True)   # stable: score descending within a position
        kept = sorted(kept, key=lambda x: x[0])
        out[qid] = [[p, s] for p, s in kept]
    return out


def cross_space_queries(index: dict) -> dict:
    """write_cross_space_queries' map (queries.rs:132-162)."""
    out = {}
    for qid in sorted(index, key=lambda s: s.encode()):
        kept = [x for x in index[qid] if x[2] >= SCORE_THRESHOLD]
        if not kept:
            continue
        kept = sorted(kept, key=lambda x: x[2], reverse=True)
        kept = sorted(kept, key=lambda x: (x[0].encode(), x[1]))
        out[qid] = [[s, p, i] for s, p, i in kept]
    return out


def pack(value) -> bytes:
    return msgpack.packb(value, use_bin_type=True)


def unpack(data: bytes):
    return msgpack.unpackb(data, raw=False, strict_map_key=False)


_zstd = None


def zstd_decompress(data: bytes) -> bytes:
    """One zstd frame with its content size (ZSTD_compress writes it)."""
    global _zstd
    if _zstd is None:
        _zstd = ctypes.CDLL("libzstd.so.1")
        _zstd.ZSTD_getFrameContentSize.restype = ctypes.c_ulonglong
        _zstd.ZSTD_getFrameContentSize.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        _zstd.ZSTD_decompress.restype = ctypes.c_size_t
        _zstd.ZSTD_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
        _zstd.ZSTD_isError.argtypes = [ctypes.c_size_t]
    src = np.frombuffer(data, np.uint8)
    n = _zstd.ZSTD_getFrameContentSize(src.ctypes.data, len(src))
    if n >= (1 << 63):
        raise ValueError("zstd frame without a content size")
    dst = np.zeros(max(1, n), np.uint8)
    got = _zstd.ZSTD_decompress(dst.ctypes.data, len(dst), src.ctypes.data, len(src))
    if _zstd.ZSTD_isError(got) or got != n:
        raise ValueError("zstd decompression failed")
    return dst[:n].tobytes()
