"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, and there only as the checker or the timed CPU
baseline.  The product package (phraya_amd) never imports this module.

The oracle restates phraya-align's engines (phraya-align/src/wfa_simd.rs,
phraya-align/src/lib.rs); see phraya_oracle.cpp for the per-function
citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OK, INVALID_INPUT, ALIGNMENT_FAILED, BAD_COSTS = 0, 1, 2, 3
LINEAR, AFFINE, MYERS, MYERS_GLOBAL, MYERS_FWD, LINEAR_GLOBAL, AFFINE_GLOBAL = 0, 1, 2, 3, 4, 5, 6


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("score", ctypes.c_int64),
        ("edit_distance", ctypes.c_int64),
        ("query_start", ctypes.c_int64),
        ("query_end", ctypes.c_int64),
        ("target_start", ctypes.c_int64),
        ("target_end", ctypes.c_int64),
        ("cigar_len", ctypes.c_int64),
    ]


RESULT_DTYPE = np.dtype(
    [
        ("status", np.int32),
        ("reserved", np.int32),
        ("score", np.int64),
        ("edit_distance", np.int64),
        ("query_start", np.int64),
        ("query_end", np.int64),
        ("target_start", np.int64),
        ("target_end", np.int64),
        ("cigar_len", np.int64),
    ]
)
assert RESULT_DTYPE.itemsize == ctypes.sizeof(OrcResult)


def build() -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        sz = ctypes.c_size_t
        i32, i64, u32 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32
        L.orc_lcp.argtypes = [P, sz, P, sz]
        L.orc_lcp.restype = sz
        L.orc_lcp_scalar.argtypes = [P, sz, P, sz]
        L.orc_lcp_scalar.restype = sz
        L.orc_lcp_u64.argtypes = [P, sz, P, sz]
        L.orc_lcp_u64.restype = sz
        R = ctypes.POINTER(OrcResult)
        L.orc_wfa_extend.argtypes = [P, sz, P, sz, sz, sz, i64, R, P, sz]
        L.orc_wfa_extend_affine.argtypes = [P, sz, P, sz, sz, sz, i32, i32, i32, i64, R, P, sz]
        L.orc_myers_extend.argtypes = [P, sz, P, sz, sz, sz, R, P, sz]
        L.orc_wfa_global.argtypes = [P, sz, P, sz, R, P, sz]
        L.orc_wfa_affine_global.argtypes = [P, sz, P, sz, i32, i32, i32, R, P, sz]
        L.orc_myers_edit_distance.argtypes = [P, sz, P, sz, R, P, sz]
        L.orc_myers_extend_distance.argtypes = [P, sz, P, sz, sz, sz, R]
        L.orc_batch_run.argtypes = [ctypes.c_int, u32, P, P, P, P, P, P, i32, i32, i32, i64,
                                    ctypes.c_int, P]
        L.orc_batch_run.restype = P
        L.orc_batch_cigar_bytes.argtypes = [P]
        L.orc_batch_cigar_bytes.restype = ctypes.c_uint64
        L.orc_batch_cigar_copy.argtypes = [P, P, P]
        L.orc_batch_free.argtypes = [P]
        L.orc_hardware_threads.restype = ctypes.c_int
        u64 = ctypes.c_uint64
        L.orc_extract_variants.argtypes = [u32, P, P, P, P, P, P, u64, P, P, P, u64, P, u64, u32,
                                           ctypes.c_int, P, u64, P, u64, P, P, P]
        L.orc_extract_variants.restype = ctypes.c_int
        L.orc_seed_occurrence_cap.argtypes = [P, P, u64, u64]
        L.orc_seed_occurrence_cap.restype = ctypes.c_uint64
        L.orc_seed_and_chain.argtypes = [P, P, u64, u64, P, P, P, u32, P, ctypes.c_int, P, P, u64,
                                         P, P, u64, P, P]
        L.orc_seed_and_chain.restype = ctypes.c_int
        L.orc_seed_and_chain_timed.argtypes = [P, P, u64, u64, P, P, P, u32, P, ctypes.c_int,
                                               ctypes.c_int, P]
        L.orc_seed_and_chain_timed.restype = ctypes.c_double
        L.orc_sketch.argtypes = [P, u64, u32, u32, ctypes.c_int, P, P]
        L.orc_sketch.restype = i64
        L.orc_sketch_rolling.argtypes = [P, u64, u32, u32, ctypes.c_int, P, P]
        L.orc_sketch_rolling.restype = i64
        L.orc_sketch_timed.argtypes = [P, P, u32, u32, u32, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
        L.orc_sketch_timed.restype = ctypes.c_double
        _lib = L
    return _lib


@dataclass
class Result:
    status: int
    cigar: str
    score: int
    edit_distance: int
    query_start: int
    query_end: int
    target_start: int
    target_end: int


def _buf(b: bytes):
    return ctypes.c_char_p(b) if b else ctypes.c_char_p(b"\0")


def _call(fn, *args) -> Result:
    cap = 1 << 12
    while True:
        out = ctypes.create_string_buffer(cap)
        r = OrcResult()
        fn(*args, ctypes.byref(r), out, cap)
        if r.status != OK or r.cigar_len < cap:
            return Result(r.status, out.value.decode() if r.status == OK else "", r.score,
                          r.edit_distance, r.query_start, r.query_end, r.target_start, r.target_end)
        cap = int(r.cigar_len) + 1


def wfa_extend(q: bytes, t: bytes, seed=(0, 0), cap: int = -1) -> Result:
    return _call(lib().orc_wfa_extend, _buf(q), len(q), _buf(t), len(t), seed[0], seed[1], cap)


def wfa_extend_affine(q: bytes, t: bytes, seed=(0, 0), costs=(2, 4, 1), cap: int = -1) -> Result:
    x, o, e = costs
    return _call(lib().orc_wfa_extend_affine, _buf(q), len(q), _buf(t), len(t), seed[0], seed[1],
                 x, o, e, cap)


def wfa_global(q: bytes, t: bytes) -> Result:
    """fill_wfa (wfa_simd.rs:231-345): the global branch whatever the lengths."""
    return _call(lib().orc_wfa_global, _buf(q), len(q), _buf(t), len(t))


def wfa_affine_global(q: bytes, t: bytes, costs=(2, 4, 1)) -> Result:
    """fill_wfa_affine_generic with fitting = false (wfa_simd.rs:581-779)."""
    x, o, e = costs
    return _call(lib().orc_wfa_affine_global, _buf(q), len(q), _buf(t), len(t), x, o, e)


def myers_extend(q: bytes, t: bytes, seed=(0, 0)) -> Result:
    return _call(lib().orc_myers_extend, _buf(q), len(q), _buf(t), len(t), seed[0], seed[1])


def myers_edit_distance(q: bytes, t: bytes) -> tuple[int, str]:
    r = _call(lib().orc_myers_edit_distance, _buf(q), len(q), _buf(t), len(t))
    return r.edit_distance, r.cigar


def lcp(a: bytes, b: bytes, tier: str = "sse2") -> int:
    fn = {"sse2": lib().orc_lcp, "scalar": lib().orc_lcp_scalar, "u64": lib().orc_lcp_u64}[tier]
    return fn(_buf(a), len(a), _buf(b), len(b))


def hardware_threads() -> int:
    return lib().orc_hardware_threads()


def batch(engine: int, q_seq: np.ndarray, q_off: np.ndarray, t_seq: np.ndarray, t_off: np.ndarray,
          seed_q=None, seed_t=None, costs=(2, 4, 1), cap: int = -1, threads: int = 1,
          want_cigars: bool = True):
    """Run the oracle over a concatenated batch.  Returns (results, cigars or None)."""
    L = lib()
    n = len(q_off) - 1
    q_seq = np.ascontiguousarray(q_seq, dtype=np.uint8)
    t_seq = np.ascontiguousarray(t_seq, dtype=np.uint8)
    q_off = np.ascontiguousarray(q_off, dtype=np.uint64)
    t_off = np.ascontiguousarray(t_off, dtype=np.uint64)
    sq = None if seed_q is None else np.ascontiguousarray(seed_q, dtype=np.uint64)
    st = None if seed_t is None else np.ascontiguousarray(seed_t, dtype=np.uint64)
    res = np.zeros(n, dtype=RESULT_DTYPE)
    x, o, e = costs
    h = L.orc_batch_run(engine, n, q_seq.ctypes.data, q_off.ctypes.data, t_seq.ctypes.data,
                        t_off.ctypes.data, None if sq is None else sq.ctypes.data,
                        None if st is None else st.ctypes.data, x, o, e, cap, threads,
                        res.ctypes.data)
    try:
        cigars = None
        if want_cigars:
            total = L.orc_batch_cigar_bytes(h)
            arena = np.zeros(max(int(total), 1), dtype=np.uint8)
            off = np.zeros(n + 1, dtype=np.uint64)
            L.orc_batch_cigar_copy(h, arena.ctypes.data, off.ctypes.data)
            raw = arena.tobytes()
            cigars = [raw[int(off[i]):int(off[i + 1])].decode() for i in range(n)]
    finally:
        L.orc_batch_free(h)
    return res, cigars


# Record layout of orc_variant (the same as phx_variant in the C ABI).
VARIANT_DTYPE = np.dtype([("alignment", np.uint32), ("position", np.uint32), ("type", np.uint8),
                          ("ref_base", np.uint8), ("alt_base", np.uint8), ("flags", np.uint8),
                          ("ins_query_pos", np.uint32), ("ins_len", np.uint32),
                          ("window_start", np.uint32), ("window_len", np.uint32),
                          ("window_offset", np.uint32), ("coverage_off", np.uint64)], align=True)


def extract_variants(runs, run_off, target_start, q_seq, q_off, target, spans, span_off,
                     repeats=None, hotspots=None, radius=25, threads=None):
    """extract_variants_from_cigar over a batch (executor.rs:814-1044); see
    orc_extract_variants.  Returns (variants, var_off, coverage) with the
    records as VARIANT_DTYPE."""
    L = lib()
    n = len(target_start)
    runs = np.ascontiguousarray(runs, np.uint32)
    run_off = np.ascontiguousarray(run_off, np.uint64)
    target_start = np.ascontiguousarray(target_start, np.uint64)
    q_seq = np.ascontiguousarray(q_seq, np.uint8)
    q_off = np.ascontiguousarray(q_off, np.uint64)
    target = np.frombuffer(bytes(target), np.uint8) if not isinstance(target, np.ndarray) else \
        np.ascontiguousarray(target, np.uint8)
    spans = np.ascontiguousarray(spans, np.uint64).reshape(-1)
    span_off = np.ascontiguousarray(span_off, np.uint64)
    rep = np.ascontiguousarray(np.zeros(0, np.uint64) if repeats is None else repeats, np.uint64).reshape(-1)
    hot = np.ascontiguousarray(np.zeros(0, np.uint32) if hotspots is None else hotspots, np.uint32).reshape(-1)
    threads = threads or hardware_threads()
    nv = ctypes.c_uint64(0)
    nc = ctypes.c_uint64(0)
    var_off = np.zeros(n + 1, np.uint64)
    out = np.zeros(1, VARIANT_DTYPE)
    cov = np.zeros(1, np.uint32)
    for _ in range(2):
        rc = L.orc_extract_variants(
            n, runs.ctypes.data, run_off.ctypes.data, target_start.ctypes.data, q_seq.ctypes.data,
            q_off.ctypes.data, target.ctypes.data, len(target), spans.ctypes.data,
            span_off.ctypes.data, rep.ctypes.data, len(rep) // 2, hot.ctypes.data, len(hot) // 2,
            radius, threads, out.ctypes.data, len(out), cov.ctypes.data, len(cov),
            var_off.ctypes.data, ctypes.byref(nv), ctypes.byref(nc))
        if rc == 0:
            return out[:nv.value], var_off, cov[:nc.value]
        out = np.zeros(max(1, nv.value), VARIANT_DTYPE)
        cov = np.zeros(max(1, nc.value), np.uint32)
    raise RuntimeError("orc_extract_variants: capacity")


# --- pre-extension: seeding and chaining (seeding.rs, chaining.rs) ----------

SEED_DTYPE = np.dtype([("query_pos", np.uint32), ("target_pos", np.uint32), ("minimizer", np.uint64)])
CHAIN_DTYPE = np.dtype([("score", np.int64), ("target_start", np.uint64), ("seed_off", np.uint64),
                        ("n_seeds", np.uint32), ("reserved", np.uint32)])


class OrcChainParams(ctypes.Structure):
    _fields_ = [("seed_len", ctypes.c_int64), ("max_band", ctypes.c_int64),
                ("max_dist_x", ctypes.c_int64), ("max_iter", ctypes.c_uint64),
                ("max_skip", ctypes.c_uint64), ("gap_penalty", ctypes.c_double),
                ("min_chain_score", ctypes.c_int64)]


def chain_params(**kw) -> OrcChainParams:
    """ChainParams::default() (chaining.rs:69-85) with overrides."""
    d = dict(seed_len=21, max_band=100, max_dist_x=100, max_iter=5000, max_skip=25,
             gap_penalty=0.21, min_chain_score=25)
    d.update(kw)
    return OrcChainParams(**d)


def _sketch(mins):
    a = np.asarray(mins, dtype=np.uint64).reshape(-1, 2) if len(mins) else np.zeros((0, 2), np.uint64)
    return np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1].astype(np.uint32))


def seed_occurrence_cap(target_mins, floor: int) -> int:
    """seed_occurrence_cap (seeding.rs:90-97) of the index built from a
    target sketch [(value, position), ...]."""
    v, p = _sketch(target_mins)
    return int(lib().orc_seed_occurrence_cap(v.ctypes.data, p.ctypes.data, len(v), floor))


def seed_and_chain(target_mins, queries, max_occ=2**64 - 1, params=None, threads=None):
    """seed_and_chain (executor.rs:511-519) per query sketch: returns
    (n_seeds, chains, chain_off, seeds) with chains as CHAIN_DTYPE (query i's
    at chain_off[i] .. chain_off[i + 1], score descending) and their seeds
    as SEED_DTYPE."""
    L = lib()
    tv, tp = _sketch(target_mins)
    qv = np.concatenate([_sketch(q)[0] for q in queries] + [np.zeros(0, np.uint64)])
    qp = np.concatenate([_sketch(q)[1] for q in queries] + [np.zeros(0, np.uint32)])
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.uint64)
    nq = len(queries)
    params = params or chain_params()
    threads = threads or hardware_threads()
    n_seeds = np.zeros(max(nq, 1), np.uint64)
    chain_off = np.zeros(nq + 1, np.uint64)
    tc, ts = ctypes.c_uint64(0), ctypes.c_uint64(0)
    chains = np.zeros(1, CHAIN_DTYPE)
    seeds = np.zeros(1, SEED_DTYPE)
    for _ in range(2):
        rc = L.orc_seed_and_chain(tv.ctypes.data, tp.ctypes.data, len(tv), max_occ, qv.ctypes.data,
                                  qp.ctypes.data, q_off.ctypes.data, nq, ctypes.byref(params), threads,
                                  n_seeds.ctypes.data, chains.ctypes.data, len(chains),
                                  chain_off.ctypes.data, seeds.ctypes.data, len(seeds),
                                  ctypes.byref(tc), ctypes.byref(ts))
        if rc == 0:
            return n_seeds[:nq], chains[:tc.value], chain_off, seeds[:ts.value]
        chains = np.zeros(max(1, tc.value), CHAIN_DTYPE)
        seeds = np.zeros(max(1, ts.value), SEED_DTYPE)
    raise RuntimeError("orc_seed_and_chain: capacity")


def seed_and_chain_timed(target_mins, queries, max_occ=2**64 - 1, params=None, threads=None, passes=1):
    """Seconds for `passes` rounds of seed_and_chain over the queries, the
    index built once beforehand (CPU baseline)."""
    tv, tp = _sketch(target_mins)
    qv = np.concatenate([_sketch(q)[0] for q in queries] + [np.zeros(0, np.uint64)])
    qp = np.concatenate([_sketch(q)[1] for q in queries] + [np.zeros(0, np.uint32)])
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.uint64)
    params = params or chain_params()
    chk = ctypes.c_uint64(0)
    return float(lib().orc_seed_and_chain_timed(tv.ctypes.data, tp.ctypes.data, len(tv), max_occ,
                                                qv.ctypes.data, qp.ctypes.data, q_off.ctypes.data,
                                                len(queries), ctypes.byref(params),
                                                threads or hardware_threads(), passes, ctypes.byref(chk)))


def sketch(seq, k: int = 21, w: int = 11, canonical: bool = True, rolling: bool = False):
    """Minimizer sketch (sketch_alphabet, phraya-core/src/types.rs:1846-1867)
    of one byte sequence: (values u64, positions u32) in window order; the
    scheme is stated in phraya_oracle.cpp.  rolling: the CPU baseline's
    implementation of the same scheme."""
    s = np.frombuffer(bytes(seq), np.uint8) if not isinstance(seq, np.ndarray) else np.ascontiguousarray(seq, np.uint8)
    cap = max(1, len(s) - k - w + 2)
    vals = np.zeros(cap, np.uint64)
    pos = np.zeros(cap, np.uint32)
    fn = lib().orc_sketch_rolling if rolling else lib().orc_sketch
    n = fn(s.ctypes.data, len(s), k, w, 1 if canonical else 0, vals.ctypes.data, pos.ctypes.data)
    if n < 0:
        raise ValueError("orc_sketch: invalid k / w")
    return vals[:n], pos[:n]


def sketch_timed(seqs, k: int = 21, w: int = 11, canonical: bool = True, threads=None, passes: int = 1) -> float:
    """Seconds for `passes` rounds of sketching every sequence (CPU baseline,
    the rolling implementation, sequences split over threads)."""
    arrs = [np.frombuffer(bytes(s), np.uint8) if not isinstance(s, np.ndarray) else np.ascontiguousarray(s, np.uint8)
            for s in seqs]
    buf = np.ascontiguousarray(np.concatenate(arrs + [np.zeros(1, np.uint8)]))
    off = np.concatenate([[0], np.cumsum([len(a) for a in arrs])]).astype(np.uint64)
    chk = ctypes.c_uint64(0)
    return float(lib().orc_sketch_timed(buf.ctypes.data, off.ctypes.data, len(arrs), k, w, 1 if canonical else 0,
                                        threads or hardware_threads(), passes, ctypes.byref(chk)))
