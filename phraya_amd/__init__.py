"""phraya_amd -- MI355X (gfx950) extension engines of phraya-align.

Host mirror of phraya-align's extension API (phraya-align/src/lib.rs:23-167)
over the C ABI of libphraya_hip.so:

    wfa_extend(query, target, seed)               lib.rs:74-76
    wfa_extend_naive(query, target, seed)         lib.rs:63-65
    wfa_extend_affine(query, target, seed, costs) lib.rs:139-167
    myers_extend(query, target, seed)             lib.rs:106-127
    myers_edit_distance(query, target)            lib.rs:94-96

Each returns what the upstream returns (an Alignment, or raises WfaError's
InvalidInput / AlignmentFailed where the upstream returns Err), computed on
the GPU.  align_batch() is the batched form used by a batched executor; Plan
keeps a batch resident in HBM for repeated launches; Server serves requests
as they arrive (continuous batching against a resident reference).
"""
from __future__ import annotations

import atexit
import ctypes
import threading
import weakref
from dataclasses import dataclass
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _lib as L

__all__ = [
    "SeedAnchor", "Alignment", "WfaError", "InvalidInput", "AlignmentFailed", "AffineCosts",
    "Context", "wfa_extend", "wfa_extend_naive", "wfa_extend_affine", "myers_extend",
    "myers_edit_distance", "align_batch", "BatchResult", "Plan", "pack_batch", "format_cigar",
    "ENGINE_LINEAR", "ENGINE_AFFINE", "ENGINE_MYERS", "ENGINE_MYERS_GLOBAL", "ENGINE_MYERS_DIST",
    "Server", "Response", "ContigParams", "ContigPlan", "contig_anchors",
]

ENGINE_LINEAR = L.ENGINE_LINEAR
ENGINE_AFFINE = L.ENGINE_AFFINE
ENGINE_MYERS = L.ENGINE_MYERS
ENGINE_MYERS_GLOBAL = L.ENGINE_MYERS_GLOBAL
ENGINE_MYERS_DIST = L.ENGINE_MYERS_DIST

_OPS = "MXID"


@dataclass(frozen=True)
class SeedAnchor:
    """SeedAnchor (phraya-align/src/lib.rs:24-28)."""
    query_pos: int = 0
    target_pos: int = 0


@dataclass
class Alignment:
    """Alignment (phraya-align/src/lib.rs:31-39); `score` is the engine's search score."""
    cigar: str
    edit_distance: int
    query_start: int
    query_end: int
    target_start: int
    target_end: int
    score: int = 0


class WfaError(Exception):
    """WfaError (phraya-align/src/lib.rs:45-49)."""


class InvalidInput(WfaError):
    pass


class AlignmentFailed(WfaError):
    pass


@dataclass(frozen=True)
class AffineCosts:
    """AffineCosts (phraya-align/src/wfa_simd.rs:520-540); defaults are upstream's 2/4/1."""
    mismatch: int = 2
    gap_open: int = 4
    gap_extend: int = 1


def format_cigar(runs: np.ndarray) -> str:
    """Runs (length << 4 | op) -> phraya CIGAR text."""
    return "".join(f"{int(r) >> 4}{_OPS[int(r) & 3]}" for r in runs)


# Live handles by teardown rank: plans and contig plans, then servers, then
# contexts.  At interpreter exit they are released in that order while the
# HIP runtime is fully up, instead of whenever finalization collects them.
_live = (weakref.WeakSet(), weakref.WeakSet(), weakref.WeakSet())


def _track(obj, rank: int) -> None:
    _live[rank].add(obj)


@atexit.register
def _release_all() -> None:
    for rank, attr in ((0, "close"), (1, "stop"), (2, "close")):
        for obj in list(_live[rank]):
            try:
                getattr(obj, attr)()
            except Exception:
                pass


class Context:
    """A device context (phx_ctx): stream, buffers, submission lock."""

    def __init__(self, device: int = 0, tuning: Optional[dict] = None):
        lib = L.lib()
        tun = None
        if tuning:
            t = L.CTuning(**tuning)
            tun = ctypes.byref(t)
        self._h = lib.phx_ctx_create(device, 0, tun)
        if not self._h:
            raise RuntimeError(f"phx_ctx_create({device}) failed: no usable GPU")
        _track(self, 2)
        self.device = device

    @property
    def handle(self):
        return self._h

    def last_error(self) -> str:
        return L.lib().phx_ctx_last_error(self._h).decode()

    def close(self):
        if getattr(self, "_h", None):
            L.lib().phx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None
_ctx_lock = threading.Lock()


def default_context(tuning: Optional[dict] = None) -> Context:
    """The process-wide context; `tuning` (phx_tuning fields) applies when it
    is first created."""
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0, tuning)
        return _default_ctx


# --------------------------------------------------------------------------- batches

@dataclass
class PackedBatch:
    """Concatenated host arrays in the phx_batch layout."""
    q_seq: np.ndarray
    q_off: np.ndarray
    t_seq: np.ndarray
    t_off: np.ndarray
    seed_q: np.ndarray
    seed_t: np.ndarray

    @property
    def n(self) -> int:
        return len(self.q_off) - 1


def _concat(seqs: Sequence[bytes]) -> tuple[np.ndarray, np.ndarray]:
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(s) for s in seqs], dtype=np.uint64)
    buf = np.frombuffer(b"".join(seqs), dtype=np.uint8) if off[-1] else np.zeros(1, np.uint8)
    return np.ascontiguousarray(buf), off


def pack_batch(queries: Sequence[bytes], targets: Sequence[bytes],
               seeds: Optional[Iterable] = None) -> PackedBatch:
    if len(queries) != len(targets):
        raise ValueError("queries and targets differ in length")
    q_seq, q_off = _concat(queries)
    t_seq, t_off = _concat(targets)
    n = len(queries)
    sq = np.zeros(n, dtype=np.uint64)
    st = np.zeros(n, dtype=np.uint64)
    if seeds is not None:
        for i, s in enumerate(seeds):
            if isinstance(s, SeedAnchor):
                sq[i], st[i] = s.query_pos, s.target_pos
            else:
                sq[i], st[i] = s
    return PackedBatch(q_seq, q_off, t_seq, t_off, sq, st)


def _cbatch(pb: PackedBatch, engine: int, costs: AffineCosts, cap: Optional[int]) -> L.CBatch:
    return L.CBatch(
        pb.n, engine, L.Costs(costs.mismatch, costs.gap_open, costs.gap_extend),
        -1 if cap is None else int(cap),
        pb.q_seq.ctypes.data, pb.q_off.ctypes.data, pb.t_seq.ctypes.data, pb.t_off.ctypes.data,
        pb.seed_q.ctypes.data, pb.seed_t.ctypes.data)


@dataclass
class BatchResult:
    """Per-pair results: `aln` is a structured array (status, score, edit_distance,
    query/target start/end); pair i's CIGAR runs are runs[run_off[i]:run_off[i+1]]."""
    aln: np.ndarray
    runs: np.ndarray
    run_off: np.ndarray

    def cigar(self, i: int) -> str:
        return format_cigar(self.runs[int(self.run_off[i]):int(self.run_off[i + 1])])

    def cigars(self) -> list[str]:
        return [self.cigar(i) for i in range(len(self.aln))]

    def alignment(self, i: int) -> Alignment:
        a = self.aln[i]
        st = int(a["status"])
        if st == L.PHX_INVALID_INPUT:
            raise InvalidInput("Seed position beyond sequence length")
        if st == L.PHX_ALIGNMENT_FAILED:
            raise AlignmentFailed("alignment abandoned: max edit distance exceeded")
        if st != L.PHX_OK:
            raise RuntimeError(f"pair {i}: device status {st}")
        return Alignment(self.cigar(i), int(a["edit_distance"]), int(a["query_start"]),
                         int(a["query_end"]), int(a["target_start"]), int(a["target_end"]),
                         int(a["score"]))


class Plan:
    """A batch resident in HBM (phx_plan): run() launches the kernels only."""

    def __init__(self, pb: PackedBatch, engine: int, costs: AffineCosts = AffineCosts(),
                 cap: Optional[int] = None, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.pb = pb
        self.engine = engine
        self._cb = _cbatch(pb, engine, costs, cap)
        self._h = L.lib().phx_plan_create(self.ctx.handle, ctypes.byref(self._cb))
        if not self._h:
            raise RuntimeError(f"phx_plan_create failed: {self.ctx.last_error()}")
        _track(self, 0)

    def run(self) -> None:
        rc = L.lib().phx_plan_run(self._h)
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_plan_run failed ({rc}): {self.ctx.last_error()}")

    def last_ms(self) -> float:
        return L.lib().phx_plan_last_ms(self._h)

    def last_kernel_ms(self) -> float:
        return L.lib().phx_plan_last_kernel_ms(self._h)

    def last_launches(self) -> int:
        return int(L.lib().phx_plan_last_launches(self._h))

    def last_launch_info(self) -> list[dict]:
        """Per launch of the last run: tier (0 LDS wave, 1 HBM wave, 2 256-thread,
        3 1024-thread, 4 512-thread and 5 one-wave workgroup, 8 Myers), storage of the workgroup tiers
        (0 HBM ring, 1 LDS ring, 2 register wavefront) and of the LDS tier (0 a pair
        per wave, 3-4 banded rows), grid, pairs, device ms, and the
        algorithmic bytes, cells and steps of its pairs."""
        out = []
        for i in range(self.last_launches()):
            info = L.CLaunchInfo()
            if L.lib().phx_plan_last_launch(self._h, i, ctypes.byref(info)) != L.PHX_OK:
                break
            out.append({f: getattr(info, f) for f, _ in L.CLaunchInfo._fields_})
        return out

    def tuning(self) -> dict:
        t = L.CTuning()
        L.lib().phx_plan_get_tuning(self._h, ctypes.byref(t))
        return {name: getattr(t, name) for name, _ in L.CTuning._fields_}

    def set_tuning(self, tuning: dict) -> None:
        t = L.CTuning(**tuning)
        rc = L.lib().phx_plan_set_tuning(self._h, ctypes.byref(t))
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_plan_set_tuning: status {rc}")

    def autotune(self, max_trials: int = 0) -> tuple[dict, float]:
        """Time candidate tunings on this plan, keep the fastest (phx_plan_autotune)."""
        return _autotune(self._h, max_trials)

    def use_graph(self, enable: bool = True) -> None:
        """Replay the run's launch sequence as one HIP graph (phx_plan_use_graph)."""
        rc = L.lib().phx_plan_use_graph(self._h, 1 if enable else 0)
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_plan_use_graph: status {rc}")

    def algorithmic_bytes(self) -> int:
        return int(L.lib().phx_plan_algorithmic_bytes(self._h))

    def results(self) -> BatchResult:
        n = self.pb.n
        aln = np.zeros(n, dtype=L.ALN_DTYPE)
        off = np.zeros(n + 1, dtype=np.uint64)
        # First call sizes the run arena (run_off[n]), the second fills it.
        res = L.CResults(aln.ctypes.data, None, 0, off.ctypes.data)
        rc = L.lib().phx_plan_results(self._h, ctypes.byref(res))
        runs = np.zeros(max(int(off[n]), 1), dtype=np.uint32)
        if rc == L.PHX_E_CAPACITY:
            res = L.CResults(aln.ctypes.data, runs.ctypes.data, len(runs), off.ctypes.data)
            rc = L.lib().phx_plan_results(self._h, ctypes.byref(res))
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_plan_results failed ({rc}): {self.ctx.last_error()}")
        return BatchResult(aln, runs[: int(off[-1])], off)

    def close(self):
        if getattr(self, "_h", None):
            L.lib().phx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _autotune(handle, max_trials: int) -> tuple[dict, float]:
    t = L.CTuning()
    ms = ctypes.c_double(0)
    rc = L.lib().phx_plan_autotune(handle, max_trials, ctypes.byref(t), ctypes.byref(ms))
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_plan_autotune: status {rc}")
    return {name: getattr(t, name) for name, _ in L.CTuning._fields_}, ms.value


def align_batch(engine: int, queries: Sequence[bytes], targets: Sequence[bytes],
                seeds: Optional[Iterable] = None, costs: AffineCosts = AffineCosts(),
                cap: Optional[int] = None, ctx: Optional[Context] = None) -> BatchResult:
    """Align many pairs in one submission (phx_align_batch)."""
    pb = pack_batch(queries, targets, seeds)
    plan = Plan(pb, engine, costs, cap, ctx)
    try:
        plan.run()
        return plan.results()
    finally:
        plan.close()


# ------------------------------------------------------------------ per-pair API

def _one(engine: int, query: bytes, target: bytes, seed: SeedAnchor,
         costs: AffineCosts = AffineCosts(), cap: Optional[int] = None) -> Alignment:
    res = align_batch(engine, [bytes(query)], [bytes(target)], [seed], costs, cap)
    return res.alignment(0)


def wfa_extend(query: bytes, target: bytes, seed: SeedAnchor = SeedAnchor(),
               cap: Optional[int] = None) -> Alignment:
    """wfa_extend (lib.rs:74-76): linear WFA, fitting iff tn > qn + qn/2 + 10."""
    return _one(ENGINE_LINEAR, query, target, seed, cap=cap)


wfa_extend_naive = wfa_extend


def wfa_extend_affine(query: bytes, target: bytes, seed: SeedAnchor = SeedAnchor(),
                      costs: AffineCosts = AffineCosts(), cap: Optional[int] = None) -> Alignment:
    """wfa_extend_affine (lib.rs:139-167)."""
    return _one(ENGINE_AFFINE, query, target, seed, costs, cap)


def myers_extend(query: bytes, target: bytes, seed: SeedAnchor = SeedAnchor()) -> Alignment:
    """myers_extend (lib.rs:106-127)."""
    return _one(ENGINE_MYERS, query, target, seed)


def get_device_dispatch_target() -> str:
    """Implementation the calling thread's last per-pair C-ABI call ran
    ("hip-<arch>", "" before any; phx_last_impl): the device counterpart of
    upstream's get_active_dispatch_target (wfa_simd.rs:25-29, :181-184)."""
    return L.lib().phx_last_impl().decode()


def myers_edit_distance(query: bytes, target: bytes) -> tuple[int, str]:
    """myers_edit_distance (lib.rs:94-96): global distance and CIGAR."""
    res = align_batch(ENGINE_MYERS_GLOBAL, [bytes(query)], [bytes(target)])
    a = res.alignment(0)
    return a.edit_distance, a.cigar


# ---------------------------------------------------------------------- serving

_REQ_DTYPE = np.dtype([(n, np.uint64) for n in ("tag", "query", "query_len", "target",
                                                "target_len", "ref_start", "ref_len",
                                                "seed_q", "seed_t")])
_RESP_DTYPE = np.dtype([("tag", np.uint64), ("status", np.int32), ("reserved", np.int32),
                        ("score", np.int64), ("edit_distance", np.int64),
                        ("query_start", np.int64), ("query_end", np.int64),
                        ("target_start", np.int64), ("target_end", np.int64),
                        ("latency_ms", np.float64)])
assert _REQ_DTYPE.itemsize == ctypes.sizeof(L.CRequest)
assert _RESP_DTYPE.itemsize == ctypes.sizeof(L.CResponse)


@dataclass
class Response:
    """One served result: the pair's status and alignment fields, its CIGAR runs
    and the time from submission to the result being ready."""
    tag: int
    status: int
    score: int
    edit_distance: int
    query_start: int
    query_end: int
    target_start: int
    target_end: int
    runs: np.ndarray
    latency_ms: float

    @property
    def cigar(self) -> str:
        return format_cigar(self.runs)

    def alignment(self) -> Alignment:
        if self.status == L.PHX_INVALID_INPUT:
            raise InvalidInput("Seed position beyond sequence length")
        if self.status == L.PHX_ALIGNMENT_FAILED:
            raise AlignmentFailed("alignment abandoned: max edit distance exceeded")
        if self.status != L.PHX_OK:
            raise RuntimeError(f"request {self.tag}: status {self.status}")
        return Alignment(self.cigar, self.edit_distance, self.query_start, self.query_end,
                         self.target_start, self.target_end, self.score)


class Server:
    """Request server (phx_serve_*): requests are admitted against an HBM budget
    for wavefront state and launched in length tiers as they arrive; results come
    back through poll() in completion order."""

    def __init__(self, engine: int = ENGINE_AFFINE, costs: AffineCosts = AffineCosts(),
                 cap: Optional[int] = None, reference=None, hbm_budget_mib: int = 0,
                 max_group: int = 0, lanes_per_tier: int = 0, ctx: Optional[Context] = None):
        self._ctx = ctx or default_context()
        ref = None if reference is None else np.frombuffer(bytes(reference), np.uint8) \
            if not isinstance(reference, np.ndarray) else np.ascontiguousarray(reference, np.uint8)
        cfg = L.CServeConfig(engine, L.Costs(costs.mismatch, costs.gap_open, costs.gap_extend),
                             -1 if cap is None else int(cap),
                             None if ref is None else ref.ctypes.data,
                             0 if ref is None else len(ref), hbm_budget_mib, max_group,
                             lanes_per_tier)
        self._h = L.lib().phx_serve_start(self._ctx.handle, ctypes.byref(cfg))
        if not self._h:
            raise RuntimeError(f"phx_serve_start failed: {self._ctx.last_error()}")
        _track(self, 1)
        self._runs = np.zeros(1 << 16, np.uint32)

    def submit_arrays(self, tags, q_seq: np.ndarray, q_off: np.ndarray, ref_start=None,
                      ref_len=None, t_seq: Optional[np.ndarray] = None,
                      t_off: Optional[np.ndarray] = None, seed_q=None, seed_t=None,
                      idx=None) -> None:
        """Submit requests whose queries are q_seq[q_off[i]:q_off[i+1]] and whose
        targets are reference windows (ref_start, ref_len) or t_seq slices; `idx`
        selects which i to submit."""
        q_seq = np.ascontiguousarray(q_seq, np.uint8)
        q_off = np.asarray(q_off, np.uint64)
        idx = np.arange(len(q_off) - 1) if idx is None else np.asarray(idx)
        r = np.zeros(len(idx), _REQ_DTYPE)
        r["tag"] = np.asarray(tags, np.uint64) if np.ndim(tags) else np.uint64(tags)
        r["query"] = np.uint64(q_seq.ctypes.data) + q_off[idx]
        r["query_len"] = q_off[idx + 1] - q_off[idx]
        if t_seq is not None:
            t_seq = np.ascontiguousarray(t_seq, np.uint8)
            t_off = np.asarray(t_off, np.uint64)
            r["target"] = np.uint64(t_seq.ctypes.data) + t_off[idx]
            r["target_len"] = t_off[idx + 1] - t_off[idx]
        else:
            r["ref_start"] = np.asarray(ref_start, np.uint64)[idx]
            r["ref_len"] = np.asarray(ref_len, np.uint64)[idx]
        if seed_q is not None:
            r["seed_q"] = np.asarray(seed_q, np.uint64)[idx]
        if seed_t is not None:
            r["seed_t"] = np.asarray(seed_t, np.uint64)[idx]
        rc = L.lib().phx_serve_submit(self._h, r.ctypes.data, len(r))
        if rc != L.PHX_OK:
            raise ValueError(f"phx_serve_submit: status {rc}")

    def submit(self, tag: int, query: bytes, target: Optional[bytes] = None,
               ref_start: int = 0, ref_len: int = 0, seed: SeedAnchor = SeedAnchor()) -> None:
        q = np.frombuffer(bytes(query), np.uint8) if len(query) else np.zeros(1, np.uint8)
        qo = np.array([0, len(query)], np.uint64)
        if target is not None:
            t = np.frombuffer(bytes(target), np.uint8) if len(target) else np.zeros(1, np.uint8)
            self.submit_arrays([tag], q, qo, t_seq=t, t_off=np.array([0, len(target)], np.uint64),
                               seed_q=[seed.query_pos], seed_t=[seed.target_pos])
        else:
            self.submit_arrays([tag], q, qo, ref_start=[ref_start], ref_len=[ref_len],
                               seed_q=[seed.query_pos], seed_t=[seed.target_pos])

    def poll_raw(self, max_out: int = 4096, timeout_us: int = 0):
        """Finished responses as (structured array, runs, run_off)."""
        out = np.zeros(max_out, _RESP_DTYPE)
        off = np.zeros(max_out + 1, np.uint64)
        n = ctypes.c_uint32(0)
        while True:
            rc = L.lib().phx_serve_poll(self._h, out.ctypes.data, max_out, self._runs.ctypes.data,
                                        len(self._runs), off.ctypes.data, ctypes.byref(n),
                                        int(timeout_us))
            if rc == L.PHX_E_CAPACITY:
                self._runs = np.zeros(max(2 * len(self._runs), int(off[0]) + 1), np.uint32)
                continue
            if rc != L.PHX_OK:
                raise RuntimeError(f"phx_serve_poll: status {rc}")
            k = n.value
            return out[:k].copy(), self._runs[:int(off[k])].copy(), off[:k + 1].copy()

    def poll(self, max_out: int = 4096, timeout_us: int = 0) -> list[Response]:
        out, runs, off = self.poll_raw(max_out, timeout_us)
        return [Response(int(o["tag"]), int(o["status"]), int(o["score"]), int(o["edit_distance"]),
                         int(o["query_start"]), int(o["query_end"]), int(o["target_start"]),
                         int(o["target_end"]), runs[int(off[i]):int(off[i + 1])],
                         float(o["latency_ms"])) for i, o in enumerate(out)]

    def stats(self) -> dict:
        st = L.CServeStats()
        L.lib().phx_serve_stats_get(self._h, ctypes.byref(st))
        return {name: getattr(st, name) for name, _ in L.CServeStats._fields_}

    def stop(self) -> None:
        if getattr(self, "_h", None):
            L.lib().phx_serve_stop(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


# ------------------------------------------------------------------ contig mode

@dataclass(frozen=True)
class ContigParams:
    """Anchoring of contig mode: (k, w) minimizers, anchor spacing drawn
    log-uniform in [gap_min, gap_max] bases from a stream seeded by `seed`."""
    k: int = 15
    w: int = 10
    gap_min: int = 1000
    gap_max: int = 20000
    seed: int = 0

    def c(self) -> L.CContigParams:
        return L.CContigParams(self.k, self.w, self.gap_min, self.gap_max, self.seed)


def contig_anchors(query: bytes, target: bytes, params: ContigParams = ContigParams(),
                   pair_index: int = 0) -> list[tuple[int, int]]:
    """Kept anchors (query pos, target pos) of one pair (host only)."""
    q, t = bytes(query), bytes(target)
    p = params.c()
    cap = max(16, len(q) // max(1, params.gap_min) + 16)
    out = np.zeros(2 * cap, np.uint64)
    n = L.lib().phx_contig_anchors(q, len(q), t, len(t), ctypes.byref(p), pair_index,
                                   out.ctypes.data, cap)
    if n < 0:
        raise ValueError("bad contig parameters")
    if n > cap:
        out = np.zeros(2 * n, np.uint64)
        L.lib().phx_contig_anchors(q, len(q), t, len(t), ctypes.byref(p), pair_index,
                                   out.ctypes.data, n)
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]


class ContigPlan:
    """Contig pairs anchored, resident in HBM, with one batched plan over all
    inter-anchor gaps (phx_contig_plan_*); run() aligns and stitches."""

    def __init__(self, q_seq: np.ndarray, q_off: np.ndarray, t_seq: np.ndarray, t_off: np.ndarray,
                 engine: int = ENGINE_AFFINE, costs: AffineCosts = AffineCosts(),
                 params: ContigParams = ContigParams(), ctx: Optional[Context] = None):
        self._ctx = ctx or default_context()
        self.n = len(q_off) - 1
        self._keep = (np.ascontiguousarray(q_seq, np.uint8), np.asarray(q_off, np.uint64),
                      np.ascontiguousarray(t_seq, np.uint8), np.asarray(t_off, np.uint64))
        qs, qo, ts, to = self._keep
        b = L.CBatch(self.n, engine, L.Costs(costs.mismatch, costs.gap_open, costs.gap_extend), -1,
                     qs.ctypes.data, qo.ctypes.data, ts.ctypes.data, to.ctypes.data, None, None)
        p = params.c()
        self._h = L.lib().phx_contig_plan_create(self._ctx.handle, engine, b.costs,
                                                 ctypes.byref(b), ctypes.byref(p))
        if not self._h:
            raise RuntimeError(f"phx_contig_plan_create failed: {self._ctx.last_error()}")
        _track(self, 0)
        self.params = params

    @property
    def gaps(self) -> int:
        return int(L.lib().phx_contig_plan_gaps(self._h))

    def layout(self, i: int) -> np.ndarray:
        """Contig i's gaps as rows (q0, q1, t0, t1)."""
        n = L.lib().phx_contig_plan_layout(self._h, i, None, 0)
        out = np.zeros((max(n, 1), 4), np.uint64)
        L.lib().phx_contig_plan_layout(self._h, i, out.ctypes.data, n)
        return out[:n]

    def gap_plan_algorithmic_bytes(self) -> int:
        return int(L.lib().phx_plan_algorithmic_bytes(L.lib().phx_contig_plan_gap_plan(self._h)))

    def autotune(self, max_trials: int = 0) -> tuple[dict, float]:
        """Tune the batched gap plan (tier edges, slots) on this contig set."""
        return _autotune(L.lib().phx_contig_plan_gap_plan(self._h), max_trials)

    def gap_kernel_ms(self) -> float:
        return float(L.lib().phx_plan_last_kernel_ms(L.lib().phx_contig_plan_gap_plan(self._h)))

    def run(self) -> None:
        rc = L.lib().phx_contig_plan_run(self._h)
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_contig_plan_run: status {rc}: {self._ctx.last_error()}")

    def last_ms(self) -> float:
        return float(L.lib().phx_contig_plan_last_ms(self._h))

    def last_gpu_ms(self) -> float:
        return float(L.lib().phx_contig_plan_last_gpu_ms(self._h))

    def last_stitch_ms(self) -> float:
        return float(L.lib().phx_contig_plan_last_stitch_ms(self._h))

    def results(self) -> BatchResult:
        aln = np.zeros(self.n, L.ALN_DTYPE)
        off = np.zeros(self.n + 1, np.uint64)
        res = L.CResults(aln.ctypes.data, None, 0, off.ctypes.data)
        rc = L.lib().phx_contig_plan_results(self._h, ctypes.byref(res))
        runs = np.zeros(max(int(off[self.n]), 1), np.uint32)
        res = L.CResults(aln.ctypes.data, runs.ctypes.data, len(runs), off.ctypes.data)
        rc = L.lib().phx_contig_plan_results(self._h, ctypes.byref(res))
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_contig_plan_results: status {rc}")
        return BatchResult(aln, runs, off)

    def close(self):
        if getattr(self, "_h", None):
            L.lib().phx_contig_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- read extension

STRATEGY_FAST, STRATEGY_BALANCED, STRATEGY_SENSITIVE = 0, 1, 2
GAP_AFFINE, GAP_LINEAR = 0, 1


@dataclass
class ScoredAlignments:
    """ScoredAlignments (phraya-align/src/lib.rs:53-56)."""
    primary: Alignment
    alternatives: list


def _parse_runs(cigar: str) -> np.ndarray:
    import re
    ops = {"M": 0, "X": 1, "I": 2, "D": 3}
    return np.array([(int(n) << 4) | ops[op] for n, op in re.findall(r"(\d+)([MXID])", cigar)],
                    dtype=np.uint32)


def score_alignments(alignments: Sequence[Alignment], query_len: int) -> ScoredAlignments:
    """score_alignments (lib.rs:170-199) through phx_score_alignments (host only)."""
    if not alignments:
        raise ValueError("At least one alignment required")
    n = len(alignments)
    aln = np.zeros(n, dtype=L.ALN_DTYPE)
    runs = [_parse_runs(a.cigar) for a in alignments]
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in runs])
    flat = np.concatenate(runs + [np.zeros(1, np.uint32)]).astype(np.uint32)
    for i, a in enumerate(alignments):
        aln[i] = (0, 0, a.score, a.edit_distance, a.query_start, a.query_end, a.target_start,
                  a.target_end)
    primary = ctypes.c_int32(-1)
    alt = np.zeros(n, np.uint8)
    rc = L.lib().phx_score_alignments(aln.ctypes.data, flat.ctypes.data, off.ctypes.data, n,
                                      int(query_len), ctypes.byref(primary), alt.ctypes.data)
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_score_alignments: status {rc}")
    p = primary.value
    return ScoredAlignments(alignments[p], [alignments[i] for i in range(n) if alt[i]])


@dataclass
class ReadExtension:
    """One read's batched extend_chains result: `scored` is None when no anchor
    extended (upstream returns None); `anchors` holds (role, Alignment or
    status) per anchor in anchor order (roles: 0 failed, 1 extended, 2
    primary, 3 alternative)."""
    scored: Optional[ScoredAlignments]
    anchors: list


def chain_cap(strategy: int) -> int:
    """chain_cap (executor.rs:330-336)."""
    return int(L.lib().phx_chain_cap(strategy))


class Reference:
    """A target resident in HBM (phx_reference) for the executor entry points:
    pass it where they take the target bytes."""

    def __init__(self, target, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        t = np.frombuffer(bytes(target), np.uint8) if not isinstance(target, np.ndarray) \
            else np.ascontiguousarray(target, np.uint8)
        self.len = len(t)
        self._h = L.lib().phx_reference_create(self.ctx.handle, t.ctypes.data if len(t) else None, len(t))
        if not self._h:
            raise RuntimeError(f"phx_reference_create failed: {self.ctx.last_error()}")
        _track(self, 0)

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().phx_reference_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def extend_chains_batch(strategy: int, gap_model: int, target,
                        reads: Sequence[tuple], ctx: Optional[Context] = None) -> list:
    """extend_chains (executor.rs:528-567) for many reads at once
    (phx_extend_chains_batch, or phx_extend_chains_batch_ref when target is a
    Reference).  reads: (query bytes, chain target_starts in rank order) per
    read."""
    ref = target if isinstance(target, Reference) else None
    ctx = ref.ctx if ref else (ctx or default_context())
    n = len(reads)
    t = np.zeros(0, np.uint8) if ref else np.frombuffer(bytes(target), np.uint8)
    t_len = ref.len if ref else len(t)
    keep, rc_arr = [], (L.CReadChains * max(n, 1))()
    for i, (q, starts) in enumerate(reads):
        qa = np.frombuffer(bytes(q), np.uint8)
        sa = np.asarray(list(starts), dtype=np.uint64)
        keep += [qa, sa]
        rc_arr[i] = L.CReadChains(qa.ctypes.data if len(qa) else None, len(qa),
                                  sa.ctypes.data if len(sa) else None, len(sa))
    na, nr = ctypes.c_uint64(0), ctypes.c_uint64(0)
    rc = L.lib().phx_extend_chains_bounds(strategy, t_len, rc_arr, n, ctypes.byref(na),
                                          ctypes.byref(nr))
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_extend_chains_bounds: status {rc}")
    A, R = na.value, max(nr.value, 1)
    out = (L.CReadExtension * max(n, 1))()
    aln = np.zeros(max(A, 1), dtype=L.ALN_DTYPE)
    role = np.zeros(max(A, 1), np.uint8)
    runs = np.zeros(R, np.uint32)
    off = np.zeros(A + 1, np.uint64)
    if ref:
        rc = L.lib().phx_extend_chains_batch_ref(ctx.handle, strategy, gap_model, ref._h, rc_arr, n,
                                                 out, aln.ctypes.data, role.ctypes.data, A,
                                                 runs.ctypes.data, R, off.ctypes.data)
    else:
        rc = L.lib().phx_extend_chains_batch(ctx.handle, strategy, gap_model,
                                             t.ctypes.data if len(t) else None, len(t), rc_arr, n,
                                             out, aln.ctypes.data, role.ctypes.data, A,
                                             runs.ctypes.data, R, off.ctypes.data)
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_extend_chains_batch failed ({rc}): {ctx.last_error()}")
    return _read_extensions(out, n, aln, role, runs, off)


def _read_extensions(out, n, aln, role, runs, off) -> list:
    res = []
    for i in range(n):
        o = out[i]
        anchors = []
        for a in range(o.first_anchor, o.first_anchor + o.n_anchors):
            r = aln[a]
            if int(r["status"]) != L.PHX_OK:
                anchors.append((0, int(r["status"])))
                continue
            anchors.append((int(role[a]), Alignment(
                format_cigar(runs[int(off[a]):int(off[a + 1])]), int(r["edit_distance"]),
                int(r["query_start"]), int(r["query_end"]), int(r["target_start"]),
                int(r["target_end"]), int(r["score"]))))
        scored = None
        if o.status == L.PHX_OK:
            prim = anchors[o.primary][1]
            alts = [al for rl, al in anchors if rl == 3]
            scored = ScoredAlignments(prim, alts)
        res.append(ReadExtension(scored, anchors))
    return res


# --- pre-extension: seeding and chaining --------------------------------------

def default_chain_params(**kw) -> "L.CChainParams":
    """ChainParams::default() (chaining.rs:69-85) with overrides."""
    d = dict(seed_len=21, max_band=100, max_dist_x=100, max_iter=5000, max_skip=25,
             gap_penalty=0.21, min_chain_score=25)
    d.update(kw)
    return L.CChainParams(**d)


def _sketch_arrays(mins):
    a = np.asarray(mins, dtype=np.uint64).reshape(-1, 2) if len(mins) else np.zeros((0, 2), np.uint64)
    return np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1].astype(np.uint32))


class SeedIndex:
    """build_minimizer_index (seeding.rs:29-35) of a target sketch
    [(value, position), ...], resident on the device."""

    def __init__(self, target_mins, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        v, p = _sketch_arrays(target_mins)
        self._h = L.lib().phx_seed_index_create(self.ctx.handle, v.ctypes.data if len(v) else None,
                                                p.ctypes.data if len(p) else None, len(v))
        if not self._h:
            raise RuntimeError(f"phx_seed_index_create failed: {self.ctx.last_error()}")
        _track(self, 0)

    def occurrence_cap(self, floor: int) -> int:
        """seed_occurrence_cap (seeding.rs:90-97)."""
        return int(L.lib().phx_seed_occurrence_cap(self._h, floor))

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().phx_seed_index_destroy(self._h)
            self._h = None

    def _release(self) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Chains:
    n_seeds: np.ndarray    # per query
    chains: np.ndarray     # CHAIN_DTYPE, query i's at chain_off[i] .. chain_off[i + 1]
    chain_off: np.ndarray
    seeds: np.ndarray      # SEED_DTYPE, each chain's at seed_off .. + n_seeds


class ChainPlan:
    """seed_and_chain (executor.rs:511-519) for a batch of query sketches
    against a SeedIndex, resident on the device (phx_chain_plan_*)."""

    def __init__(self, index: SeedIndex, queries, max_occ: int = 2 ** 64 - 1, params=None):
        self.ctx = index.ctx
        self.index = index
        qs = [_sketch_arrays(q) for q in queries]
        self._qv = np.concatenate([a for a, _ in qs] + [np.zeros(0, np.uint64)])
        self._qp = np.concatenate([b for _, b in qs] + [np.zeros(0, np.uint32)])
        self._off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.uint64)
        self.n = len(queries)
        self._params = params or default_chain_params()
        self._h = L.lib().phx_chain_plan_create(self.ctx.handle, index._h, max_occ, ctypes.byref(self._params),
                                                self._qv.ctypes.data if len(self._qv) else None,
                                                self._qp.ctypes.data if len(self._qp) else None,
                                                self._off.ctypes.data, self.n)
        if not self._h:
            raise RuntimeError(f"phx_chain_plan_create failed: {self.ctx.last_error()}")
        _track(self, 0)

    @property
    def seeds(self) -> int:
        return int(L.lib().phx_chain_plan_seeds(self._h))

    def run(self) -> None:
        rc = L.lib().phx_chain_plan_run(self._h)
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_chain_plan_run failed ({rc}): {self.ctx.last_error()}")

    def last_kernel_ms(self) -> float:
        return float(L.lib().phx_chain_plan_last_kernel_ms(self._h))

    def results(self) -> Chains:
        n_seeds = np.zeros(max(self.n, 1), np.uint64)
        off = np.zeros(self.n + 1, np.uint64)
        out = L.CChainResults(None, 0, None, 0, off.ctypes.data, n_seeds.ctypes.data, 0, 0)
        rc = L.lib().phx_chain_plan_results(self._h, ctypes.byref(out))
        ch, sd = np.zeros(0, L.CHAIN_DTYPE), np.zeros(0, L.SEED_DTYPE)
        if rc == L.PHX_E_CAPACITY:
            ch = np.zeros(max(1, out.n_chains), L.CHAIN_DTYPE)
            sd = np.zeros(max(1, out.n_chain_seeds), L.SEED_DTYPE)
            out.chains, out.chains_cap = ch.ctypes.data, len(ch)
            out.seeds, out.seeds_cap = sd.ctypes.data, len(sd)
            rc = L.lib().phx_chain_plan_results(self._h, ctypes.byref(out))
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_chain_plan_results failed ({rc}): {self.ctx.last_error()}")
        return Chains(n_seeds[:self.n], ch[:out.n_chains], off, sd[:out.n_chain_seeds])

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().phx_chain_plan_destroy(self._h)
            self._h = None

    def _release(self) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def seed_and_chain(target_mins, queries, max_occ: int = 2 ** 64 - 1, params=None,
                   ctx: Optional[Context] = None) -> Chains:
    """One-shot: index the target sketch, seed and chain every query sketch."""
    ix = SeedIndex(target_mins, ctx)
    try:
        p = ChainPlan(ix, queries, max_occ, params)
        try:
            p.run()
            return p.results()
        finally:
            p.close()
    finally:
        ix.close()


# --- minimizer sketches --------------------------------------------------------

DEFAULT_K, DEFAULT_W = 21, 11                   # types.rs:1743-1746
DEFAULT_K_PROTEIN, DEFAULT_W_PROTEIN = 6, 5     # types.rs:1754-1757


def _concat(seqs):
    arrs = [np.frombuffer(bytes(s), np.uint8) if not isinstance(s, np.ndarray) else
            np.ascontiguousarray(s, np.uint8) for s in seqs]
    off = np.zeros(len(arrs) + 1, np.uint64)
    if arrs:
        off[1:] = np.cumsum([len(a) for a in arrs])
    buf = np.concatenate(arrs) if arrs else np.zeros(0, np.uint8)
    return np.ascontiguousarray(buf), off


class SketchPlan:
    """sketch_alphabet (phraya-core/src/types.rs:1846-1867) for a batch of
    sequences, resident on the device (phx_sketch_plan_*): canonical
    minimizers for DNA, forward ones for protein."""

    def __init__(self, seqs, k: int = DEFAULT_K, w: int = DEFAULT_W, canonical: bool = True,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self._buf, self._off = _concat(seqs)
        self.n = len(self._off) - 1
        self._h = L.lib().phx_sketch_plan_create(self.ctx.handle, self._buf.ctypes.data if len(self._buf) else None,
                                                 self._off.ctypes.data, self.n, k, w, 1 if canonical else 0)
        if not self._h:
            raise ValueError(f"phx_sketch_plan_create failed: {self.ctx.last_error()}")
        _track(self, 0)

    def run(self) -> None:
        rc = L.lib().phx_sketch_plan_run(self._h)
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_sketch_plan_run failed ({rc}): {self.ctx.last_error()}")

    def last_kernel_ms(self) -> float:
        return float(L.lib().phx_sketch_plan_last_kernel_ms(self._h))

    def last_launch_ms(self, i: int) -> float:
        """0 compute, 1 scan, 2 copy."""
        return float(L.lib().phx_sketch_plan_last_launch_ms(self._h, i))

    @property
    def capacity(self) -> int:
        return int(L.lib().phx_sketch_plan_capacity(self._h))

    def total(self) -> int:
        return int(L.lib().phx_sketch_plan_total(self._h))

    def results(self):
        """(values u64, positions u32, off u64[n + 1]): sequence i's
        minimizers at off[i] .. off[i + 1], in window order."""
        off = np.zeros(self.n + 1, np.uint64)
        total = ctypes.c_uint64(0)
        rc = L.lib().phx_sketch_plan_results(self._h, None, None, 0, off.ctypes.data, ctypes.byref(total))
        vals = np.zeros(max(1, total.value), np.uint64)
        pos = np.zeros(max(1, total.value), np.uint32)
        if rc in (L.PHX_OK, L.PHX_E_CAPACITY):
            rc = L.lib().phx_sketch_plan_results(self._h, vals.ctypes.data, pos.ctypes.data, len(vals),
                                                 off.ctypes.data, ctypes.byref(total))
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_sketch_plan_results failed ({rc}): {self.ctx.last_error()}")
        return vals[:total.value], pos[:total.value], off

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().phx_sketch_plan_destroy(self._h)
            self._h = None

    def _release(self) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sketch_batch(seqs, k: int = DEFAULT_K, w: int = DEFAULT_W, canonical: bool = True,
                 ctx: Optional[Context] = None):
    """Sketches of many sequences in one device pass: a list of (n, 2) u64
    arrays of (value, position) rows, MinimizerSketch's layout."""
    p = SketchPlan(seqs, k, w, canonical, ctx)
    try:
        p.run()
        v, q, off = p.results()
    finally:
        p.close()
    out = []
    for i in range(len(off) - 1):
        a, b = int(off[i]), int(off[i + 1])
        out.append(np.stack([v[a:b], q[a:b].astype(np.uint64)], axis=1) if b > a else np.zeros((0, 2), np.uint64))
    return out


def sketch_alphabet(seq, k: int, w: int, alphabet: str = "dna", ctx: Optional[Context] = None):
    """sketch_alphabet (types.rs:1846-1867): canonical minimizers for "dna",
    forward ones for "protein"."""
    if alphabet not in ("dna", "protein"):
        raise ValueError("alphabet must be 'dna' or 'protein'")
    return sketch_batch([seq], k, w, alphabet == "dna", ctx)[0]


def sketch(seq, k: int = DEFAULT_K, w: int = DEFAULT_W, ctx: Optional[Context] = None):
    """sketch (types.rs:1873-1877): canonical DNA minimizers."""
    return sketch_alphabet(seq, k, w, "dna", ctx)


# --- align_read's strand handling ---------------------------------------------

STRAND_FORWARD, STRAND_REVERSE = 0, 1
OUTCOME_PLACED, OUTCOME_NO_SEED, OUTCOME_BELOW_THRESHOLD, OUTCOME_FAST_CUTOFF, \
    OUTCOME_NO_ALIGNMENT = 0, 1, 2, 3, 4


def reverse_complement(seq: bytes) -> bytes:
    """reverse_complement (phraya-core/src/types.rs:91-107), host code of the
    library."""
    a = np.frombuffer(bytes(seq), np.uint8)
    out = np.zeros(max(len(a), 1), np.uint8)
    L.lib().phx_reverse_complement(a.ctypes.data if len(a) else None, len(a), out.ctypes.data)
    return out[:len(a)].tobytes()


@dataclass
class ReadPlacement:
    """align_read's strand handling for one read (executor.rs:573-733):
    outcome (OUTCOME_*), the kept strand (STRAND_*, -1 when none), the kept
    orientation's query bytes and ScoredAlignments (None when no alignment;
    set also for FAST_CUTOFF, where upstream returns None), its primary score,
    and each extended orientation's ReadExtension (None when skipped)."""
    outcome: int
    strand: int
    query: bytes
    scored: Optional[ScoredAlignments]
    primary_score: float
    forward: Optional[ReadExtension]
    reverse: Optional[ReadExtension]


def align_reads_batch(strategy: int, gap_model: int, target, reads: Sequence[tuple],
                      is_dna: bool = True, ctx: Optional[Context] = None) -> list:
    """phx_align_reads_batch (phx_align_reads_batch_ref when target is a
    Reference): reads are (query bytes, forward chain target_starts, reverse
    chain target_starts, had_seeds); every orientation worth extending goes
    through one batched extend_chains."""
    ref = target if isinstance(target, Reference) else None
    ctx = ref.ctx if ref else (ctx or default_context())
    n = len(reads)
    t = np.zeros(0, np.uint8) if ref else np.frombuffer(bytes(target), np.uint8)
    t_len = ref.len if ref else len(t)
    keep, arr = [], (L.CReadStrands * max(n, 1))()

    def u64(x):
        a = np.asarray(list(x), dtype=np.uint64)
        keep.append(a)
        return (a.ctypes.data if len(a) else None), len(a)

    for i, (q, fwd, rev, seeds) in enumerate(reads):
        qa = np.frombuffer(bytes(q), np.uint8)
        keep.append(qa)
        fp, fn = u64(fwd)
        rp, rn = u64(rev)
        arr[i] = L.CReadStrands(qa.ctypes.data if len(qa) else None, len(qa), fp, fn, rp, rn,
                                1 if seeds else 0)
    ne, na, nr = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
    rc = L.lib().phx_align_reads_bounds(strategy, int(is_dna), t_len, arr, n, ctypes.byref(ne),
                                        ctypes.byref(na), ctypes.byref(nr))
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_align_reads_bounds: status {rc}")
    E, A, R = ne.value, na.value, max(nr.value, 1)
    place = (L.CReadPlacement * max(n, 1))()
    ext = (L.CReadExtension * max(E, 1))()
    aln = np.zeros(max(A, 1), dtype=L.ALN_DTYPE)
    role = np.zeros(max(A, 1), np.uint8)
    runs = np.zeros(R, np.uint32)
    off = np.zeros(A + 1, np.uint64)
    if ref:
        rc = L.lib().phx_align_reads_batch_ref(ctx.handle, strategy, gap_model, int(is_dna), ref._h, arr,
                                               n, place, ext, E, aln.ctypes.data, role.ctypes.data, A,
                                               runs.ctypes.data, R, off.ctypes.data)
    else:
        rc = L.lib().phx_align_reads_batch(ctx.handle, strategy, gap_model, int(is_dna),
                                           t.ctypes.data if len(t) else None, len(t), arr, n, place,
                                           ext, E, aln.ctypes.data, role.ctypes.data, A,
                                           runs.ctypes.data, R, off.ctypes.data)
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_align_reads_batch failed ({rc}): {ctx.last_error()}")
    exts = _read_extensions(ext, E, aln, role, runs, off)
    res = []
    for i, (q, _, _, _) in enumerate(reads):
        p = place[i]
        fwd = exts[p.entry_fwd] if p.entry_fwd >= 0 else None
        rev = exts[p.entry_rev] if p.entry_rev >= 0 else None
        kept = None if p.strand < 0 else (fwd if p.strand == STRAND_FORWARD else rev)
        qb = bytes(q) if p.strand != STRAND_REVERSE else (reverse_complement(q) if is_dna else b"")
        res.append(ReadPlacement(int(p.outcome), int(p.strand), qb,
                                 kept.scored if kept else None, float(p.primary_score), fwd, rev))
    return res


# --- post-extension: variant sites from CIGAR runs ---------------------------

VAR_SNV, VAR_DELETION, VAR_INSERTION = 0, 1, 2
VAR_IN_REPEAT, VAR_IN_HOTSPOT = 1, 2


@dataclass
class Variants:
    """Variant records (L.VARIANT_DTYPE) in alignment then CIGAR order;
    alignment i's at var_off[i] .. var_off[i + 1]; coverage values of a
    record at coverage[coverage_off .. + window_len]."""
    variants: np.ndarray
    var_off: np.ndarray
    coverage: np.ndarray


class VariantPlan:
    """extract_variants_from_cigar (executor.rs:814-1044) for a batch of
    alignments against one target, resident on the device
    (phx_variant_plan_*).  runs / run_off: CIGAR runs per alignment;
    target_start: the target position of each alignment's first base;
    queries: the query bytes of each alignment; spans: (start, end) coverage
    windows of each alignment's read (compute_windowed_coverage, :1071-1090),
    one list per alignment; repeats / hotspots: (start, end) intervals
    sorted by start (creation fails otherwise), non-overlapping as upstream's
    are; radius: coverage_window_radius."""

    def __init__(self, runs, run_off, target_start, queries: Sequence[bytes], target: bytes,
                 spans: Sequence[Sequence[tuple]], repeats=None, hotspots=None, radius: int = 25,
                 ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        n = len(target_start)
        self._keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(x, dtype=dt)
            self._keep.append(a)
            return a

        runs = arr(runs, np.uint32)
        run_off = arr(run_off, np.uint64)
        ts = arr(target_start, np.uint64)
        q_seq = arr(np.frombuffer(b"".join(bytes(q) for q in queries), np.uint8), np.uint8)
        q_off = arr(np.concatenate([[0], np.cumsum([len(q) for q in queries])]), np.uint64)
        t = arr(np.frombuffer(bytes(target), np.uint8), np.uint8)
        flat = [(int(a), int(b)) for sp in spans for (a, b) in sp]
        sp = arr(np.array(flat, dtype=np.uint64).reshape(-1), np.uint64)
        sp_off = arr(np.concatenate([[0], np.cumsum([len(s) for s in spans])]), np.uint64)
        rep = arr(np.zeros(0) if repeats is None else repeats, np.uint64).reshape(-1)
        hot = arr(np.zeros(0) if hotspots is None else hotspots, np.uint32).reshape(-1)

        def ptr(a):
            return a.ctypes.data if a.size else None

        self.n = n
        b = L.CVariantBatch(n, ptr(runs), ptr(run_off), ptr(ts), ptr(q_seq), ptr(q_off), ptr(t), len(t),
                            ptr(sp), ptr(sp_off), ptr(rep), len(rep) // 2, ptr(hot), len(hot) // 2,
                            radius)
        self._h = L.lib().phx_variant_plan_create(self.ctx.handle, ctypes.byref(b))
        if not self._h:
            raise RuntimeError(f"phx_variant_plan_create failed: {self.ctx.last_error()}")
        _track(self, 0)

    def run(self) -> None:
        rc = L.lib().phx_variant_plan_run(self._h)
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_variant_plan_run failed ({rc}): {self.ctx.last_error()}")

    def last_kernel_ms(self) -> float:
        return float(L.lib().phx_variant_plan_last_kernel_ms(self._h))

    def results(self) -> Variants:
        out = L.CVariantResults(None, 0, None, 0, None, 0, 0)
        var_off = np.zeros(self.n + 1, np.uint64)
        out.var_off = var_off.ctypes.data
        rc = L.lib().phx_variant_plan_results(self._h, ctypes.byref(out))
        if rc == L.PHX_E_CAPACITY:
            v = np.zeros(max(1, out.n_variants), L.VARIANT_DTYPE)
            c = np.zeros(max(1, out.n_coverage), np.uint32)
            out.variants, out.variants_cap = v.ctypes.data, len(v)
            out.coverage, out.coverage_cap = c.ctypes.data, len(c)
            rc = L.lib().phx_variant_plan_results(self._h, ctypes.byref(out))
        else:
            v, c = np.zeros(0, L.VARIANT_DTYPE), np.zeros(0, np.uint32)
        if rc != L.PHX_OK:
            raise RuntimeError(f"phx_variant_plan_results failed ({rc}): {self.ctx.last_error()}")
        return Variants(v[:out.n_variants], var_off, c[:out.n_coverage])

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().phx_variant_plan_destroy(self._h)
            self._h = None

    def _release(self) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def extract_variants(runs, run_off, target_start, queries, target, spans, repeats=None,
                     hotspots=None, radius: int = 25, ctx: Optional[Context] = None) -> Variants:
    """One-shot VariantPlan: upload, run, download."""
    p = VariantPlan(runs, run_off, target_start, queries, target, spans, repeats, hotspots, radius, ctx)
    try:
        p.run()
        return p.results()
    finally:
        p.close()


# --- output files: .phraya and .phraya.queries (host code) ----------------------

@dataclass
class ReadMeta:
    """extract_variants_from_cigar's per-read arguments (executor.rs:814-831)
    for the observations of one alignment."""
    cigar: str
    provenance: str
    query: bytes = b""              # the aligned query bytes (insertion alleles)
    edit_distance: int = 0
    mapq: int = 60
    strand: int = 0                 # STRAND_FORWARD "+", STRAND_REVERSE "-"
    avg_base_quality: float = 0.0
    confidence: float = 1.0
    mate: Optional[dict] = None     # MateInfo fields (types.rs:2031-2050)
    insert_stats: Optional[tuple] = None   # (sum, count)


def _phraya_args(reference_length, sample_id, timestamp, variants, window_coverage, reads, coverage):
    keep = []
    hdr = L.CPhrayaHeader(reference_length, sample_id.encode(), timestamp.encode())
    v = np.ascontiguousarray(variants, dtype=L.VARIANT_DTYPE)
    wc = np.ascontiguousarray(window_coverage, dtype=np.uint32)
    cov = np.ascontiguousarray(coverage, dtype=np.uint32)
    rs = (L.CObsRead * max(1, len(reads)))()
    for i, r in enumerate(reads):
        q = np.frombuffer(bytes(r.query), np.uint8) if len(r.query) else np.zeros(1, np.uint8)
        keep.append(q)
        mate = None
        if r.mate is not None:
            m = r.mate
            mate = L.CMateInfo(m["mate_id"].encode(), m.get("insert_size", 0), int(m.get("proper_pair", False)),
                               int(m.get("is_first_in_pair", False)), int(m.get("is_second_in_pair", False)),
                               int(m.get("mate_mapped", False)))
            keep.append(mate)
        cig, prov = r.cigar.encode(), r.provenance.encode()
        keep += [cig, prov]
        s_, c_ = r.insert_stats if r.insert_stats is not None else (0, 0)
        rs[i] = L.CObsRead(cig, prov, q.ctypes.data, len(r.query), r.avg_base_quality, r.confidence, s_, c_,
                           r.edit_distance, r.mapq, r.strand, int(r.insert_stats is not None), 0,
                           ctypes.addressof(mate) if mate is not None else None)
    keep += [hdr, v, wc, cov, rs]
    args = (ctypes.byref(hdr), v.ctypes.data if len(v) else None, len(v), wc.ctypes.data if len(wc) else None,
            len(wc), rs, len(reads), cov.ctypes.data if len(cov) else None, len(cov))
    return args, keep


def _encoded(fn, *args) -> bytes:
    size = ctypes.c_uint64(0)
    rc = fn(*args, None, 0, ctypes.byref(size))
    if rc not in (L.PHX_OK, L.PHX_E_CAPACITY):
        raise ValueError(f"{fn.__name__} refused the input ({rc})")
    buf = np.zeros(max(1, size.value), np.uint8)
    rc = fn(*args, buf.ctypes.data, len(buf), ctypes.byref(size))
    if rc != L.PHX_OK:
        raise ValueError(f"{fn.__name__} failed ({rc})")
    return buf[:size.value].tobytes()


def encode_phraya(reference_length: int, sample_id: str, timestamp: str, variants, window_coverage, reads,
                  coverage) -> bytes:
    """The MessagePack payload of write_phraya (phraya-io/src/phraya.rs:67-80)
    for the device variant records (VARIANT_DTYPE, window values in
    window_coverage), one ReadMeta per alignment and the summed per-position
    coverage (CoverageTrack::new quantizes and run-length encodes it)."""
    args, keep = _phraya_args(reference_length, sample_id, timestamp, variants, window_coverage, reads, coverage)
    return _encoded(L.lib().phx_phraya_encode, *args)


def write_phraya(path: str, reference_length: int, sample_id: str, timestamp: str, variants, window_coverage,
                 reads, coverage, level: int = 3) -> None:
    """write_phraya: encode_phraya's payload, zstd-compressed, to path."""
    args, keep = _phraya_args(reference_length, sample_id, timestamp, variants, window_coverage, reads, coverage)
    rc = L.lib().phx_phraya_write(str(path).encode(), *args, level)
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_phraya_write failed ({rc})")


def _query_args(index: dict):
    keep = []
    hits = (L.CQueryHits * max(1, len(index)))()
    for i, (qid, lst) in enumerate(index.items()):
        pos = np.ascontiguousarray([p for p, _ in lst], dtype=np.uint32)
        sc = np.ascontiguousarray([s for _, s in lst], dtype=np.float64)
        b = qid.encode()
        keep += [pos, sc, b]
        hits[i] = L.CQueryHits(b, pos.ctypes.data if len(pos) else None, sc.ctypes.data if len(sc) else None,
                               len(lst))
    keep.append(hits)
    return (hits, len(index)), keep


def encode_queries(index: dict) -> bytes:
    """The MessagePack payload of write_queries (queries.rs:62-103) for
    {query_id: [(position, score_ratio), ...]}."""
    args, keep = _query_args(index)
    return _encoded(L.lib().phx_queries_encode, *args)


def write_queries(path: str, index: dict, level: int = 3) -> None:
    args, keep = _query_args(index)
    rc = L.lib().phx_queries_write(str(path).encode(), *args, level)
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_queries_write failed ({rc})")


def _cross_args(index: dict):
    keep = []
    hits = (L.CCrossSpaceHits * max(1, len(index)))()
    for i, (qid, lst) in enumerate(index.items()):
        pl = (L.CPlacement * max(1, len(lst)))()
        for j, (space, pos, ident) in enumerate(lst):
            sb = space.encode()
            keep.append(sb)
            pl[j] = L.CPlacement(sb, pos, ident)
        b = qid.encode()
        keep += [pl, b]
        hits[i] = L.CCrossSpaceHits(b, ctypes.addressof(pl), len(lst))
    keep.append(hits)
    return (hits, len(index)), keep


def encode_cross_space_queries(index: dict) -> bytes:
    """The payload of write_cross_space_queries (queries.rs:132-167) for
    {query_id: [(space, position, identity), ...]}."""
    args, keep = _cross_args(index)
    return _encoded(L.lib().phx_cross_space_queries_encode, *args)


def write_cross_space_queries(path: str, index: dict, level: int = 3) -> None:
    args, keep = _cross_args(index)
    rc = L.lib().phx_cross_space_queries_write(str(path).encode(), *args, level)
    if rc != L.PHX_OK:
        raise RuntimeError(f"phx_cross_space_queries_write failed ({rc})")
