"""Seeded synthetic workloads (host side) for benchmarks and tests.

Workloads (read = query, window = target; seed (0, 0)):
  short150   10,000 reads of 150 bp from a 5 Mbp uniform-ACGT genome, 1%
             substitutions + 0.1% indel events (length 1-3, geometric p=1/2),
             against the 170 bp window starting at the read's source position.
  window300  10,000 reads drawn as short150's, against 300 bp windows (2 x qlen, the window the
             executor cuts for an anchor): the fitting branch.
  myers150   1,000,000 reads of 150 bp, per-read event rate ~ U[0, 5%] (80%
             substitutions, 20% indels of length 1-3), against 180 bp windows.
  mixed      200,000 reads, lengths log-uniform in [100, 20000]; 70% class
             "short" (1% events, 10% of them indels) and 30% class "long" (8%
             events, 60% indels, geometric length mean 2, max 20); window =
             ceil(1.1 * L).
Event rates are per source site; see csrc/synth.cpp for the walk.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_LIB = os.path.join(HERE, "libphx_synth.so")


class SynthParams(ctypes.Structure):
    _fields_ = [
        ("genome_len", ctypes.c_uint64), ("genome_seed", ctypes.c_uint64),
        ("read_seed", ctypes.c_uint64), ("n", ctypes.c_uint32),
        ("len_lo", ctypes.c_uint32), ("len_hi", ctypes.c_uint32),
        ("log_uniform", ctypes.c_int32), ("window_scale", ctypes.c_double),
        ("window_add", ctypes.c_int32), ("class_a_frac", ctypes.c_double),
        ("a_err_lo", ctypes.c_double), ("a_err_hi", ctypes.c_double),
        ("a_indel_frac", ctypes.c_double), ("b_err_lo", ctypes.c_double),
        ("b_err_hi", ctypes.c_double), ("b_indel_frac", ctypes.c_double),
        ("indel_mean", ctypes.c_double), ("indel_max", ctypes.c_int32),
        ("indel_uniform", ctypes.c_int32),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_LIB):
            from .build import build_synth
            build_synth()
        L = ctypes.CDLL(SYNTH_LIB)
        P = ctypes.c_void_p
        L.synth_genome.argtypes = [ctypes.c_uint64, ctypes.c_uint64, P]
        L.synth_pairs.argtypes = [P, P]
        L.synth_pairs.restype = P
        L.synth_q_bytes.argtypes = [P]
        L.synth_q_bytes.restype = ctypes.c_uint64
        L.synth_t_bytes.argtypes = [P]
        L.synth_t_bytes.restype = ctypes.c_uint64
        L.synth_copy.argtypes = [P, P, P, P, P, P]
        L.synth_positions.argtypes = [P, P]
        L.synth_free.argtypes = [P]
        _lib = L
    return _lib


@dataclass
class Workload:
    name: str
    q_seq: np.ndarray
    q_off: np.ndarray
    t_seq: np.ndarray
    t_off: np.ndarray
    cls: np.ndarray
    params: dict
    pos: np.ndarray | None = None  # window start of each target in the genome

    @property
    def n(self) -> int:
        return len(self.q_off) - 1

    def pair(self, i: int) -> tuple[bytes, bytes]:
        return (self.q_seq[self.q_off[i]:self.q_off[i + 1]].tobytes(),
                self.t_seq[self.t_off[i]:self.t_off[i + 1]].tobytes())

    def subset(self, idx) -> "Workload":
        idx = np.asarray(idx)
        qs = [self.q_seq[self.q_off[i]:self.q_off[i + 1]] for i in idx]
        ts = [self.t_seq[self.t_off[i]:self.t_off[i + 1]] for i in idx]
        qo = np.zeros(len(idx) + 1, np.uint64)
        to = np.zeros(len(idx) + 1, np.uint64)
        qo[1:] = np.cumsum([len(x) for x in qs])
        to[1:] = np.cumsum([len(x) for x in ts])
        return Workload(self.name, np.concatenate(qs) if qs else np.zeros(0, np.uint8), qo,
                        np.concatenate(ts) if ts else np.zeros(0, np.uint8), to, self.cls[idx],
                        dict(self.params, subset=len(idx)),
                        None if self.pos is None else self.pos[idx])


def genome(length: int = 5_000_000, seed: int = 0x5EED0001) -> np.ndarray:
    g = np.empty(length, dtype=np.uint8)
    _load().synth_genome(length, seed, g.ctypes.data)
    return g


def minimizer_sketch(seq, k: int = 21, w: int = 11) -> np.ndarray:
    """Forward (k, w) minimizers of an ACGT byte sequence, for synthetic
    seeding / chaining workloads: per window of w consecutive k-mers the
    leftmost one of smallest 64-bit hash (the SplitMix64 finalizer of its
    2-bit packed value), one entry per selected position.  Rows are
    (hash, position), the (value, position) layout of a MinimizerSketch.
    phraya's own sketches come from simd-minimizers (canonical, k = 21,
    w = 11); seeding and chaining depend only on equal values at shared
    k-mers, which this keeps."""
    s = np.frombuffer(bytes(seq), np.uint8) if not isinstance(seq, np.ndarray) else seq.astype(np.uint8)
    n = len(s) - k + 1
    if n < w:
        return np.zeros((0, 2), np.uint64)
    code = ((s >> 1) & 3).astype(np.uint64)  # A 0, C 1, T 2, G 3
    km = np.zeros(n, np.uint64)
    for j in range(k):
        km = (km << np.uint64(2)) | code[j:j + n]
    z = km
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    h = z ^ (z >> np.uint64(31))
    win = np.lib.stride_tricks.sliding_window_view(h, w)
    am = win.argmin(axis=1) + np.arange(n - w + 1)
    keep = np.ones(len(am), bool)
    keep[1:] = am[1:] != am[:-1]
    pos = am[keep]
    return np.stack([h[pos], pos.astype(np.uint64)], axis=1)


def generate(p: dict, name: str = "custom") -> Workload:
    L = _load()
    g = genome(p.get("genome_len", 5_000_000), p.get("genome_seed", 0x5EED0001))
    sp = SynthParams(**{k: v for k, v in p.items() if k in dict(SynthParams._fields_)})
    sp.genome_len = len(g)
    h = L.synth_pairs(ctypes.byref(sp), g.ctypes.data)
    try:
        n = sp.n
        q = np.empty(max(L.synth_q_bytes(h), 1), np.uint8)
        t = np.empty(max(L.synth_t_bytes(h), 1), np.uint8)
        qo = np.empty(n + 1, np.uint64)
        to = np.empty(n + 1, np.uint64)
        cls = np.empty(max(n, 1), np.uint8)
        pos = np.empty(max(n, 1), np.uint64)
        L.synth_copy(h, q.ctypes.data, qo.ctypes.data, t.ctypes.data, to.ctypes.data, cls.ctypes.data)
        L.synth_positions(h, pos.ctypes.data)
    finally:
        L.synth_free(h)
    return Workload(name, q, qo, t, to, cls[:n], dict(p), pos[:n])


WORKLOADS = {
    "short150": dict(genome_seed=0x5EED0001, read_seed=0xC0FFEE01, n=10_000, len_lo=150,
                     len_hi=150, log_uniform=0, window_scale=1.0, window_add=20,
                     class_a_frac=1.0, a_err_lo=0.011, a_err_hi=0.011,
                     a_indel_frac=0.001 / 0.011, b_err_lo=0, b_err_hi=0, b_indel_frac=0,
                     indel_mean=2.0, indel_max=3),
    # Reads drawn as short150's, against the 2 x qlen = 300 bp windows extend_anchor
    # cuts (executor.rs:551-553): the fitting branch.
    "window300": dict(genome_seed=0x5EED0001, read_seed=0xC0FFEE01, n=10_000, len_lo=150,
                      len_hi=150, log_uniform=0, window_scale=1.0, window_add=150,
                      class_a_frac=1.0, a_err_lo=0.011, a_err_hi=0.011,
                      a_indel_frac=0.001 / 0.011, b_err_lo=0, b_err_hi=0, b_indel_frac=0,
                      indel_mean=2.0, indel_max=3),
    "myers150": dict(genome_seed=0x5EED0001, read_seed=0xC0FFEE02, n=1_000_000, len_lo=150,
                     len_hi=150, log_uniform=0, window_scale=1.0, window_add=30,
                     class_a_frac=1.0, a_err_lo=0.0, a_err_hi=0.05, a_indel_frac=0.2,
                     b_err_lo=0, b_err_hi=0, b_indel_frac=0, indel_mean=2.0, indel_max=3),
    # Contig segments 100 kbp - 1 Mbp (log-uniform), divergence U[0.5%, 2%], 10% of
    # events indels of length U[1, 50]; query = mutated segment, target = segment.
    "contig": dict(genome_seed=0x5EED0001, read_seed=0xC0FFEE04, n=2_000, len_lo=100_000,
                   len_hi=1_000_000, log_uniform=1, window_scale=1.0, window_add=0,
                   class_a_frac=1.0, a_err_lo=0.005, a_err_hi=0.02, a_indel_frac=0.1,
                   b_err_lo=0, b_err_hi=0, b_indel_frac=0, indel_mean=25.0, indel_max=50,
                   indel_uniform=1),
    "mixed": dict(genome_seed=0x5EED0001, read_seed=0xC0FFEE03, n=200_000, len_lo=100,
                  len_hi=20_000, log_uniform=1, window_scale=1.1, window_add=0,
                  class_a_frac=0.7, a_err_lo=0.01, a_err_hi=0.01, a_indel_frac=0.1,
                  b_err_lo=0.08, b_err_hi=0.08, b_indel_frac=0.6, indel_mean=2.0, indel_max=20),
}


def workload(name: str, n: int | None = None, seed: int | None = None) -> Workload:
    p = dict(WORKLOADS[name])
    if n is not None:
        p["n"] = n
    if seed is not None:
        p["read_seed"] = seed
    return generate(p, name)


# Request classes of the served stream: 150 bp reads (1% events, 10% indels)
# against 170 bp windows; long reads log-uniform 1-50 kbp (8% events, 60%
# indels, mean length 2) against 1.1x windows; contig gaps log-uniform 5-20 kbp
# (2% events, 10% indels of mean length 10, up to 50) against equal windows.
SERVE_CLASSES = {
    "short": dict(len_lo=150, len_hi=150, log_uniform=0, window_scale=1.0, window_add=20,
                  class_a_frac=1.0, a_err_lo=0.011, a_err_hi=0.011, a_indel_frac=0.001 / 0.011,
                  b_err_lo=0, b_err_hi=0, b_indel_frac=0, indel_mean=2.0, indel_max=3),
    "long": dict(len_lo=1000, len_hi=50_000, log_uniform=1, window_scale=1.1, window_add=0,
                 class_a_frac=1.0, a_err_lo=0.08, a_err_hi=0.08, a_indel_frac=0.6,
                 b_err_lo=0, b_err_hi=0, b_indel_frac=0, indel_mean=2.0, indel_max=20),
    "gap": dict(len_lo=5000, len_hi=20_000, log_uniform=1, window_scale=1.0, window_add=0,
                class_a_frac=1.0, a_err_lo=0.02, a_err_hi=0.02, a_indel_frac=0.1,
                b_err_lo=0, b_err_hi=0, b_indel_frac=0, indel_mean=10.0, indel_max=50),
}
SERVE_CLASS_NAMES = ("short", "long", "gap")


@dataclass
class RequestStream:
    """Requests against one resident genome: query i is q_seq[q_off[i]:q_off[i+1]]
    and its target is genome[pos[i] : pos[i] + win[i]]."""
    genome: np.ndarray
    q_seq: np.ndarray
    q_off: np.ndarray
    pos: np.ndarray
    win: np.ndarray
    cls: np.ndarray

    @property
    def n(self) -> int:
        return len(self.q_off) - 1

    def query(self, i: int) -> bytes:
        return self.q_seq[self.q_off[i]:self.q_off[i + 1]].tobytes()

    def target(self, i: int) -> bytes:
        return self.genome[self.pos[i]:self.pos[i] + self.win[i]].tobytes()


def serve_stream(n: int, mix=(0.6, 0.3, 0.1), seed: int = 0xC0FFEE05,
                 genome_seed: int = 0x5EED0001) -> RequestStream:
    """The served request mix, classes interleaved by a seeded draw."""
    rng = np.random.default_rng(seed)
    cls = rng.choice(3, size=n, p=np.asarray(mix, dtype=np.float64) / sum(mix)).astype(np.uint8)
    g = genome(5_000_000, genome_seed)
    parts = []
    for c, name in enumerate(SERVE_CLASS_NAMES):
        k = int((cls == c).sum())
        if k == 0:
            parts.append(None)
            continue
        p = dict(SERVE_CLASSES[name], n=k, read_seed=seed + 17 * (c + 1), genome_seed=genome_seed)
        parts.append(generate(p, name))
    idx = [0, 0, 0]
    qs, pos, win = [], np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    for i in range(n):
        c = int(cls[i])
        w = parts[c]
        j = idx[c]
        idx[c] += 1
        qs.append(w.q_seq[w.q_off[j]:w.q_off[j + 1]])
        pos[i] = w.pos[j]
        win[i] = w.t_off[j + 1] - w.t_off[j]
    q_off = np.zeros(n + 1, np.uint64)
    q_off[1:] = np.cumsum([len(x) for x in qs])
    q_seq = np.concatenate(qs) if qs else np.zeros(0, np.uint8)
    return RequestStream(g, q_seq, q_off, pos, win, cls)
