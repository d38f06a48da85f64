"""ctypes binding of tests/native/libalncheck.so, the full-size alignment
checker (test infrastructure only): CIGAR replay, cost == score, edit
distance == non-M bases, whole target consumed, and score == the optimum of
a banded Gotoh DP whose band holds every alignment of cost <= score."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")
_LIB_PATH = os.path.join(_HERE, "libalncheck.so")
F_REPLAY, F_CONSUME, F_COST, F_ED, F_OPT, F_TEND = 1, 2, 4, 8, 16, 32
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ctypes.CDLL(_LIB_PATH)
        P, i64, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64
        L.chk_band_global.argtypes = [ctypes.c_int, i64, i64, i64, ctypes.c_char_p, u64,
                                      ctypes.c_char_p, u64, i64, P]
        L.chk_band_global.restype = i64
        L.chk_alignments.argtypes = [ctypes.c_int, i64, i64, i64, ctypes.c_uint32, P, P, P, P, P, P,
                                     P, P, P, ctypes.c_int, ctypes.c_int, P]
        L.chk_alignments.restype = u64
        _lib = L
    return _lib


def band_global(q: bytes, t: bytes, S: int, costs=None) -> tuple[int, int]:
    """(banded optimum or -1, cells) for the global alignment of q and t;
    costs None = linear (unit) costs."""
    cells = ctypes.c_uint64(0)
    mode, (x, o, e) = (0, (1, 0, 1)) if costs is None else (1, costs)
    r = lib().chk_band_global(mode, x, o, e, q, len(q), t, len(t), S, ctypes.byref(cells))
    return int(r), int(cells.value)


def check(q_seq, q_off, t_seq, t_off, score, ed, t_end, runs, run_off, costs=None,
          optimal=True, threads=None) -> tuple[np.ndarray, int]:
    """Per-pair failure bits (0 = all checks passed) and the DP cells used."""
    n = len(q_off) - 1
    arrs = [np.ascontiguousarray(q_seq, np.uint8), np.ascontiguousarray(q_off, np.uint64),
            np.ascontiguousarray(t_seq, np.uint8), np.ascontiguousarray(t_off, np.uint64),
            np.ascontiguousarray(score, np.int64), np.ascontiguousarray(ed, np.int64),
            np.ascontiguousarray(t_end, np.int64), np.ascontiguousarray(runs, np.uint32),
            np.ascontiguousarray(run_off, np.uint64)]
    if arrs[7].size == 0:
        arrs[7] = np.zeros(1, np.uint32)
    flags = np.zeros(max(n, 1), np.uint8)
    mode, (x, o, e) = (0, (1, 0, 1)) if costs is None else (1, costs)
    threads = threads or min(16, os.cpu_count() or 1)
    cells = lib().chk_alignments(mode, x, o, e, n, *[a.ctypes.data for a in arrs], int(optimal),
                                 threads, flags.ctypes.data)
    return flags[:n], int(cells)
