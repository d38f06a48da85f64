"""Memory model of a WFA request (phx_estimate_request; host only).

The estimate sizes HBM history before a search runs.  An estimate below the
true step count costs a re-run, one far above it wastes admission budget, so
it is checked against the true score of each pair (from the oracle)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def est():
    from phraya_amd import _lib as L
    lib = L.lib()

    def f(q: bytes, t: bytes, costs=(4, 6, 2), engine=1):
        s = ctypes.c_int32(0)
        b = lib.phx_estimate_request(engine, L.Costs(*costs), q, len(q), t, len(t),
                                     ctypes.byref(s))
        return s.value, b
    return f


def _pairs(kind: str, n: int):
    from phraya_amd import synth
    if kind == "mixed":
        w = synth.workload("mixed", n=400)
        L = np.diff(w.q_off)
        idx = [i for i in range(w.n) if 600 < L[i] <= 4000][:n]
        return [w.pair(i) for i in idx]
    rs = synth.serve_stream(600)
    c = synth.SERVE_CLASS_NAMES.index(kind)
    cap = 8000 if kind == "gap" else 5000  # gaps are 5-20 kbp
    idx = [i for i in range(rs.n) if rs.cls[i] == c and len(rs.query(i)) <= cap][:n]
    assert len(idx) >= min(n, 8), f"only {len(idx)} {kind} pairs"
    return [(rs.query(i), rs.target(i)) for i in idx]


@pytest.mark.parametrize("kind", ["mixed", "long", "gap"])
def test_affine_estimate_brackets_true_steps(est, kind):
    ratios = []
    for q, t in _pairs(kind, 25):
        steps, nbytes = est(q, t)
        true = O.wfa_extend_affine(q, t, (0, 0), (4, 6, 2)).score // 2  # g = gcd(4, 6, 2)
        ratios.append(steps / max(true, 1))
        # Bytes cover one history cell per diagonal of every estimated step.
        assert nbytes >= 4 * steps
    ratios = np.array(ratios)
    assert (ratios >= 1.0).mean() >= 0.95, ratios
    assert np.median(ratios) <= 2.0, ratios


def test_linear_estimate_brackets_true_steps(est):
    ratios = []
    for q, t in _pairs("gap", 12):
        steps, _ = est(q, t, engine=0)
        true = O.wfa_extend(q, t).score
        ratios.append(steps / max(true, 1))
    ratios = np.array(ratios)
    assert (ratios >= 1.0).all() and np.median(ratios) <= 2.5, ratios


def test_estimate_rejects_bad_arguments(est):
    from phraya_amd import _lib as L
    lib = L.lib()
    s = ctypes.c_int32(0)
    assert lib.phx_estimate_request(3, L.Costs(4, 6, 2), b"ACGT", 4, b"ACGT", 4,
                                    ctypes.byref(s)) == -1
    assert lib.phx_estimate_request(1, L.Costs(0, 6, 2), b"ACGT", 4, b"ACGT", 4,
                                    ctypes.byref(s)) == -1
    # Too short for the k-mer probe: the length-only model answers.
    steps, nbytes = est(b"ACGTACGT", b"ACGTACGTAA")
    assert steps > 0 and nbytes > 0
