"""Pins the full-size checker (tests/native/alncheck.cpp) before the GPU
full-size tests trust it: its banded DP equals the full Gotoh / Levenshtein
DP (tests/dp.py) at, above and below the optimum, its two implementations
(by anti-diagonal, by row) agree on kilobase pairs, the oracle's own
alignments pass every check, and each planted mistake raises its flag."""
from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

import alncheck as C
import dp
from oracle import oracle as O
from test_oracle import mutate, rand_seq

COSTS = [None, (4, 6, 2), (2, 4, 1), (1, 0, 1), (3, 5, 2)]


def _rows(q, t, S, costs):
    L = C.lib()
    L.chk_band_global_rows.argtypes = L.chk_band_global.argtypes
    L.chk_band_global_rows.restype = ctypes.c_int64
    mode, (x, o, e) = (0, (1, 0, 1)) if costs is None else (1, costs)
    return int(L.chk_band_global_rows(mode, x, o, e, q, len(q), t, len(t), S, None))


def test_band_equals_full_dp():
    rng = random.Random(3)
    for _ in range(1500):
        q = rand_seq(rng, rng.randint(0, 30))
        t = mutate(rng, q, rng.choice([0.0, 0.05, 0.2, 0.5])) + rand_seq(rng, rng.randint(0, 8))
        if rng.random() < 0.2:
            t = t[:rng.randint(0, len(t))]
        costs = rng.choice(COSTS)
        x, o, e = (1, 0, 1) if costs is None else costs
        opt = int(dp.gotoh(q, t, x, o, e)[len(q), len(t)]) if (q or t) else 0
        for S in (opt, opt + rng.randint(1, 20)):
            assert C.band_global(q, t, S, costs)[0] == opt
            assert _rows(q, t, S, costs) == opt
        if opt > 0:  # below the optimum nothing in the band reaches S
            r = C.band_global(q, t, opt - 1, costs)[0]
            assert r == -1 or r > opt - 1


@pytest.mark.parametrize("costs", [None, (4, 6, 2)])
def test_band_implementations_agree_on_kilobase_pairs(costs):
    rng = random.Random(4)
    x, o, e = (1, 0, 1) if costs is None else costs
    for _ in range(6):
        L = rng.randint(800, 3000)
        t = rand_seq(rng, int(L * 1.1))
        q = mutate(rng, t[:L], rng.choice([0.01, 0.08]))
        loose = 2 * o + e * (len(q) + len(t))  # every alignment costs at most this
        opt = C.band_global(q, t, loose, costs)[0]
        assert opt == _rows(q, t, loose, costs) > 0
        assert C.band_global(q, t, opt, costs)[0] == opt == _rows(q, t, opt, costs)
        below = C.band_global(q, t, opt - 1, costs)[0]
        assert below == -1 or below > opt - 1


def _oracle_batch(rng, n, costs):
    qs, ts = [], []
    for _ in range(n):
        q = rand_seq(rng, rng.randint(1, 120))
        t = mutate(rng, q, rng.choice([0.0, 0.02, 0.1, 0.3])) + rand_seq(rng, rng.randint(0, 12))
        qs.append(q)
        ts.append(t)
    outs = [O.wfa_affine_global(q, t, costs) if costs else O.wfa_global(q, t) for q, t in zip(qs, ts)]
    runs, run_off = [], [0]
    for o in outs:
        for n_, op in dp.runs(o.cigar):
            runs.append((n_ << 4) | "MXID".index(op))
        run_off.append(len(runs))
    cat = lambda xs: np.frombuffer(b"".join(xs), np.uint8)
    off = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.uint64)
    arrays = dict(q_seq=cat(qs), q_off=off(qs), t_seq=cat(ts), t_off=off(ts),
                  score=np.array([o.score for o in outs], np.int64),
                  ed=np.array([o.edit_distance for o in outs], np.int64),
                  t_end=np.array([o.target_end for o in outs], np.int64),
                  runs=np.array(runs, np.uint32), run_off=np.array(run_off, np.uint64))
    return qs, ts, outs, arrays


@pytest.mark.parametrize("costs", [None, (4, 6, 2), (2, 4, 1)])
def test_oracle_alignments_pass(costs):
    rng = random.Random(5)
    _, _, _, a = _oracle_batch(rng, 300, costs)
    flags, cells = C.check(**a, costs=costs)
    assert not flags.any() and cells > 0


def test_planted_mistakes_are_caught():
    rng = random.Random(6)
    costs = (4, 6, 2)
    qs, ts, outs, a = _oracle_batch(rng, 200, costs)
    i = next(k for k, o in enumerate(outs) if "X" in o.cigar and "M" in o.cigar)
    lo, hi = int(a["run_off"][i]), int(a["run_off"][i + 1])

    def flags_with(**change):
        b = {k: v.copy() for k, v in a.items()}
        for k, fn in change.items():
            fn(b[k])
        return int(C.check(**b, costs=costs)[0][i])

    def swap_m_x(r):  # an M run relabelled X: replay fails (and cost, ed)
        j = next(j for j in range(lo, hi) if (r[j] & 15) == 0)
        r[j] = (r[j] & ~np.uint32(15)) | 1
    assert flags_with(runs=swap_m_x) & C.F_REPLAY
    # A wrong score is neither the CIGAR's cost nor the banded optimum.
    assert flags_with(score=lambda s: s.__setitem__(i, s[i] + 1)) == C.F_COST | C.F_OPT
    assert flags_with(ed=lambda s: s.__setitem__(i, s[i] - 1)) == C.F_ED
    assert flags_with(t_end=lambda s: s.__setitem__(i, s[i] - 1)) & (C.F_CONSUME | C.F_TEND)
    # A valid but suboptimal alignment whose score is its own cost: one
    # mismatch (cost 4) replaced by a query-only and a target-only base
    # (cost 8 + 8): replay, cost and distance bookkeeping all hold, only the
    # banded optimum shows that it is not optimal.
    runs = a["runs"]
    j = next(j for j in range(lo, hi) if (runs[j] & 15) == 1)
    new = list(runs[lo:j]) + ([int(((int(runs[j]) >> 4) - 1) << 4) | 1] if runs[j] >> 4 > 1 else []) + \
        [(1 << 4) | 3, (1 << 4) | 2] + list(runs[j + 1:hi])
    b = {k: v.copy() for k, v in a.items()}
    b["runs"] = np.concatenate([runs[:lo], np.array(new, np.uint32), runs[hi:]])
    b["run_off"] = a["run_off"].copy()
    b["run_off"][i + 1:] += len(new) - (hi - lo)
    b["score"][i] += 8 + 8 - 4
    b["ed"][i] += 1
    assert int(C.check(**b, costs=costs)[0][i]) == C.F_OPT
