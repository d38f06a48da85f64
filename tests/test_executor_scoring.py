"""score_alignments (phraya-align/src/lib.rs:170-199) through
phx_score_alignments (host code of libphraya_hip.so, no GPU needed), and the
anchor bookkeeping of the batched extend_chains (chain_cap, bounds)."""
from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

import phraya_amd as px
from phraya_amd import _lib as L

A = px.Alignment


def test_single_alignment():
    # lib.rs:205-219
    a = A("100M", 0, 0, 100, 0, 100)
    s = px.score_alignments([a], 100)
    assert s.primary == a and s.alternatives == []


def test_similar_alignments_both_stored():
    # lib.rs:221-244: (1 - 3/100) / (1 - 2/100) = 0.9898 >= 0.95
    p = A("100M", 2, 0, 100, 0, 100)
    alt = A("98M1I1D", 3, 0, 100, 0, 100)
    s = px.score_alignments([p, alt], 100)
    assert s.primary == p and s.alternatives == [alt]


def test_distant_alignments_only_primary():
    # lib.rs:246-267: 0.816 < 0.95
    p = A("100M", 2, 0, 100, 0, 100)
    s = px.score_alignments([p, A("80M10I10D", 20, 0, 100, 0, 100)], 100)
    assert s.primary == p and s.alternatives == []


def test_query_len_normalization():
    # lib.rs:269-315: equal distances give ratio 1.0 at both lengths
    s = px.score_alignments([A("10M", 1, 0, 10, 0, 10), A("9M1D", 1, 0, 10, 0, 10)], 10)
    assert len(s.alternatives) == 1
    s = px.score_alignments([A("100M", 1, 0, 100, 0, 100), A("99M1D", 1, 0, 100, 0, 100)], 100)
    assert len(s.alternatives) == 1


def test_empty_is_an_error():
    # lib.rs:317-320 panics; the C ABI reports PHX_E_ARG
    with pytest.raises(ValueError):
        px.score_alignments([], 100)
    p = ctypes.c_int32(0)
    rc = L.lib().phx_score_alignments(None, None, np.zeros(1, np.uint64).ctypes.data, 0, 100,
                                      ctypes.byref(p), np.zeros(1, np.uint8).ctypes.data)
    assert rc == L.PHX_E_ARG


def test_first_minimum_and_copies_of_the_primary():
    # min_by_key returns the first minimum; `*a != &primary` drops every copy.
    p = A("50M", 1, 0, 50, 10, 60)
    same = A("50M", 1, 0, 50, 10, 60)
    other_pos = A("50M", 1, 0, 50, 20, 70)
    s = px.score_alignments([A("48M2X", 2, 0, 50, 0, 50), p, same, other_pos], 50)
    assert s.primary is p or s.primary == p
    assert s.alternatives == [A("48M2X", 2, 0, 50, 0, 50), other_pos]


def brute(alns, qlen):
    """Direct restatement of the rules with IEEE doubles (numpy), for random
    cross-checks; the vectors above pin it to upstream's tests."""
    p = min(range(len(alns)), key=lambda i: (alns[i].edit_distance, i))
    with np.errstate(divide="ignore", invalid="ignore"):
        pn = np.float64(1.0) - np.float64(alns[p].edit_distance) / np.float64(qlen)
        out = []
        for a in alns:
            if a == alns[p]:
                continue
            an = np.float64(1.0) - np.float64(a.edit_distance) / np.float64(qlen)
            if an / pn >= 0.95:
                out.append(a)
    return alns[p], out


@pytest.mark.parametrize("seed", range(5))
def test_random_sets_match_the_rules(seed):
    rng = random.Random(seed)
    for _ in range(300):
        qlen = rng.choice([0, 1, 10, 100, 150, 1000])
        n = rng.randint(1, 8)
        pool = [A(rng.choice(["10M", "5M1X4M", "3M2I5M", "9M1D"]), rng.randint(0, 40),
                  0, qlen, rng.randint(0, 3), rng.randint(3, 9)) for _ in range(3)]
        alns = [rng.choice(pool) if rng.random() < 0.3 else
                A(rng.choice(["10M", "8M2X"]), rng.randint(0, 40), 0, qlen, 0, 10) for _ in range(n)]
        s = px.score_alignments(alns, qlen)
        bp, balt = brute(alns, qlen)
        assert s.primary == bp and s.alternatives == balt


def test_chain_cap_and_bounds():
    # chain_cap (executor.rs:330-336)
    assert [px.chain_cap(s) for s in (px.STRATEGY_FAST, px.STRATEGY_BALANCED,
                                      px.STRATEGY_SENSITIVE)] == [1, 2, 50]
    q = np.frombuffer(b"A" * 100, np.uint8)
    starts = np.arange(0, 6000, 100, dtype=np.uint64)  # 60 chains
    reads = (L.CReadChains * 3)()
    reads[0] = L.CReadChains(q.ctypes.data, 100, starts.ctypes.data, 60)
    reads[1] = L.CReadChains(q.ctypes.data, 100, None, 0)            # no chain: anchor (0, 0)
    reads[2] = L.CReadChains(q.ctypes.data, 100, starts.ctypes.data, 1)
    for strat, want in ((px.STRATEGY_FAST, 3), (px.STRATEGY_BALANCED, 4), (px.STRATEGY_SENSITIVE, 52)):
        na, nr = ctypes.c_uint64(0), ctypes.c_uint64(0)
        assert L.lib().phx_extend_chains_bounds(strat, 10_000, reads, 3, ctypes.byref(na),
                                                ctypes.byref(nr)) == L.PHX_OK
        assert na.value == want
        assert nr.value == want * (100 + 200 + 1)  # every window is full length here
