"""Contig mode on the GPU: every stitched contig CIGAR equals the oracle's gap
alignments on the same layout stitched the same way, and replays exactly.
Gaps are pinned by anchors (or a contig end) on both sides, so each is the
global branch alone (fill_wfa / fill_wfa_affine_generic with fitting = false,
phraya-align/src/wfa_simd.rs:231-345, :581-779), also for gap shapes on which
the per-pair wrappers would switch to fitting; the contig score is therefore
the cost of its stitched CIGAR."""
from __future__ import annotations

import random
import re
import sys

import numpy as np
import pytest

import dp
from oracle import oracle as O
from test_contig_anchors import make_pair

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-contig] {msg}", flush=True)
    sys.stdout.flush()


def parse(cigar: str):
    return [(int(n), op) for n, op in re.findall(r"(\d+)([MXID])", cigar)]


def stitch(parts, k):
    """parts: list of gap CIGAR runs; anchors of k bases between parts."""
    out = []

    def put(n, op):
        if n <= 0:
            return
        if out and out[-1][1] == op:
            out[-1] = (out[-1][0] + n, op)
        else:
            out.append((n, op))
    for j, runs in enumerate(parts):
        for n, op in runs:
            put(n, op)
        if j + 1 < len(parts):
            put(k, "M")
    return "".join(f"{n}{op}" for n, op in out)


def replay_ok(q: bytes, t: bytes, cigar: str) -> bool:
    i = j = 0
    for n, op in parse(cigar):
        if op == "M":
            if q[i:i + n] != t[j:j + n]:
                return False
            i += n
            j += n
        elif op == "X":
            if any(q[i + d] == t[j + d] for d in range(n)):
                return False
            i += n
            j += n
        elif op == "I":
            j += n
        else:
            i += n
    return i == len(q) and j == len(t)


@pytest.fixture(scope="module")
def contigs():
    rng = random.Random(31)
    qs, ts = [], []
    for s in range(6):
        q, t = make_pair(100 + s, n=rng.randint(6000, 30000), rate=rng.choice([0.005, 0.01, 0.02]),
                         with_n=(s == 3))
        qs.append(q)
        ts.append(t)
    # A 1,500-base target insertion: the gap holding it has tn > qn + qn/2 + 10
    # (a shape on which the wrappers would take the fitting branch).
    q, t = make_pair(200, n=20000, rate=0.01)
    ins = bytes(rng.choice(b"ACGT") for _ in range(1500))
    qs.append(q)
    ts.append(bytes(rng.choice(b"ACGT") for _ in range(300)) + t[:10000] + ins + t[10000:])
    # A query that ends where its last anchor ends (with this pair's spacing
    # draws, which depend on its index 7 in the batch): the final gap has no
    # query bases and 5,843 target bases (global: one I run).
    q, t = make_pair(300, n=20000, rate=0.005)
    qs.append(q[:14209])
    ts.append(t)
    return qs, ts


@pytest.mark.parametrize("engine", ["linear", "affine"])
def test_contig_stitch_matches_oracle(contigs, engine):
    import phraya_amd as px
    qs, ts = contigs
    q_off = np.zeros(len(qs) + 1, np.uint64)
    t_off = np.zeros(len(ts) + 1, np.uint64)
    q_off[1:] = np.cumsum([len(x) for x in qs])
    t_off[1:] = np.cumsum([len(x) for x in ts])
    q_seq = np.frombuffer(b"".join(qs), np.uint8)
    t_seq = np.frombuffer(b"".join(ts), np.uint8)
    params = px.ContigParams(gap_min=400, gap_max=3000, seed=5)
    eng = px.ENGINE_AFFINE if engine == "affine" else px.ENGINE_LINEAR
    plan = px.ContigPlan(q_seq, q_off, t_seq, t_off, eng, px.AffineCosts(4, 6, 2), params)
    say(f"{engine}: {plan.gaps} gaps over {len(qs)} contigs")
    plan.run()
    res = plan.results()
    shapes = {"fitting_shaped": 0, "empty_query": 0}
    for c in range(len(qs)):
        lay = plan.layout(c)
        assert int(lay[0][0]) == 0 and int(lay[-1][1]) == len(qs[c]) and int(lay[-1][3]) == len(ts[c])
        parts, score = [], 0
        for q0, q1, t0, t1 in lay.astype(np.int64):
            qw, tw = qs[c][q0:q1], ts[c][t0:t1]
            if t1 - t0 > (q1 - q0) + (q1 - q0) // 2 + 10:
                shapes["fitting_shaped"] += 1
            if q1 == q0 and t1 > t0:
                shapes["empty_query"] += 1
            o = O.wfa_affine_global(qw, tw, (4, 6, 2)) if engine == "affine" else O.wfa_global(qw, tw)
            assert o.status == O.OK and o.target_end == t1 - t0
            parts.append(parse(o.cigar))
            score += o.score
        want = stitch(parts, params.k)
        got = res.cigar(c)
        assert got == want, f"contig {c}: stitched CIGAR differs"
        assert int(res.aln[c]["score"]) == score
        assert replay_ok(qs[c], ts[c], got), f"contig {c}: CIGAR does not replay"
        # Anchors are exact matches: the contig score is the cost of its CIGAR.
        cost = dp.affine_cost(got, 4, 6, 2) if engine == "affine" else dp.unit_cost(got)
        assert int(res.aln[c]["score"]) == cost, f"contig {c}: score {res.aln[c]['score']} != cost {cost}"
        ed = sum(n for n, op in parse(got) if op != "M")
        assert int(res.aln[c]["edit_distance"]) == ed
    assert shapes["fitting_shaped"] >= 1 and shapes["empty_query"] >= 1, shapes


def test_contig_gap_longer_than_device_format_is_refused():
    """A contig pair with no shared minimizers (here: no k-mers at all) is one
    gap as long as the pair;
    beyond 2^27 - 2 bases plan creation refuses it (before any upload or
    launch) instead of wrapping the 32-bit lengths."""
    import phraya_amd as px
    n = (1 << 27) + 64
    q_seq = np.full(n, ord("N"), np.uint8)
    t_seq = np.full(n, ord("N"), np.uint8)
    off = np.array([0, n], np.uint64)
    with pytest.raises(RuntimeError, match="longer than the device format"):
        px.ContigPlan(q_seq, off, t_seq, off, px.ENGINE_AFFINE, px.AffineCosts(4, 6, 2),
                      px.ContigParams(gap_min=400, gap_max=3000, seed=5))
