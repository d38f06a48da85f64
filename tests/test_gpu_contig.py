"""Contig mode on the GPU: every stitched contig CIGAR equals the oracle's gap
alignments on the same layout stitched the same way, and replays exactly."""
from __future__ import annotations

import random
import re
import sys

import numpy as np
import pytest

from oracle import oracle as O
from test_contig_anchors import make_pair

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-contig] {msg}", flush=True)
    sys.stdout.flush()


def parse(cigar: str):
    return [(int(n), op) for n, op in re.findall(r"(\d+)([MXID])", cigar)]


def stitch(parts, k):
    """parts: list of (runs, t_len, t_end); anchors of k bases between parts."""
    out = []

    def put(n, op):
        if n <= 0:
            return
        if out and out[-1][1] == op:
            out[-1] = (out[-1][0] + n, op)
        else:
            out.append((n, op))
    for j, (runs, tl, te) in enumerate(parts):
        for n, op in runs:
            put(n, op)
        put(tl - te, "I")
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
    for c in range(len(qs)):
        lay = plan.layout(c)
        assert int(lay[0][0]) == 0 and int(lay[-1][1]) == len(qs[c]) and int(lay[-1][3]) == len(ts[c])
        parts, score = [], 0
        for q0, q1, t0, t1 in lay.astype(np.int64):
            qw, tw = qs[c][q0:q1], ts[c][t0:t1]
            o = O.wfa_extend_affine(qw, tw, (0, 0), (4, 6, 2)) if engine == "affine" \
                else O.wfa_extend(qw, tw)
            assert o.status == O.OK
            parts.append((parse(o.cigar), t1 - t0, o.target_end))
            score += o.score
        want = stitch(parts, params.k)
        got = res.cigar(c)
        assert got == want, f"contig {c}: stitched CIGAR differs"
        assert int(res.aln[c]["score"]) == score
        assert replay_ok(qs[c], ts[c], got), f"contig {c}: CIGAR does not replay"
        ed = sum(n for n, op in parse(got) if op != "M")
        assert int(res.aln[c]["edit_distance"]) == ed
