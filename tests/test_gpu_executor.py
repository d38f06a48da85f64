"""Batched extend_chains (phx_extend_chains_batch) vs the oracle run anchor by
anchor the way phraya-align's executor does it (executor.rs:347-362,
:379-404, :528-567), with score_alignments' selection (lib.rs:170-199)."""
from __future__ import annotations

import random
import sys

import pytest

from oracle import oracle as O
from test_executor_scoring import brute

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-executor] {msg}", flush=True)
    sys.stdout.flush()


def mutate(rng, s, rate):
    out = bytearray()
    i = 0
    while i < len(s):
        r = rng.random()
        if r < rate * 0.2:
            out += bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 3)))
            continue
        if r < rate * 0.4:
            i += rng.randint(1, 3)
            continue
        out.append(rng.choice([c for c in b"ACGT" if c != s[i]]) if r < rate else s[i])
        i += 1
    return bytes(out)


def make_reads(seed, genome, n, long_every=5):
    rng = random.Random(seed)
    reads = []
    G = len(genome)
    for i in range(n):
        L = 600 if i % long_every == 0 else 150
        pos = rng.randrange(0, G - L - 50)
        q = mutate(rng, genome[pos:pos + L], rng.choice([0.0, 0.01, 0.03]))
        kind = i % 7
        if kind == 0:
            starts = []                                  # no chain: anchor (0, 0)
        elif kind == 1:
            starts = [pos, pos]                          # duplicate anchors: copies of the primary
        elif kind == 2:
            starts = [G + 5, pos]                        # first anchor past the target: InvalidInput
        elif kind == 3:
            starts = [G - 40, pos]                       # truncated window at the end
        else:
            starts = [pos + rng.randint(-3, 3)] + [rng.randrange(0, G) for _ in range(rng.randint(0, 4))]
            rng.shuffle(starts)
        reads.append((q, [max(0, s) for s in starts]))
    return reads


def reference(strategy, gap, target, q, starts):
    cap = {0: 1, 1: 2, 2: 50}[strategy]
    anchors = [(0, 0)] if not starts else [(0, s) for s in starts[:cap]]
    out = []
    for qp, tp in anchors:
        win = target[:min(tp + 2 * len(q), len(target))]
        if strategy == 2:
            r = O.wfa_extend_affine(q, win, (qp, tp), (2, 4, 1)) if gap == 0 else O.wfa_extend(q, win, (qp, tp))
        else:
            r = O.myers_extend(q, win, (qp, tp)) if len(q) <= 500 else O.wfa_extend(q, win, (qp, tp))
        out.append(r)
    return out


@pytest.fixture(scope="module")
def genome():
    rng = random.Random(77)
    return bytes(rng.choice(b"ACGT") for _ in range(60_000))


@pytest.fixture(scope="module")
def genome_bytes():
    """A target with N runs and lowercase stretches: windows holding them run
    in the 8-bit encoding (resident references use their 8-bit copy)."""
    rng = random.Random(78)
    g = bytearray(rng.choice(b"ACGT") for _ in range(60_000))
    for _ in range(40):
        p = rng.randrange(0, len(g) - 30)
        k = rng.randint(1, 25)
        g[p:p + k] = b"N" * k
    for _ in range(20):
        p = rng.randrange(0, len(g) - 200)
        g[p:p + 100] = bytes(g[p:p + 100]).lower()
    return bytes(g)


@pytest.mark.parametrize("resident", [False, True], ids=["host_target", "resident_reference"])
@pytest.mark.parametrize("strategy,gap", [(2, 0), (2, 1), (1, 0), (0, 0)])
def test_extend_chains_matches_the_executor(genome, genome_bytes, strategy, gap, resident):
    import phraya_amd as px
    reads = make_reads(100 + 10 * strategy + gap, genome, 210)
    # A third of the batch against the target with N and lowercase bytes.
    extra = make_reads(900 + 10 * strategy + gap, genome_bytes, 70)
    say(f"strategy {strategy} gap {gap} resident {resident}: {len(reads)} + {len(extra)} reads")
    check_batch(px, strategy, gap, genome, reads, resident)
    check_batch(px, strategy, gap, genome_bytes, extra, resident)


def check_batch(px, strategy, gap, genome, reads, resident):
    target = px.Reference(genome) if resident else genome
    got = px.extend_chains_batch(strategy, gap, target, reads)
    n_alt = n_none = 0
    for i, ((q, starts), g) in enumerate(zip(reads, got)):
        want = reference(strategy, gap, genome, q, starts)
        assert len(g.anchors) == len(want), f"read {i}: anchor count"
        ok = []
        for j, (w, (role, a)) in enumerate(zip(want, g.anchors)):
            if w.status != O.OK:
                assert role == 0 and a == w.status, f"read {i} anchor {j}: status"
                continue
            assert role != 0, f"read {i} anchor {j}: extended upstream, failed here"
            got_f = (a.cigar, a.edit_distance, a.query_start, a.query_end, a.target_start, a.target_end)
            want_f = (w.cigar, w.edit_distance, w.query_start, w.query_end, w.target_start, w.target_end)
            assert got_f == want_f, f"read {i} anchor {j}: {got_f[1:]} vs {want_f[1:]}"
            ok.append(a)
        if not ok:
            assert g.scored is None
            n_none += 1
            continue
        bp, balt = brute(ok, len(q))
        assert g.scored is not None and g.scored.primary == bp, f"read {i}: primary"
        assert g.scored.alternatives == balt, f"read {i}: alternatives"
        n_alt += len(balt)
    say(f"checked; {n_alt} alternatives, {n_none} reads without an alignment")
