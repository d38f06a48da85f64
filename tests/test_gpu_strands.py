"""align_read's strand handling on the GPU (phx_align_reads_batch through the
C ABI): upstream's strand and outcome scenarios (executor.rs:1334-1473,
:1969-1995) with the chains their seeding would give spelled out, and a
random batch against the oracle run orientation by orientation, anchor by
anchor, with upstream's selection (executor.rs:676-733)."""
from __future__ import annotations

import random
import sys

import numpy as np
import pytest

from oracle import oracle as O
from test_executor_scoring import brute
from test_gpu_executor import mutate, reference

pytestmark = pytest.mark.gpu

FAST, BALANCED, SENSITIVE = 0, 1, 2


def say(msg):
    print(f"[gpu-strands] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def px():
    import phraya_amd
    return phraya_amd


def dna(n, seed):
    rng = random.Random(seed)
    return bytes(rng.choice(b"ACGT") for _ in range(n))


def snp_at(px, placed, target, pos):
    """The kept primary's variant sites at pos (extract_variants on its runs)."""
    prim = placed.scored.primary
    ops = {"M": 0, "X": 1, "I": 2, "D": 3}
    runs, num = [], ""
    for ch in prim.cigar:
        if ch.isdigit():
            num += ch
        else:
            runs.append((int(num) << 4) | ops[ch])
            num = ""
    v = px.extract_variants(runs, [0, len(runs)], [prim.target_start], [placed.query], target,
                            [[(prim.target_start, prim.target_end)]], None, None, 25)
    return [(int(r["position"]), int(r["type"]), chr(r["ref_base"]), chr(r["alt_base"]))
            for r in v.variants if int(r["position"]) == pos]


def fields(a):
    return (a.cigar, a.edit_distance, a.query_start, a.query_end, a.target_start, a.target_end)


def rc_ref(q):
    """reverse_complement restated (types.rs:91-107; the library's own is
    pinned by tests/test_strands.py)."""
    table = {65: 84, 84: 65, 67: 71, 71: 67, 97: 116, 116: 97, 99: 103, 103: 99}
    return bytes(table.get(b, b) for b in reversed(q))


@pytest.mark.parametrize("strategy", [BALANCED, SENSITIVE])
def test_reverse_and_forward_reads(px, strategy):
    # executor.rs:1334-1387: the reverse complement of a mutated region
    # [100, 200) places at 100 on the reverse strand with the forward-strand
    # SNP at 150; :1441-1473: the forward read is marked forward.  Seeding the
    # wrong orientation chains to nothing (:1475-1500), so the chains are
    # [] / [100] and [100] / [].
    target = dna(300, 7)
    region = bytearray(target[100:200])
    orig = region[50]
    region[50] = ord("C") if orig == ord("A") else ord("A")
    rev_read = px.reverse_complement(bytes(region))
    got = px.align_reads_batch(strategy, 0, target, [(rev_read, [], [100], True),
                                                     (bytes(region), [100], [], True)])
    r, f = got
    assert (r.strand, r.outcome) == (px.STRAND_REVERSE, px.OUTCOME_PLACED)
    assert r.forward is None and r.reverse is not None
    assert abs(r.scored.primary.target_start - 100) <= 1 and r.primary_score > 0.9
    assert snp_at(px, r, target, 150) == [(150, 0, chr(orig), chr(region[50]))]
    assert (f.strand, f.outcome) == (px.STRAND_FORWARD, px.OUTCOME_PLACED)
    assert f.reverse is None
    assert snp_at(px, f, target, 150) == [(150, 0, chr(orig), chr(region[50]))]


def test_protein_never_extends_the_reverse_strand(px):
    # executor.rs:1969-1995: with forward chains, a protein read is aligned
    # forward only, whatever reverse chains the caller passes.
    rng = random.Random(11)
    target = bytes(rng.choice(b"ACDEFGHIKLMNPQRSTVWY") for _ in range(300))
    region = bytearray(target[100:200])
    region[50] = ord("C") if region[50] == ord("A") else ord("A")
    (p,) = px.align_reads_batch(SENSITIVE, 0, target, [(bytes(region), [100], [100], True)],
                                is_dna=False)
    assert (p.strand, p.reverse, p.outcome) == (px.STRAND_FORWARD, None, px.OUTCOME_PLACED)
    assert p.scored.primary.edit_distance == 1


def test_outcomes(px):
    # executor.rs:1389-1439: a perfect read is placed, a seeded read with 15
    # clustered mismatches in 150 falls below 0.95, an unrelated read with no
    # seeds is no-seed; :716-723: Fast drops divergence above 0.20.
    target = dna(300, 3)
    good = target[50:150]
    div = bytearray(target[50:200])
    for i in range(67, 82):
        div[i] = ord("C") if div[i] == ord("A") else ord("A")
    junk = dna(120, 987_654)
    got = px.align_reads_batch(BALANCED, 0, target, [(good, [50], [], True),
                                                     (bytes(div), [50], [], True),
                                                     (junk, [], [], False)])
    assert [g.outcome for g in got] == [px.OUTCOME_PLACED, px.OUTCOME_BELOW_THRESHOLD,
                                        px.OUTCOME_NO_SEED]
    assert got[2].forward is not None and got[2].reverse is not None  # both orientations tried
    far = bytearray(target[50:150])
    for i in range(0, 100, 4):
        far[i] = ord("C") if far[i] == ord("A") else ord("A")      # 25 mismatches / 100
    (x,) = px.align_reads_batch(FAST, 0, target, [(bytes(far), [50], [], True)])
    assert x.outcome == px.OUTCOME_FAST_CUTOFF and x.primary_score < 0.8
    (y,) = px.align_reads_batch(BALANCED, 0, target, [(bytes(far), [50], [], True)])
    assert y.outcome == px.OUTCOME_BELOW_THRESHOLD


def expected(strategy, gap, target, q, fwd, rev, seeds, is_dna):
    """align_read after seed_and_chain, from the oracle: orientations per
    :676-698, each orientation's extend_chains anchor by anchor, the better
    one per :700-711 (ties forward), then :716-733."""
    nr = rev if is_dna else []
    rcq = bytes(rc_ref(q)) if is_dna else b""
    todo = {"f": not (not fwd and nr), "r": not (fwd and not nr)}
    res = {}
    for k, (qq, ch) in {"f": (q, fwd), "r": (rcq, nr)}.items():
        if not todo[k]:
            res[k] = None
            continue
        ok = [a for a in reference(strategy, gap, target, qq, ch) if a.status == O.OK]
        res[k] = brute(ok, len(qq)) if ok else None
    f, r = res["f"], res["r"]
    if f and r and r[0].edit_distance < f[0].edit_distance:
        strand, kept = 1, r
    elif f:
        strand, kept = 0, f
    elif r:
        strand, kept = 1, r
    else:
        return (-1, 4, None, todo)
    ql = np.float64(max(len(q), 1))
    ed = np.float64(kept[0].edit_distance)
    score = np.float64(1.0) - ed / ql
    if strategy == FAST and ed / ql > 0.20:
        outcome = 3
    elif score >= 0.95:
        outcome = 0
    else:
        outcome = 2 if seeds else 1
    return (strand, outcome, kept, todo)


@pytest.mark.parametrize("resident", [False, True], ids=["host_target", "resident_reference"])
@pytest.mark.parametrize("strategy,gap,is_dna", [(SENSITIVE, 0, True), (BALANCED, 0, True),
                                                 (FAST, 0, True), (SENSITIVE, 1, False)])
def test_random_reads_match_the_executor(px, strategy, gap, is_dna, resident):
    rng = random.Random(500 + 10 * strategy + gap + int(is_dna))
    G = 20_000
    genome = dna(G, 41)
    reads = []
    for i in range(160):
        L = rng.choice([0, 1, 60, 150, 150, 600])
        pos = rng.randrange(0, G - L - 50)
        src = mutate(rng, genome[pos:pos + L], rng.choice([0.0, 0.02, 0.1, 0.3]))
        q = rc_ref(src) if i % 3 == 0 else src           # every third read from the reverse strand
        kind = i % 7
        right = [pos + rng.randint(-2, 2)]
        noise = [rng.randrange(0, G) for _ in range(rng.randint(0, 3))]
        fwd, rev = {0: (right + noise, []), 1: ([], right + noise), 2: (right, noise),
                    3: (noise, right), 4: ([], []), 5: (right, right),
                    6: ([G + 5], [G + 9])}[kind]                     # 6: every anchor fails
        seeds = kind != 4 or i % 14 == 4
        reads.append((q, [max(0, s) for s in fwd], [max(0, s) for s in rev], seeds))
    say(f"strategy {strategy} gap {gap} dna {is_dna}: {len(reads)} reads")
    got = px.align_reads_batch(strategy, gap, px.Reference(genome) if resident else genome, reads,
                               is_dna=is_dna)
    counts = [0] * 5
    for i, ((q, fwd, rev, seeds), g) in enumerate(zip(reads, got)):
        strand, outcome, kept, todo = expected(strategy, gap, genome, q, fwd, rev, seeds, is_dna)
        assert (g.forward is not None, g.reverse is not None) == (todo["f"], todo["r"]), f"read {i}"
        assert (g.strand, g.outcome) == (strand, outcome), f"read {i}"
        counts[outcome] += 1
        if kept is None:
            assert g.scored is None
            continue
        assert fields(g.scored.primary) == fields(kept[0]), f"read {i}: primary"
        assert [fields(a) for a in g.scored.alternatives] == [fields(a) for a in kept[1]], \
            f"read {i}: alternatives"
        assert g.query == (q if strand == 0 else (rc_ref(q) if is_dna else b""))
    say(f"outcomes placed/no-seed/below/fast/none: {counts}")
