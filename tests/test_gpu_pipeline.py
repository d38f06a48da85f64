"""The per-read path end to end on the GPU, as align_read composes it
(phraya-align/src/executor.rs:573-788): canonical (21, 11) minimizer
sketches of each read, of its reverse complement and of the genome
(phx_sketch_batch, as sketch() does, phraya-core/src/types.rs:1873-1877) ->
seed_and_chain (phx_seed_and_chain_batch) -> the orientations worth
extending, extended and chosen (phx_align_reads_batch) -> variant sites of
the kept primary with the read's coverage windows (phx_extract_variants) ->
the .phraya payload of those observations (phx_phraya_encode, as the align
command writes it, phraya-cli/src/main.rs:556-639).  Every stage is compared
with the oracle composed the same way."""
from __future__ import annotations

import random
import sys

import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_executor import mutate
from test_gpu_strands import expected, fields, rc_ref

pytestmark = pytest.mark.gpu

OPS = {"M": 0, "X": 1, "I": 2, "D": 3}


def say(msg):
    print(f"[gpu-pipeline] {msg}", file=sys.stderr, flush=True)


def runs_of(cigar):
    out, num = [], ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
        else:
            out.append((int(num) << 4) | OPS[ch])
            num = ""
    return out


@pytest.mark.parametrize("strategy", [1, 2])
def test_reads_from_sketch_to_variant_sites(strategy):
    import phraya_amd as px
    from phraya_amd import synth
    rng = random.Random(20 + strategy)
    genome = synth.genome(200_000, 0x5EED0042).tobytes()
    reads = []
    for i in range(300):
        pos = rng.randrange(0, len(genome) - 200)
        src = mutate(rng, genome[pos:pos + 150], rng.choice([0.0, 0.01, 0.03, 0.3]))
        reads.append(rc_ref(src) if i % 3 == 0 else src)
    reads.append(bytes(rng.choice(b"ACGT") for _ in range(150)))  # unrelated: no seeds either way
    say(f"strategy {strategy}: {len(reads)} reads")
    seqs = [genome] + reads + [rc_ref(r) for r in reads]
    allsk = px.sketch_batch(seqs)
    for s_, k_ in zip(seqs[::37], allsk[::37]):
        v_, p_ = O.sketch(s_, 21, 11, True)
        assert np.array_equal(k_[:, 0], v_) and np.array_equal(k_[:, 1], p_.astype(np.uint64))
    gs, sk = allsk[0], allsk[1:]
    cap = O.seed_occurrence_cap(gs, 256)
    n = len(reads)

    # Seeding and chaining, both orientations in one batch.
    got = px.seed_and_chain(gs, sk, cap)
    ns, ch, off, sd = O.seed_and_chain(gs, sk, cap)
    assert np.array_equal(got.n_seeds, ns) and np.array_equal(got.chain_off, off)
    assert np.array_equal(got.chains, ch) and np.array_equal(got.seeds, sd)
    starts = [[int(c["target_start"]) for c in ch[int(off[k]):int(off[k + 1])]] for k in range(2 * n)]
    batch = [(reads[i], starts[i], starts[n + i], int(ns[i]) + int(ns[n + i]) > 0) for i in range(n)]

    # Extension with align_read's strand rule.
    placed = px.align_reads_batch(strategy, 0, genome, batch)
    kept = []
    for i, ((q, fwd, rev, seeds), g) in enumerate(zip(batch, placed)):
        strand, outcome, want, _ = expected(strategy, 0, genome, q, fwd, rev, seeds, True)
        assert (g.strand, g.outcome) == (strand, outcome), f"read {i}"
        if want is None:
            assert g.scored is None
            continue
        assert fields(g.scored.primary) == fields(want[0]), f"read {i}"
        assert [fields(a) for a in g.scored.alternatives] == [fields(a) for a in want[1]], f"read {i}"
        kept.append(g)
    counts = {k: sum(1 for g in placed if g.outcome == k) for k in range(5)}
    say(f"outcomes {counts}, reverse kept {sum(1 for g in placed if g.strand == 1)}")
    assert counts[0] > 150 and counts[1] > 20 and sum(1 for g in placed if g.strand == 1) > 50

    # Variant sites of every kept primary, coverage windows = its alignments.
    runs, run_off, ts, qs, spans = [], [0], [], [], []
    for g in kept:
        runs += runs_of(g.scored.primary.cigar)
        run_off.append(len(runs))
        ts.append(g.scored.primary.target_start)
        qs.append(g.query)
        spans.append([(a.target_start, a.target_end) for a in [g.scored.primary] + g.scored.alternatives])
    rep = np.array([[k, k + 30] for k in range(0, len(genome), 900)], np.uint64)
    hot = np.array([[k, k + 80] for k in range(400, len(genome), 1700)], np.uint32)
    v = px.extract_variants(runs, run_off, ts, qs, genome, spans, rep, hot, 25)
    q_seq = np.frombuffer(b"".join(qs), np.uint8)
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in qs])]).astype(np.uint64)
    flat = np.array([p for sp in spans for p in sp], np.uint64).reshape(-1, 2)
    sp_off = np.concatenate([[0], np.cumsum([len(s) for s in spans])]).astype(np.uint64)
    ov, ooff, ocov = O.extract_variants(np.asarray(runs, np.uint32), np.asarray(run_off, np.uint64),
                                        np.asarray(ts, np.uint64), q_seq, q_off, genome, flat, sp_off,
                                        rep, hot, 25)
    assert np.array_equal(v.var_off, ooff) and np.array_equal(v.coverage, ocov)
    for f in O.VARIANT_DTYPE.names:
        assert np.array_equal(v.variants[f], ov[f]), f
    say(f"{len(kept)} kept reads, {len(v.variants)} variant sites")
    assert len(v.variants) > 100

    # The .phraya payload: observations with each read's fields, the
    # genome coverage summed over every alignment window (merge_result).
    from oracle import formats as F
    cov = np.zeros(len(genome), np.uint32)
    for sp in spans:
        for a, b in sp:
            cov[min(a, len(genome)):min(b, len(genome))] += 1
    metas = [px.ReadMeta(cigar=g.scored.primary.cigar, provenance=f"sample1:read{i}", query=g.query,
                         edit_distance=g.scored.primary.edit_distance, mapq=60, strand=g.strand,
                         avg_base_quality=40.0, confidence=1.0) for i, g in enumerate(kept)]
    doc = F.unpack(px.encode_phraya(len(genome), "worker_0", "t0", v.variants, v.coverage, metas, cov))
    want = F.phraya_file(len(genome), "worker_0", "t0", ov, ocov,
                         [dict(cigar=m.cigar, provenance=m.provenance, query=m.query,
                               edit_distance=m.edit_distance, mapq=60, strand=m.strand, avg_base_quality=40.0,
                               confidence=1.0) for m in metas], cov)
    assert doc == want
