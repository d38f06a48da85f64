"""Parity at full workload sizes, where the oracle (which keeps upstream's
full-width score frames: about 9 GB for one 20 kbp pair at 8%) cannot run.

Every alignment the GPU returns is checked by tests/native/alncheck.cpp
(pinned in tests/test_alncheck.py): CIGAR replay over the real bytes, cost of
the CIGAR == score, edit distance == non-M bases, the whole target consumed,
and score == the optimum of a banded Gotoh / Levenshtein DP whose band holds
every alignment of cost <= score, so the score is the true optimum.
Covered: all 200,000 pairs of the mixed-length workload (100 bp - 20 kbp,
long reads at 8%), contig gaps of 3-20 kbp of 100 kbp - 1 Mbp contigs (both
engines) with their stitched contigs, served long requests up to 50 kbp, and
all 1,000,000 Myers pairs against the oracle's forward-only engine (which
runs at full size).  Caps that are not multiples of the score step g = 2
(costs 4/6/2) are compared with the oracle at sizes it handles."""
from __future__ import annotations

import ctypes
import os
import sys
import time

import numpy as np
import pytest

import alncheck as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
FLAG_NAMES = {C.F_REPLAY: "replay", C.F_CONSUME: "consumed", C.F_COST: "cost!=score",
              C.F_ED: "edit distance", C.F_OPT: "not optimal", C.F_TEND: "t_end!=tn"}


def say(msg):
    print(f"[gpu-fullsize] {msg}", flush=True)
    sys.stdout.flush()


def check_all(label, q_seq, q_off, t_seq, t_off, aln, runs, run_off, costs, optimal=True, chunks=8):
    """Checks every pair in `chunks` interleaved slices (progress after each)."""
    n = len(q_off) - 1
    assert (aln["status"] == 0).all(), f"{label}: statuses {np.unique(aln['status'])}"
    t_end = aln["target_end"] - aln["target_start"]
    q_len = np.diff(q_off.astype(np.int64))
    assert (aln["query_end"] - aln["query_start"] == q_len).all(), f"{label}: query ends"
    flags = np.zeros(n, np.uint8)
    cells = 0
    t0 = time.time()
    for c in range(chunks):
        idx = np.arange(c, n, chunks)
        if len(idx) == 0:
            continue
        qs = [q_seq[q_off[i]:q_off[i + 1]] for i in idx]
        ts = [t_seq[t_off[i]:t_off[i + 1]] for i in idx]
        rs = [runs[run_off[i]:run_off[i + 1]] for i in idx]
        off = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.uint64)
        cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)
        f, k = C.check(cat(qs, np.uint8), off(qs), cat(ts, np.uint8), off(ts), aln["score"][idx],
                       aln["edit_distance"][idx], t_end[idx], cat(rs, np.uint32), off(rs),
                       costs=costs, optimal=optimal, threads=THREADS)
        flags[idx] = f
        cells += k
        say(f"{label}: checked {c + 1}/{chunks} ({cells / 1e9:.1f} G DP cells, {time.time() - t0:.0f} s)")
    bad = np.nonzero(flags)[0]
    if len(bad):
        kinds = {name: int(((flags & bit) != 0).sum()) for bit, name in FLAG_NAMES.items()}
        first = ", ".join(f"pair {i} ({int(q_len[i])} bp): flags {int(flags[i])}" for i in bad[:5])
        pytest.fail(f"{label}: {len(bad)} of {n} alignments fail {kinds}; {first}")
    return cells


def test_mixed_all_200k_pairs():
    """C3 in full: 200,000 pairs, 100 bp - 20 kbp, 70% at 1% and 30% at 8%
    (60% indels), affine 4/6/2, global (window 1.1 L)."""
    import phraya_amd as px
    from phraya_amd import synth
    t0 = time.time()
    w = synth.workload("mixed")
    qn = np.diff(w.q_off.astype(np.int64))
    say(f"mixed: {w.n} pairs generated in {time.time() - t0:.0f} s; {(qn > 2500).sum()} above 2.5 kbp, "
        f"{((qn > 6000) & (w.cls == 1)).sum()} long-class above 6 kbp, longest {qn.max()}")
    pb = px.PackedBatch(w.q_seq, w.q_off, w.t_seq, w.t_off, np.zeros(w.n, np.uint64),
                        np.zeros(w.n, np.uint64))
    plan = px.Plan(pb, px.ENGINE_AFFINE, px.AffineCosts(4, 6, 2))
    plan.run()
    res = plan.results()
    plan.close()
    say(f"mixed: aligned on the GPU ({time.time() - t0:.0f} s)")
    cells = check_all("mixed", w.q_seq, w.q_off, w.t_seq, w.t_off, res.aln, res.runs, res.run_off,
                      (4, 6, 2))
    say(f"mixed: all {w.n} pairs bit-checked and optimal ({cells / 1e12:.2f} T DP cells)")


def test_mixed_default_costs_other_seed():
    """The same mixed-length regime under AffineCosts::default() (2/4/1,
    whose score step is 1 and whose register wavefront keeps five m frames)
    on another read seed: 100,000 pairs, every one checked optimal, and the
    one-wave, 256-, 512- and 1024-thread register tiers all in the run."""
    import phraya_amd as px
    from phraya_amd import synth
    w = synth.workload("mixed", n=100_000, seed=12)
    pb = px.PackedBatch(w.q_seq, w.q_off, w.t_seq, w.t_off, np.zeros(w.n, np.uint64),
                        np.zeros(w.n, np.uint64))
    plan = px.Plan(pb, px.ENGINE_AFFINE, px.AffineCosts(2, 4, 1))
    plan.run()
    res = plan.results()
    regs = {li["tier"] for li in plan.last_launch_info() if li["storage"] == 2}
    plan.close()
    say(f"mixed 2/4/1: register tiers {sorted(regs)}")
    assert {2, 3, 4, 5} <= regs, f"register tiers {sorted(regs)}"
    check_all("mixed 2/4/1", w.q_seq, w.q_off, w.t_seq, w.t_off, res.aln, res.runs, res.run_off, (2, 4, 1),
              chunks=4)


def test_myers150_all_1m_pairs():
    """C2 in full: every one of the 1,000,000 distance + end-column results
    equals the oracle's forward-only Myers (myers_extend without backtrace)."""
    import phraya_amd as px
    from phraya_amd import synth
    w = synth.workload("myers150")
    pb = px.PackedBatch(w.q_seq, w.q_off, w.t_seq, w.t_off, np.zeros(w.n, np.uint64),
                        np.zeros(w.n, np.uint64))
    plan = px.Plan(pb, px.ENGINE_MYERS_DIST)
    plan.run()
    res = plan.results()
    plan.close()
    say(f"myers150: {w.n} pairs aligned; oracle forward-only on {THREADS} threads")
    ores, _ = O.batch(O.MYERS_FWD, w.q_seq, w.q_off, w.t_seq, w.t_off, threads=THREADS,
                      want_cigars=False)
    assert (res.aln["status"] == 0).all() and (ores["status"] == 0).all()
    for f in ("edit_distance", "target_end"):
        bad = np.nonzero(res.aln[f] != ores[f])[0]
        assert len(bad) == 0, f"{f}: {len(bad)} mismatches, first {bad[:5]}"
    say("myers150: all 1,000,000 distances and end columns equal the oracle's")


def _gap_results(plan, n_gaps):
    """Per-gap alignments of a contig plan's batched gap plan."""
    from phraya_amd import _lib as L
    lib = L.lib()
    gp = lib.phx_contig_plan_gap_plan(plan._h)
    aln = np.zeros(n_gaps, L.ALN_DTYPE)
    off = np.zeros(n_gaps + 1, np.uint64)
    r = L.CResults(aln.ctypes.data, None, 0, off.ctypes.data)
    rc = lib.phx_plan_results(gp, ctypes.byref(r))
    runs = np.zeros(max(int(off[-1]), 1), np.uint32)
    if rc == L.PHX_E_CAPACITY:
        r = L.CResults(aln.ctypes.data, runs.ctypes.data, len(runs), off.ctypes.data)
        rc = lib.phx_plan_results(gp, ctypes.byref(r))
    assert rc == L.PHX_OK, rc
    return aln, runs, off


@pytest.mark.parametrize("engine", ["affine", "linear"])
def test_contig_gaps_3_to_20_kbp(engine):
    """C4 gaps at their real lengths: 40 contig pairs of 100 kbp - 1 Mbp
    (0.5-2% divergence, indels up to 50 bp) anchored every 3-20 kbp; every gap
    alignment is checked (global, optimal) and every stitched contig replays
    with score == cost."""
    import phraya_amd as px
    from phraya_amd import synth
    w = synth.workload("contig", n=40)
    costs = (4, 6, 2) if engine == "affine" else None
    eng = px.ENGINE_AFFINE if engine == "affine" else px.ENGINE_LINEAR
    params = px.ContigParams(gap_min=3000, gap_max=20000, seed=21)
    plan = px.ContigPlan(w.q_seq, w.q_off, w.t_seq, w.t_off, eng, px.AffineCosts(4, 6, 2), params)
    plan.run()
    res = plan.results()
    G = plan.gaps
    aln, runs, off = _gap_results(plan, G)
    qs, ts = [], []
    lens = []
    for c in range(w.n):
        qc = w.q_seq[w.q_off[c]:w.q_off[c + 1]]
        tc = w.t_seq[w.t_off[c]:w.t_off[c + 1]]
        for q0, q1, t0, t1 in plan.layout(c).astype(np.int64):
            qs.append(qc[q0:q1])
            ts.append(tc[t0:t1])
            lens.append(q1 - q0)
    assert len(qs) == G
    lens = np.array(lens)
    say(f"contig {engine}: {G} gaps, {(lens >= 3000).sum()} of 3 kbp or more, longest {lens.max()}")
    o = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.uint64)
    check_all(f"contig {engine} gaps", np.concatenate(qs), o(qs), np.concatenate(ts), o(ts), aln, runs,
              off, costs, chunks=2)
    check_all(f"contig {engine} stitched", w.q_seq, w.q_off, w.t_seq, w.t_off, res.aln, res.runs,
              res.run_off, costs, optimal=False, chunks=1)
    plan.close()


def test_served_long_requests_to_50_kbp():
    """C5's long classes through the request server: reads of 1-50 kbp at 8%
    (indel-heavy) and contig gaps of 5-20 kbp against windows of the resident
    5 Mbp reference, every response checked."""
    import phraya_amd as px
    from phraya_amd import synth
    st = synth.serve_stream(240, mix=(0.0, 0.8, 0.2), seed=0xC0FFEE15)
    qn = np.diff(st.q_off.astype(np.int64))
    say(f"serve: {st.n} requests, {(qn >= 30000).sum()} of 30 kbp or more, longest {qn.max()}")
    costs = px.AffineCosts(4, 6, 2)
    with px.Server(px.ENGINE_AFFINE, costs, reference=st.genome) as srv:
        srv.submit_arrays(np.arange(st.n), st.q_seq, st.q_off, ref_start=st.pos, ref_len=st.win)
        got = {}
        t0 = time.time()
        while len(got) < st.n:
            for r in srv.poll(4096, timeout_us=20000):
                got[r.tag] = r
            assert time.time() - t0 < 240, f"{len(got)} of {st.n} responses"
    say(f"serve: {st.n} responses in {time.time() - t0:.1f} s")
    from phraya_amd import _lib as L
    aln = np.zeros(st.n, L.ALN_DTYPE)
    rs = []
    for i in range(st.n):
        r = got[i]
        aln[i] = (r.status, 0, r.score, r.edit_distance, r.query_start, r.query_end, r.target_start,
                  r.target_end)
        rs.append(np.asarray(r.runs, np.uint32))
    ts = [st.genome[st.pos[i]:st.pos[i] + st.win[i]] for i in range(st.n)]
    o = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.uint64)
    check_all("serve long", st.q_seq, st.q_off, np.concatenate(ts), o(ts), aln, np.concatenate(rs),
              o(rs), (4, 6, 2), chunks=2)


@pytest.mark.parametrize("cap", [0, 1, 7, 13, 45, 101])
def test_caps_not_multiples_of_the_score_step(cap):
    """Costs 4/6/2 explore in score steps of g = 2; an odd cap floors to the
    last whole step below it, as upstream's loop bound `s <= cap` does.
    Global and fitting pairs, against the oracle with the same cap."""
    from test_gpu_parity import compare, random_batch
    import phraya_amd as px
    qs, ts, _ = random_batch(700 + cap, 300, 200)
    compare(px, "affine", qs, ts, costs=(4, 6, 2), cap=cap, label=f"cap {cap}")


@pytest.mark.parametrize("cap", [501, 2001])
def test_caps_on_workgroup_tiers(cap):
    """Odd caps on pairs long enough for the HBM and workgroup tiers (3-9 kbp
    at 2-8%), some reaching their cap, some not."""
    from test_gpu_parity import compare, mutate, rand_seq
    import random
    import phraya_amd as px
    rng = random.Random(cap)
    qs, ts = [], []
    for k in range(12):
        L = rng.choice([3000, 5000, 9000])
        t = rand_seq(rng, int(L * 1.1))
        qs.append(mutate(rng, t[:L], rng.choice([0.02, 0.08])))
        ts.append(t)
    compare(px, "affine", qs, ts, costs=(4, 6, 2), cap=cap, label=f"long cap {cap}")


def test_queries_past_the_register_frames():
    """The register wavefront keeps offset + 1 in 16-bit frame halves, so a
    query longer than 65,534 bases must run on a ring tier even when its band
    fits the register window and its sequences fit in LDS (a 70 kbp pair at
    0.2% does both).  Pairs on either side of the limit, every one checked
    optimal, and none past it on a register launch."""
    import random
    import phraya_amd as px
    from test_gpu_parity import mutate, rand_seq
    rng = random.Random(65535)
    qs, ts = [], []
    for L in (60000, 65534, 65535, 65536, 70000, 72000):
        q = rand_seq(rng, L)
        qs.append(q)
        ts.append(mutate(rng, q, 0.002))
    pb = px.pack_batch(qs, ts)
    plan = px.Plan(pb, px.ENGINE_AFFINE, px.AffineCosts(4, 6, 2))
    plan.run()
    res = plan.results()
    for li in plan.last_launch_info():
        say(f"register frames: launch tier {li['tier']} storage {li['storage']} pairs {li['pairs']}")
    regs = sum(li["pairs"] for li in plan.last_launch_info() if li["tier"] >= 2 and li["storage"] == 2)
    plan.close()
    assert regs <= 2, f"{regs} pairs on register launches; only the two of <= 65,534 bases may be"
    o = lambda xs: np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.uint64)
    cat = lambda xs: np.frombuffer(b"".join(xs), np.uint8)
    check_all("register frames", cat(qs), o(qs), cat(ts), o(ts), res.aln, res.runs, res.run_off,
              (4, 6, 2), chunks=1)
