"""HIP path vs the CPU oracle, bit-exact (scores, edit distances, ends, CIGARs).

Every test goes through the C ABI (libphraya_hip.so via phraya_amd)."""
from __future__ import annotations

import random
import sys

import numpy as np
import pytest

import kat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ENG = {"linear": (0, O.LINEAR), "affine": (1, O.AFFINE), "myers": (2, O.MYERS),
       "myers_global": (3, O.MYERS_GLOBAL), "myers_dist": (4, O.MYERS_FWD)}


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def px():
    import phraya_amd
    return phraya_amd


class GpuImpl(kat.Impl):
    def __init__(self, px):
        self.px = px

    def _d(self, fn, *a, **kw):
        try:
            r = fn(*a, **kw)
        except self.px.InvalidInput:
            return {"status": "InvalidInput"}
        except self.px.AlignmentFailed:
            return {"status": "AlignmentFailed"}
        return dict(status="ok", cigar=r.cigar, edit_distance=r.edit_distance,
                    query_start=r.query_start, query_end=r.query_end,
                    target_start=r.target_start, target_end=r.target_end, score=r.score)

    def wfa_extend(self, q, t, seed, cap=None):
        return self._d(self.px.wfa_extend, q, t, self.px.SeedAnchor(*seed), cap)

    def wfa_extend_affine(self, q, t, seed, costs):
        return self._d(self.px.wfa_extend_affine, q, t, self.px.SeedAnchor(*seed),
                       self.px.AffineCosts(*costs))

    def myers_extend(self, q, t, seed):
        return self._d(self.px.myers_extend, q, t, self.px.SeedAnchor(*seed))

    def myers_edit_distance(self, q, t):
        return self.px.myers_edit_distance(q, t)

    def lcp(self, a, b):
        # The match-extend primitive is exercised through a perfect-prefix alignment:
        # wfa_extend of a vs b (fitting off) starts with an M run of LCP length.
        r = self.px.wfa_extend(a, b)
        first = r.cigar
        n = 0
        num = ""
        for ch in first:
            if ch.isdigit():
                num += ch
            else:
                n = int(num) if ch == "M" else 0
                break
        return n


@pytest.mark.parametrize("k", kat.load(), ids=lambda k: k["src"].split("/")[-1])
def test_gpu_kat(px, k):
    say(f"kat {k['src']}")
    kat.check(k, GpuImpl(px))


# ------------------------------------------------------------------- batch compare

def compare(px, engine, qs, ts, seeds=None, costs=(2, 4, 1), cap=None, ctx=None, label="", res=None):
    """GPU results (res, else one align_batch call) against the oracle."""
    eng_gpu, eng_orc = ENG[engine]
    say(f"compare {engine} {label} n={len(qs)} costs={costs} cap={cap}")
    if res is None:
        res = px.align_batch(eng_gpu, qs, ts, seeds, px.AffineCosts(*costs), cap, ctx)
    pb = px.pack_batch(qs, ts, seeds)
    ores, ocig = O.batch(eng_orc, pb.q_seq, pb.q_off, pb.t_seq, pb.t_off, pb.seed_q, pb.seed_t,
                         costs=costs, cap=-1 if cap is None else cap, threads=8)
    g = res.aln
    bad = []
    for i in range(len(qs)):
        gs, os_ = int(g["status"][i]), int(ores["status"][i])
        if gs != os_:
            bad.append((i, "status", gs, os_))
            continue
        if gs != 0:
            continue
        for f in ("edit_distance", "query_start", "query_end", "target_start", "target_end"):
            if int(g[f][i]) != int(ores[f][i]):
                bad.append((i, f, int(g[f][i]), int(ores[f][i])))
        if engine in ("linear", "affine") and int(g["score"][i]) != int(ores["score"][i]):
            bad.append((i, "score", int(g["score"][i]), int(ores["score"][i])))
        if engine != "myers_dist" and res.cigar(i) != ocig[i]:
            bad.append((i, "cigar", res.cigar(i)[:80], ocig[i][:80]))
    assert not bad, f"{len(bad)} mismatches, first: {bad[:5]}"
    return res


def rand_seq(rng, n, alphabet=b"ACGT"):
    return bytes(rng.choice(alphabet) for _ in range(n))


def mutate(rng, s, rate, alphabet=b"ACGT"):
    out = bytearray()
    i = 0
    while i < len(s):
        r = rng.random()
        if r < rate * 0.25:
            out += rand_seq(rng, rng.randint(1, 3), alphabet)
            continue
        if r < rate * 0.5:
            i += rng.randint(1, 3)
            continue
        out.append(rng.choice([c for c in alphabet if c != s[i]]) if r < rate else s[i])
        i += 1
    return bytes(out)


def random_batch(seed, n, qmax, alphabet=b"ACGT", fitting_frac=0.3, seeds=False):
    rng = random.Random(seed)
    qs, ts, ss = [], [], []
    for _ in range(n):
        qn = rng.randint(0, qmax)
        q = rand_seq(rng, qn, alphabet)
        t = mutate(rng, q, rng.choice([0.0, 0.01, 0.05, 0.15, 0.3]), alphabet)
        if rng.random() < fitting_frac:
            t += rand_seq(rng, qn // 2 + rng.randint(11, 40), alphabet)
        if seeds:
            sq = rng.randint(0, min(len(q), 5))
            st = rng.randint(0, min(len(t), 5))
            if rng.random() < 0.05:
                sq = len(q) + 1  # InvalidInput
            q = rand_seq(rng, sq, alphabet) + q if rng.random() < 0.5 else q
            ss.append((sq, st))
        qs.append(q)
        ts.append(t)
    return qs, ts, (ss if seeds else None)


@pytest.mark.parametrize("engine", ["linear", "affine", "myers", "myers_global", "myers_dist"])
def test_random_small(px, engine):
    qs, ts, ss = random_batch(100 + len(engine), 600, 120, seeds=True)
    compare(px, engine, qs, ts, ss, costs=(4, 6, 2), label="small")


@pytest.mark.parametrize("costs", [(2, 4, 1), (4, 6, 2), (1, 0, 1), (3, 5, 2), (1, 30, 1), (1, 70, 1)])
def test_affine_cost_sets(px, costs):
    qs, ts, _ = random_batch(7 + costs[0] * 10 + costs[1], 300, 90)
    compare(px, "affine", qs, ts, costs=costs, label="costs")


@pytest.mark.parametrize("engine", ["linear", "affine", "myers", "myers_global", "myers_dist"])
def test_bytes_path(px, engine):
    """Pairs with bytes outside uppercase ACGT take the 8-bit encoding."""
    qs, ts, _ = random_batch(55, 300, 100, alphabet=b"ACGTNacgt")
    compare(px, engine, qs, ts, costs=(4, 6, 2), label="bytes")
    prot = b"ACDEFGHIKLMNPQRSTVWY"
    qs, ts, _ = random_batch(56, 100, 80, alphabet=prot)
    compare(px, engine, qs, ts, costs=(4, 6, 2), label="protein")


@pytest.mark.parametrize("engine", ["linear", "affine"])
def test_caps(px, engine):
    rng = random.Random(9)
    qs, ts, _ = random_batch(77, 300, 60)
    for cap in (0, 1, 3, 8, 20):
        compare(px, engine, qs, ts, costs=(2, 4, 1), cap=cap, label=f"cap{cap}")


def rows_batch(seed, n, alphabet=b"ACGT"):
    """Global-branch pairs around the banded-row limits: |qn - tn| from 0 to
    60 in both directions (rows take |qn - tn| <= 48), error rates whose
    optimum lands below and above the band bound of |qn - tn| + 15 steps."""
    rng = random.Random(seed)
    qs, ts = [], []
    while len(qs) < n:
        q = rand_seq(rng, rng.randint(1, 300), alphabet)
        t = mutate(rng, q, rng.choice([0.0, 0.005, 0.02, 0.05, 0.12]), alphabet)
        d = rng.randint(0, 60)
        if rng.random() < 0.6:
            t += rand_seq(rng, d, alphabet)
        else:
            q += rand_seq(rng, d, alphabet)
        t = t[:len(q) + len(q) // 2 + 10]  # global branch
        if q and t:
            qs.append(q)
            ts.append(t)
    return qs, ts


def rows_launches(px, qs, ts, costs, ctx=None, cap=None):
    """(results, storage codes of the run's LDS-tier launches)."""
    pb = px.pack_batch(qs, ts)
    plan = px.Plan(pb, px.ENGINE_AFFINE, px.AffineCosts(*costs), cap, ctx)
    try:
        plan.run()
        kinds = sorted(li["storage"] for li in plan.last_launch_info() if li["tier"] == 0)
        return plan.results(), kinds
    finally:
        plan.close()


@pytest.mark.parametrize("costs", [(4, 6, 2), (2, 4, 1), (1, 0, 1)])
def test_banded_rows(px, costs):
    """Banded rows (16- or 32-diagonal windows sliding with the score step,
    four or two pairs per wave): both window directions, pairs whose optimum
    lies inside and beyond the 16-lane band (routed to 32-lane rows) and
    beyond the 32-lane band (finished on the one-pair path in the same
    launch), 2-bit and byte encodings, caps inside and beyond the band, and
    partial sets of rows; every result equals the oracle's and the one-pair-
    per-wave path's bytes, with the routing on and with every pair in 32-lane
    rows."""
    off_ctx = px.Context(0, dict(lds_rows=2))
    wide_ctx = px.Context(0, dict(lds_rows=3))
    try:
        for label, alphabet, n in (("2bit", b"ACGT", 803), ("bytes", b"ACGTNacgt", 301), ("tiny", b"ACGT", 3)):
            qs, ts = rows_batch(1000 * costs[0] + 100 * costs[1] + n, n, alphabet)
            if n < 4:  # a partial quad of pairs the rows take
                qs, ts = map(list, zip(*[(q, t) for q, t in zip(*rows_batch(n, 40, alphabet))
                                         if abs(len(q) - len(t)) <= 48][:n]))
            for cap in (None, 3, 40):
                res, kinds = rows_launches(px, qs, ts, costs, cap=cap)
                assert any(k >= 3 for k in kinds), f"no banded-row launch: {kinds}"
                compare(px, "affine", qs, ts, costs=costs, cap=cap, label=f"rows-{label}-cap{cap}", res=res)
                off, off_kinds = rows_launches(px, qs, ts, costs, ctx=off_ctx, cap=cap)
                assert all(k == 0 for k in off_kinds)
                assert np.array_equal(res.aln, off.aln) and np.array_equal(res.runs, off.runs) \
                    and np.array_equal(res.run_off, off.run_off)
                wide, wide_kinds = rows_launches(px, qs, ts, costs, ctx=wide_ctx, cap=cap)
                assert any(k >= 3 for k in wide_kinds)
                assert np.array_equal(res.aln, wide.aln) and np.array_equal(res.runs, wide.runs)
    finally:
        off_ctx.close()
        wide_ctx.close()


def test_long_pairs_escalate_tiers(px):
    """Long and divergent pairs overflow the one-wave tier and are re-run on
    the wider tiers; results stay bit-exact."""
    rng = random.Random(1234)
    qs, ts = [], []
    for L in (300, 700, 1500, 2500, 4000, 6000):
        for rate in (0.01, 0.08):
            q = rand_seq(rng, L)
            qs.append(q)
            ts.append(mutate(rng, q, rate) + rand_seq(rng, L // 10))
    for eng in ("affine", "linear"):
        compare(px, eng, qs, ts, costs=(4, 6, 2), label="long")


def test_block_tiers_for_wide_pairs(px):
    """Pairs with qn + tn above 8192 run a workgroup per pair (256 or 1024
    threads); low divergence keeps the oracle's full-width frames small."""
    rng = random.Random(777)
    qs, ts = [], []
    for L, rate in ((5200, 0.03), (6000, 0.01), (9000, 0.01), (9500, 0.005)):
        q = rand_seq(rng, L)
        qs.append(q)
        ts.append(mutate(rng, q, rate) + rand_seq(rng, L // 20))
    for eng in ("affine", "linear"):
        compare(px, eng, qs, ts, costs=(4, 6, 2), label="block-tier")


def test_tiny_tiers_force_every_escalation(px):
    """A context with deliberately small tiers makes most pairs move through
    several tiers; results must not depend on the route."""
    ctx = px.Context(0, dict(wave_lds_kib=1, wave_slots=64, hbm_slots=16, hbm_slot_mib=1,
                             hbm_budget_mib=2048))
    qs, ts, _ = random_batch(4321, 200, 400)
    for eng in ("affine", "linear"):
        compare(px, eng, qs, ts, costs=(4, 6, 2), ctx=ctx, label="tiny-tiers")
    ctx.close()


def test_short150_full_batch(px):
    """The 150 bp / 170 bp affine workload (4/6/2), all 10,000 pairs."""
    from phraya_amd import synth
    w = synth.workload("short150")
    qs = [w.pair(i)[0] for i in range(w.n)]
    ts = [w.pair(i)[1] for i in range(w.n)]
    compare(px, "affine", qs, ts, costs=(4, 6, 2), label="short150")


def test_myers150_subset(px):
    from phraya_amd import synth
    w = synth.workload("myers150", n=50_000)
    qs = [w.pair(i)[0] for i in range(w.n)]
    ts = [w.pair(i)[1] for i in range(w.n)]
    compare(px, "myers_dist", qs, ts, label="myers150")
    compare(px, "myers", qs[:5000], ts[:5000], label="myers150-cigar")


def test_mixed_subset(px):
    from phraya_amd import synth
    w = synth.workload("mixed", n=400)
    idx = [i for i in range(w.n) if w.q_off[i + 1] - w.q_off[i] <= 2500][:120]
    qs = [w.pair(i)[0] for i in idx]
    ts = [w.pair(i)[1] for i in idx]
    compare(px, "affine", qs, ts, costs=(4, 6, 2), label="mixed")


def test_per_pair_entry_points_and_dispatch_name(px):
    """The C per-pair entry points (what an unchanged extend_anchor binds)
    equal the batched path, size the CIGAR text as documented (full length
    returned when the buffer is short), and name the implementation they ran
    per thread (phx_last_impl, upstream's LAST_IMPL)."""
    import ctypes
    from phraya_amd import _lib as L
    lib = L.lib()
    ctx = px.default_context()
    qs, ts, _ = random_batch(31, 40, 150)
    ref = px.align_batch(1, qs, ts, costs=px.AffineCosts(2, 4, 1))
    for i, (q, t) in enumerate(zip(qs, ts)):
        out = L.CAlignment()
        cap = 8
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_size_t(0)
        rc = lib.phx_wfa_extend_affine(ctx.handle, q, len(q), t, len(t), 0, 0, L.Costs(2, 4, 1),
                                       ctypes.byref(out), buf, cap, ctypes.byref(n))
        assert rc == int(ref.aln["status"][i])
        if rc != 0:
            continue
        if n.value >= cap:  # one retry with the full length, as the Rust glue does
            buf = ctypes.create_string_buffer(n.value + 1)
            rc = lib.phx_wfa_extend_affine(ctx.handle, q, len(q), t, len(t), 0, 0, L.Costs(2, 4, 1),
                                           ctypes.byref(out), buf, n.value + 1, ctypes.byref(n))
        assert buf.value.decode() == ref.cigar(i) and n.value == len(ref.cigar(i))
        assert out.score == int(ref.aln["score"][i]) and out.target_end == int(ref.aln["target_end"][i])
    assert px.get_device_dispatch_target().startswith("hip-gfx9")


def test_deterministic_and_isolated(px):
    """Same batch twice gives identical bytes; a pair alone equals the same
    pair inside a mixed batch."""
    qs, ts, _ = random_batch(999, 200, 200)
    a = px.align_batch(1, qs, ts, costs=px.AffineCosts(4, 6, 2))
    b = px.align_batch(1, qs, ts, costs=px.AffineCosts(4, 6, 2))
    assert np.array_equal(a.aln, b.aln) and np.array_equal(a.runs, b.runs)
    for i in (0, 17, 123):
        one = px.wfa_extend_affine(qs[i], ts[i], px.SeedAnchor(), px.AffineCosts(4, 6, 2))
        assert one == a.alignment(i)


def test_results_do_not_depend_on_tuning(px):
    """Regrouping a plan under other tier edges / heavy shares / slot counts,
    and autotuning it, leaves every result byte unchanged."""
    from phraya_amd import synth
    w = synth.workload("mixed", n=300)
    idx = [i for i in range(w.n) if w.q_off[i + 1] - w.q_off[i] <= 6000][:150]
    sub = w.subset(idx)
    pb = px.PackedBatch(sub.q_seq, sub.q_off, sub.t_seq, sub.t_off,
                        np.zeros(sub.n, np.uint64), np.zeros(sub.n, np.uint64))
    plan = px.Plan(pb, px.ENGINE_AFFINE, px.AffineCosts(4, 6, 2))
    plan.run()
    ref = plan.results()
    for t in (dict(block256_min_len=2048, block1024_min_len=4096, heavy_permille=300),
              dict(block256_min_len=1 << 20, block1024_min_len=1 << 21, heavy_permille=1),
              dict(hbm_slots=64, wave_lds_kib=5, wave_slots=256)):
        say(f"tuning {t}")
        plan.set_tuning(t)
        plan.run()
        r = plan.results()
        assert np.array_equal(r.aln, ref.aln) and np.array_equal(r.runs, ref.runs)
    tuned, ms = plan.autotune(max_trials=6)
    say(f"autotuned {tuned} {ms:.2f} ms")
    plan.run()
    r = plan.results()
    assert np.array_equal(r.aln, ref.aln) and np.array_equal(r.runs, ref.runs)
    plan.close()


def test_graph_replays_match_eager_runs(px):
    """Graph mode: the launch sequence recorded once and replayed gives the
    eager run's results on every replay, with other eager work (another plan
    on the same context, result downloads) between replays, across a retune
    (re-recorded graph), and with overflow re-runs after the replay."""
    from phraya_amd import synth
    w = synth.workload("mixed", n=300)
    idx = [i for i in range(w.n) if w.q_off[i + 1] - w.q_off[i] <= 6000][:120]
    sub = w.subset(idx)
    pb = px.PackedBatch(sub.q_seq, sub.q_off, sub.t_seq, sub.t_off,
                        np.zeros(sub.n, np.uint64), np.zeros(sub.n, np.uint64))
    # A narrow LDS ring makes some workgroup-tier pairs overflow every run.
    ctx = px.Context(0, dict(block256_min_len=2048, block1024_min_len=6000, block_lds_ring=128))
    short = synth.workload("short150", n=2000)
    pb2 = px.PackedBatch(short.q_seq, short.q_off, short.t_seq, short.t_off,
                         np.zeros(short.n, np.uint64), np.zeros(short.n, np.uint64))
    other = px.Plan(pb2, px.ENGINE_LINEAR, ctx=ctx)
    other.run()
    other_ref = other.results()
    for eng in (px.ENGINE_AFFINE, px.ENGINE_LINEAR):
        plan = px.Plan(pb, eng, px.AffineCosts(4, 6, 2), ctx=ctx)
        plan.run()
        ref = plan.results()
        launches = plan.last_launches()
        plan.use_graph(True)
        for rep in range(5):
            if rep == 3:
                plan.set_tuning(dict(heavy_permille=300))
            plan.run()
            say(f"engine {eng} replay {rep}: {plan.last_launches()} launches "
                f"{plan.last_kernel_ms():.3f} ms kernels, {plan.last_ms():.3f} ms run")
            assert plan.last_kernel_ms() > 0
            if rep in (0, 3):  # recorded runs execute eagerly (3: a new grouping)
                launches = plan.last_launches()
            else:
                assert plan.last_launches() >= launches - 1
            r = plan.results()
            assert np.array_equal(r.aln, ref.aln) and np.array_equal(r.runs, ref.runs), f"replay {rep}"
            other.run()
            o = other.results()
            assert np.array_equal(o.aln, other_ref.aln) and np.array_equal(o.runs, other_ref.runs)
        plan.use_graph(False)
        plan.run()
        r = plan.results()
        assert np.array_equal(r.aln, ref.aln) and np.array_equal(r.runs, ref.runs)
        plan.close()
    other.close()
    ctx.close()


def test_lds_ring_workgroup_tier_and_width_overflow(px):
    """Workgroup tiers with the score ring in LDS: pairs routed there by their
    estimate, and (with a deliberately narrow ring) pairs that outgrow it and
    finish on the HBM ring, all bit-exact."""
    rng = random.Random(4242)
    qs, ts = [], []
    for L, rate in ((3000, 0.005), (4500, 0.01), (5200, 0.02), (6000, 0.03), (7000, 0.01)):
        for _ in range(3):
            q = rand_seq(rng, L)
            qs.append(q)
            ts.append(mutate(rng, q, rate) + rand_seq(rng, L // 50))
    for ring in (1024, 128):
        ctx = px.Context(0, dict(block256_min_len=4096, block1024_min_len=12000, block_lds_ring=ring))
        for eng in ("affine", "linear"):
            compare(px, eng, qs, ts, costs=(4, 6, 2), ctx=ctx, label=f"lds-ring-{ring}")
        ctx.close()


@pytest.mark.parametrize("costs", [(4, 6, 2), (2, 4, 1)])
def test_workgroup_register_wavefront(px, costs):
    """Workgroup tiers with the wavefront in registers (16 chunks of 64
    diagonals per wave, chunk ends exchanged through LDS): pairs of 1-8 kbp
    on both workgroup tiers, global and fitting, targets much shorter than
    the query (window placement), byte sequences, 0.5-25% divergence, with
    band pruning to each pair's step estimate; a pair whose optimum is far
    above its estimate (the band empties) and, with every pair forced onto
    the register path, pairs whose range outgrows the 4,096-diagonal window
    of the 256-thread tier stop and are re-run on the ring path.  By default
    the 1024-thread tier's register pairs run on the narrowest register tier
    whose window holds their band (256 or 512 threads).  With the
    band's bound cut to a thousandth of the estimate (band_permille = 1, set
    at context creation and again through set_tuning on a plan) every
    register pair's band empties and the whole set is re-run on the
    register path with its estimate as bound.  Every result equals the
    oracle's, and the launches show the register storage was used (and the
    re-runs)."""
    rng = random.Random(costs[0] * 100 + costs[1])
    qs, ts = [], []
    for L, rate in ((1100, 0.01), (1500, 0.05), (2000, 0.15), (2600, 0.08), (3300, 0.02),
                    (4200, 0.01), (5000, 0.08), (6500, 0.03), (8000, 0.005)):
        for v in range(3):
            q = rand_seq(rng, L)
            t = mutate(rng, q, rate)
            if v == 1:
                t = t + rand_seq(rng, L // 2 + 40)  # fitting
            if v == 2:
                t = t[: len(t) * 3 // 4]  # target shorter than the query
            qs.append(q)
            ts.append(t)
    bq = [q.lower() if i % 4 == 0 else q for i, q in enumerate(qs)]
    # Wide divergence: the range passes 4,096 diagonals, so the register pass
    # of the 256-thread tier overflows when every pair is forced onto it.
    # The first pair has the largest estimate of its length bucket, so it is
    # the one "heavy" pair moved up to the 1024-thread tier (window 16,384).
    wq, wt = [], []
    for L, rate in ((7000, 0.3), (4600, 0.25), (5200, 0.3)):
        q = rand_seq(rng, L)
        wq.append(q)
        wt.append(mutate(rng, q, rate))
    # Band bound below the optimum: the target repeats a block the query has
    # once and the query repeats a block the target lacks, so nearly every
    # query k-mer occurs in the target and the step estimate (the band's
    # bound) is far below the optimum (two gaps of 800 bases); the band
    # empties before the end and the pair is re-run without a bound.
    P, D, Q, R = (rand_seq(rng, n) for n in (1500, 800, 1500, 400))
    uq, ut = [P + D + Q + R + R, qs[20]], [P + D + D + Q, ts[20]]
    regs_pairs = {}
    for tun, pairs, label in (
            (dict(block256_min_len=1024, block1024_min_len=6000), (uq, ut), "underestimated"),
            (dict(block256_min_len=1024, block1024_min_len=6000, band_permille=1000, regs_narrow=2), (qs, ts),
             "tiers"),
            (dict(block256_min_len=1024, block1024_min_len=6000, band_permille=1000), (qs, ts), "narrowest"),
            (dict(block256_min_len=1024, block1024_min_len=6000, band_permille=1), (qs, ts), "relaxed"),
            (dict(block256_min_len=1024, block1024_min_len=6000), (bq, ts), "bytes"),
            (dict(block256_min_len=1024, block1024_min_len=20000, block_regs=1, heavy_permille=1),
             (qs[:6] + wq, ts[:6] + wt), "forced")):
        ctx = px.Context(0, tun)
        pb = px.pack_batch(*pairs)
        plan = px.Plan(pb, px.ENGINE_AFFINE, px.AffineCosts(*costs), ctx=ctx)
        plan.run()
        res = plan.results()
        info = plan.last_launch_info()
        plan.close()
        ctx.close()
        regs = [r for r in info if r["storage"] == 2]
        say(f"{label}: launches {[(r['tier'], r['storage'], r['pairs']) for r in info]}")
        assert regs, "no launch used the register wavefront"
        regs_pairs[label] = sum(r["pairs"] for r in regs)
        if label == "relaxed":
            # every register pair ran twice: its band, then the estimate
            assert regs_pairs["relaxed"] == 2 * regs_pairs["tiers"], info
        if label == "tiers":
            assert {r["tier"] for r in regs} == {2, 3}
        if label == "narrowest":
            # the 1024-thread tier's pairs run on the 256- or 512-thread one
            assert 4 in {r["tier"] for r in regs} and 3 not in {r["tier"] for r in regs}, info
            assert regs_pairs["narrowest"] == regs_pairs["tiers"]
        if label == "underestimated":
            assert any(r["storage"] == 0 and r["tier"] >= 2 for r in info), info
        if label == "forced":
            # the two wide pairs overflow the register window: one ring re-run
            # launch takes both (other ring launches here hold one pair)
            assert max(r["pairs"] for r in info if r["storage"] == 0 and r["tier"] >= 2) >= 2
        compare(px, "affine", *pairs, costs=costs, label=f"block regs {label}", res=res)
    # The bound follows set_tuning on a live plan, and back.
    ctx = px.Context(0, dict(block256_min_len=1024, block1024_min_len=6000))
    plan = px.Plan(px.pack_batch(qs, ts), px.ENGINE_AFFINE, px.AffineCosts(*costs), ctx=ctx)
    for pm, runs in ((1, 2), (1000, 1)):
        plan.set_tuning(dict(band_permille=pm))
        plan.run()
        res = plan.results()
        n_regs = sum(r["pairs"] for r in plan.last_launch_info() if r["storage"] == 2)
        assert n_regs == runs * regs_pairs["tiers"], (pm, plan.last_launch_info())
        compare(px, "affine", qs, ts, costs=costs, label=f"block regs set_tuning {pm}", res=res)
    plan.close()
    ctx.close()


@pytest.mark.parametrize("costs", [(4, 6, 2), (2, 4, 1), (1, 0, 1), (3, 5, 2), (1, 30, 1)])
def test_register_wavefront_phases(px, costs):
    """LDS-tier pairs whose step counts cross the register windows (64 and
    128 diagonals) and the 128-diagonal limit (re-run on an HBM tier), with
    short queries or targets (window placement), fitting targets and byte
    sequences; the last cost set has no register wavefront (LDS ring)."""
    rng = random.Random(sum(costs) * 7 + 1)
    qs, ts = [], []
    for qn in (5, 20, 33, 63, 64, 65, 100, 150, 300, 600, 900):
        for rate in (0.0, 0.05, 0.15, 0.3):
            q = rand_seq(rng, qn)
            t = mutate(rng, q, rate)
            qs.append(q)
            ts.append(t)
            qs.append(q)
            ts.append(t + rand_seq(rng, qn // 2 + 20))  # fitting
            qs.append(q[: max(1, qn // 4)])
            ts.append(t)  # target much longer than the query
            qs.append(q)
            ts.append(t[: max(1, len(t) // 5)])  # target much shorter
    for eng in ("affine", "linear"):
        compare(px, eng, qs, ts, costs=costs, label=f"regs {costs}")
    bq = [q.lower() if i % 3 == 0 else q for i, q in enumerate(qs[:120])]
    compare(px, "affine", bq, ts[:120], costs=costs, label=f"regs bytes {costs}")


def test_run_timings_belong_to_their_plan(px):
    """Each plan times its runs with its own events, read on first query:
    a run of another plan on the same context in between leaves them as
    they were; kernel time never exceeds the run time; a one-launch run and
    a multi-group run both report positive times."""
    from phraya_amd import synth
    ctx = px.Context(0)
    short = synth.workload("short150", n=3000)
    pb1 = px.PackedBatch(short.q_seq, short.q_off, short.t_seq, short.t_off,
                         np.zeros(short.n, np.uint64), np.zeros(short.n, np.uint64))
    w = synth.workload("mixed", n=200)
    idx = [i for i in range(w.n) if w.q_off[i + 1] - w.q_off[i] <= 5000][:80]
    sub = w.subset(idx)
    pb2 = px.PackedBatch(sub.q_seq, sub.q_off, sub.t_seq, sub.t_off,
                         np.zeros(sub.n, np.uint64), np.zeros(sub.n, np.uint64))
    a = px.Plan(pb1, px.ENGINE_AFFINE, px.AffineCosts(4, 6, 2), ctx=ctx)
    b = px.Plan(pb2, px.ENGINE_AFFINE, px.AffineCosts(4, 6, 2), ctx=ctx)
    for _ in range(3):
        a.run()
        b.run()  # runs between a's run and the queries of a's times
        a_ms, a_k, a_n = a.last_ms(), a.last_kernel_ms(), a.last_launches()
        b_ms, b_k, b_n = b.last_ms(), b.last_kernel_ms(), b.last_launches()
        say(f"a: {a_n} launch(es) {a_k:.4f} / {a_ms:.4f} ms; b: {b_n} launches {b_k:.4f} / {b_ms:.4f} ms")
        assert a_n == 1 and b_n >= 2
        assert 0 < a_k <= a_ms * 1.0001 and a_ms < 50
        assert 0 < b_k and 0 < b_ms
        # Queried again: the same numbers (read once per run).
        assert a.last_ms() == a_ms and a.last_kernel_ms() == a_k
    a.close()
    b.close()
    ctx.close()
