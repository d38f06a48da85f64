"""Variant sites from CIGAR runs on the GPU (phx_variant_plan_*, through the
C ABI) against the oracle (orc_extract_variants): every record field and
every coverage value, bit-exact."""
from __future__ import annotations

import random
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPS = {"M": 0, "X": 1, "I": 2, "D": 3}


def say(msg):
    print(f"[gpu-test] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def px():
    import phraya_amd
    return phraya_amd


def compare(px, runs, run_off, target_start, queries, target, spans, repeats, hotspots, radius,
            label=""):
    say(f"variants {label}: {len(target_start)} alignments, radius {radius}")
    got = px.extract_variants(runs, run_off, target_start, queries, target, spans, repeats,
                              hotspots, radius)
    q_seq = np.frombuffer(b"".join(queries), np.uint8)
    q_off = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.uint64)
    flat = np.array([p for sp in spans for p in sp], np.uint64).reshape(-1, 2)
    sp_off = np.concatenate([[0], np.cumsum([len(s) for s in spans])]).astype(np.uint64)
    ov, ooff, ocov = O.extract_variants(
        np.asarray(runs, np.uint32), np.asarray(run_off, np.uint64), np.asarray(target_start, np.uint64),
        q_seq, q_off, target, flat, sp_off, repeats, hotspots, radius)
    assert np.array_equal(got.var_off, ooff), label
    assert len(got.variants) == len(ov), label
    for f in O.VARIANT_DTYPE.names:
        assert np.array_equal(got.variants[f], ov[f]), (label, f)
    assert np.array_equal(got.coverage, ocov), label
    return got


def runs_of(cigar: str):
    out, num = [], ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
        else:
            out.append((int(num) << 4) | OPS[ch])
            num = ""
    return out


def test_known_cases(px):
    items = [("2M1X4M", 0, b"ACTACGT", [(0, 7)], b"ACCACGT"),
             ("3M1I4M", 0, b"ACGACGT", [(0, 8)], b"ACGTACGT"),
             ("4M1D4M", 0, b"ACGTTACGT", [(0, 8)], b"ACGTACGT"),
             ("4M4I", 0, b"ACGT", [(0, 8)], b"ACGTAAAA")]
    for cigar, start, q, spans, target in items:
        r = runs_of(cigar)
        compare(px, r, [0, len(r)], [start], [q], target, [spans], None, None, 50, cigar)
    target = b"TTAACCGGTA" + b"ATATATATATATATATATATAT" + b"CGTACCGATT"
    q = bytearray(target)
    q[2], q[15] = ord("C"), ord("T")
    r = runs_of("2M1X12M1X26M")
    got = compare(px, r, [0, len(r)], [0], [bytes(q)], target, [[(0, 42)]],
                  np.array([[10, 32]], np.uint64), None, 25, "tandem repeat")
    assert {int(x["position"]): int(x["flags"]) for x in got.variants} == {2: 0, 15: 1}


def test_empty_batch_and_alignments_without_runs(px):
    compare(px, np.zeros(0, np.uint32), [0], [], [], b"ACGT", [], None, None, 25, "empty")
    r = runs_of("3M")
    compare(px, r, [0, 0, len(r), len(r)], [0, 1, 2], [b"", b"AAA", b"A"], b"AAAAAA",
            [[], [(0, 3)], [(2, 3)]], None, None, 25, "no runs")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_runs(px, seed):
    """Random CIGARs that overrun queries and targets, starts past the target,
    overlapping / empty / clipped coverage windows, random repeat and
    hotspot intervals, radii 0..150."""
    rng = random.Random(seed)
    tlen = rng.choice([50, 300, 5000])
    target = bytes(rng.choice(b"ACGTN") for _ in range(tlen))
    runs, run_off, starts, queries, spans = [], [0], [], [], []
    for _ in range(400):
        nr = rng.randint(0, 12)
        for _ in range(nr):
            op = rng.choice([0, 0, 1, 1, 2, 3])
            runs.append((rng.randint(1, 40) << 4) | op)
        run_off.append(len(runs))
        starts.append(rng.randint(0, tlen + 20))
        queries.append(bytes(rng.choice(b"ACGT") for _ in range(rng.randint(0, 120))))
        sp = []
        for _ in range(rng.randint(0, 4)):
            a = rng.randint(0, tlen + 30)
            sp.append((a, a + rng.randint(0, 80)))
        spans.append(sp)

    def intervals(k):
        pts = sorted(rng.sample(range(0, tlen + 10), min(2 * k, tlen + 10)))
        return np.array([[pts[i], pts[i + 1]] for i in range(0, len(pts) - 1, 2)], np.uint64)

    rep = intervals(rng.randint(0, 20))
    hot = intervals(rng.randint(0, 20)).astype(np.uint32)
    for radius in (0, 1, 25, 150):
        compare(px, runs, run_off, starts, queries, target, spans, rep, hot, radius,
                f"seed {seed} target {tlen}")


def test_real_alignments(px):
    """CIGARs of the GPU affine engine on 2000 reads placed in a 100 kbp
    target (insertions, deletions, substitutions), each with its alignment
    as coverage window and a second window for some."""
    from phraya_amd import synth
    rng = random.Random(7)
    g = synth.genome(100_000, 0x5EED0003)
    target = bytes(g)
    qs, ts, pos = [], [], []
    for _ in range(2000):
        p = rng.randint(0, len(target) - 400)
        win = target[p:p + 170]
        q = bytearray(win[:150])
        for _ in range(rng.randint(0, 4)):
            i = rng.randrange(len(q))
            kind = rng.random()
            if kind < 0.6:
                q[i] = rng.choice([c for c in b"ACGT" if c != q[i]])
            elif kind < 0.8:
                del q[i]
            else:
                q.insert(i, rng.choice(b"ACGT"))
        qs.append(bytes(q))
        ts.append(win)
        pos.append(p)
    res = px.align_batch(px.ENGINE_AFFINE, qs, ts, None, px.AffineCosts(4, 6, 2))
    runs, run_off, starts, spans = [], [0], [], []
    for i in range(len(qs)):
        a = res.aln[i]
        assert int(a["status"]) == 0
        runs += [int(x) for x in res.runs[int(res.run_off[i]):int(res.run_off[i + 1])]]
        run_off.append(len(runs))
        ts0 = pos[i] + int(a["target_start"])
        starts.append(ts0)
        sp = [(ts0, pos[i] + int(a["target_end"]))]
        if i % 5 == 0:
            sp.append((ts0 + 40, ts0 + 300))
        spans.append(sp)
    rep = np.array([[k, k + 25] for k in range(0, len(target), 700)], np.uint64)
    hot = np.array([[k, k + 60] for k in range(300, len(target), 1100)], np.uint32)
    got = compare(px, runs, run_off, starts, qs, target, spans, rep, hot, 25, "real")
    say(f"{len(got.variants)} variants, {len(got.coverage)} coverage words")
    assert len(got.variants) > 1000


def test_long_alignments_and_several_tiles(px):
    """3000 alignments (three scan tiles of 1024), some with up to 200 runs
    (several 64-run chunks per alignment), long X runs on a 400 bp target
    with radius 150 (windows clipped at both ends), and intervals the lookup
    directory has to get right: duplicate and nested starts, empty and
    touching intervals, starts past the target, one interval over most of
    it."""
    rng = random.Random(11)
    tlen = 400
    target = bytes(rng.choice(b"ACGT") for _ in range(tlen))
    runs, run_off, starts, queries, spans = [], [0], [], [], []
    for k in range(3000):
        nr = rng.choice([0, 3, 8, 70, 200]) if k % 7 == 0 else rng.randint(0, 10)
        for _ in range(nr):
            op = rng.choice([0, 1, 1, 2, 3])
            ln = rng.randint(1, 300) if op == 1 and rng.random() < 0.1 else rng.randint(1, 6)
            runs.append((ln << 4) | op)
        run_off.append(len(runs))
        starts.append(rng.choice([0, 1, rng.randint(0, tlen - 1), tlen - 2, tlen + 5]))
        queries.append(bytes(rng.choice(b"ACGT") for _ in range(rng.randint(0, 400))))
        spans.append([(a, a + rng.randint(0, 200)) for a in (rng.randint(0, tlen) for _ in range(rng.randint(0, 3)))])
    rep = np.array([[5, 5], [5, 40], [5, 12], [30, 31], [31, 60], [100, 390], [395, 396],
                    [399, 400], [400, 420], [500, 600]], np.uint64)
    hot = np.array([[0, 1], [2, 3], [3, 4], [150, 160], [155, 170], [390, 1000]], np.uint32)
    for radius in (0, 150):
        compare(px, runs, run_off, starts, queries, target, spans, rep, hot, radius, "long/tiles")


def _batch(px, runs, run_off, starts, queries, target, spans, rep, hot, radius):
    from phraya_amd import _lib as L
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(x, dtype=dt).reshape(-1)
        keep.append(a)
        return a.ctypes.data if a.size else None

    qs = b"".join(queries)
    b = L.CVariantBatch(len(starts), arr(runs, np.uint32), arr(run_off, np.uint64), arr(starts, np.uint64),
                        arr(np.frombuffer(qs, np.uint8), np.uint8),
                        arr(np.concatenate([[0], np.cumsum([len(q) for q in queries])]), np.uint64),
                        arr(np.frombuffer(target, np.uint8), np.uint8), len(target),
                        arr([p for sp in spans for ab in sp for p in ab], np.uint64),
                        arr(np.concatenate([[0], np.cumsum([len(s) for s in spans])]), np.uint64),
                        arr(rep, np.uint64), len(rep), arr(hot, np.uint32), len(hot), radius)
    return b, keep


def test_one_shot_entry_capacity_and_unsorted_intervals(px):
    """phx_extract_variants through the C ABI: the same records as the plan;
    too small an output reports PHX_E_CAPACITY with the totals set; interval
    starts out of order are refused (PHX_E_ARG), by the one-shot call and by
    plan creation alike."""
    import ctypes
    from phraya_amd import _lib as L
    runs = runs_of("2M1X4M1I3M2D5M")
    target = b"ACCACGTTTACGTACGTAAAA"
    args = (runs, [0, len(runs)], [0], [b"ACTACGTTACGGGTACGTA"], target, [[(0, 20)]],
            np.array([[2, 4], [10, 12]], np.uint64), np.array([[0, 3]], np.uint32), 5)
    want = compare(px, *args, label="one-shot")
    b, keep = _batch(px, *args)
    lib = L.lib()
    h = px.default_context().handle
    v = np.zeros(len(want.variants), L.VARIANT_DTYPE)
    c = np.zeros(len(want.coverage), np.uint32)
    off = np.zeros(2, np.uint64)
    out = L.CVariantResults(v.ctypes.data, len(v), c.ctypes.data, len(c), off.ctypes.data, 0, 0)
    assert lib.phx_extract_variants(h, ctypes.byref(b), ctypes.byref(out)) == L.PHX_OK
    assert (out.n_variants, out.n_coverage) == (len(v), len(c))
    assert np.array_equal(off, want.var_off) and np.array_equal(c, want.coverage)
    for f in L.VARIANT_DTYPE.names:
        assert np.array_equal(v[f], want.variants[f]), f
    small = L.CVariantResults(v.ctypes.data, len(v) - 1, c.ctypes.data, len(c), off.ctypes.data, 0, 0)
    assert lib.phx_extract_variants(h, ctypes.byref(b), ctypes.byref(small)) == L.PHX_E_CAPACITY
    assert (small.n_variants, small.n_coverage) == (len(v), len(c))
    bad = list(args)
    bad[6] = np.array([[10, 12], [2, 4]], np.uint64)
    b2, keep2 = _batch(px, *bad)
    assert lib.phx_extract_variants(h, ctypes.byref(b2), ctypes.byref(out)) == L.PHX_E_ARG
    assert "not sorted" in px.default_context().last_error()
    with pytest.raises(RuntimeError, match="not sorted"):
        px.extract_variants(*bad)
