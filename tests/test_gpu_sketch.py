"""Minimizer sketches on the GPU (phx_sketch_plan_*, phx_sketch_batch through
the C ABI) against the oracle (orc_sketch), bit-exact: every value and
position of every sequence, canonical and forward, across task boundaries
(253 - w windows per task), at the 5 Mbp genome's full size, plus strand
symmetry on the device's own output."""
from __future__ import annotations

import ctypes
import random
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu



def task_windows(w):
    return 253 - w


def say(msg):
    print(f"[gpu-sketch] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def px():
    import phraya_amd
    return phraya_amd


def rc(b: bytes) -> bytes:
    return b[::-1].translate(bytes.maketrans(b"ACGTacgt", b"TGCAtgca"))


def check_batch(px, seqs, k, w, canonical, label):
    say(f"{label}: {len(seqs)} sequences, {sum(map(len, seqs))} bases, k {k} w {w} canonical {canonical}")
    got = px.sketch_batch(seqs, k, w, canonical)
    assert len(got) == len(seqs)
    for i, s in enumerate(seqs):
        v, p = O.sketch(s, k, w, canonical)
        assert np.array_equal(got[i][:, 0], v), (label, i, len(s))
        assert np.array_equal(got[i][:, 1], p.astype(np.uint64)), (label, i, len(s))
    return got


def lengths_around_tasks(k, w):
    l = k + w - 1
    out = [0, 1, l - 1, l, l + 1, 150, 300]
    tw = task_windows(w)
    for j in (1, 2, 3, 7):
        out += [l - 2 + tw * j, l - 1 + tw * j, l + tw * j]
    return out


@pytest.mark.parametrize("k,w,canonical", [(21, 11, True), (21, 5, True), (15, 9, True), (32, 32, True),
                                           (1, 1, True), (21, 11, False), (6, 5, False), (15, 10, False),
                                           (3, 128, False), (32, 1, False)])
def test_batch_matches_oracle(px, k, w, canonical):
    rng = random.Random(k * 1000 + w * 10 + canonical)
    seqs = []
    for n in lengths_around_tasks(k, w):
        seqs.append(bytes(rng.choice(b"ACGT") for _ in range(n)))
    # tandem repeats (equal k-mers inside a window: the tie rules), lower
    # case, N and other bytes, protein letters
    unit = b"ACGTTGCA"[: rng.randint(1, 4)]
    seqs.append((unit * 400)[:1500])
    seqs.append(bytes(rng.choice(b"ACGTacgtN") for _ in range(700)))
    seqs.append(bytes(rng.choice(b"ACDEFGHIKLMNPQRSTVWY") for _ in range(900)))
    seqs.append(b"A" * 600)
    rng.shuffle(seqs)
    check_batch(px, seqs, k, w, canonical, f"mixed k{k} w{w}")


def test_reads_batch_and_genome_full_size(px):
    from phraya_amd import synth
    g = synth.genome(5_000_000)
    wl = synth.workload("short150")
    reads = [wl.q_seq[wl.q_off[i]:wl.q_off[i + 1]].tobytes() for i in range(wl.n)]
    check_batch(px, reads, 21, 11, True, "short150 reads")
    got = check_batch(px, [g], 21, 11, True, "5 Mbp genome")[0]
    # strand symmetry on the device's own output at full size
    say("5 Mbp genome reverse complement")
    gr = px.sketch_batch([rc(g.tobytes())], 21, 11, True)[0]
    assert np.array_equal(gr[::-1, 0], got[:, 0])
    assert np.array_equal((len(g) - 21 - gr[::-1, 1].astype(np.int64)), got[:, 1].astype(np.int64))


def test_plan_reruns_capacity_and_empty(px):
    import phraya_amd._lib as L
    rng = random.Random(5)
    seqs = [bytes(rng.choice(b"ACGT") for _ in range(n)) for n in (400, 10, 1000, 0)]
    say("plan rerun / capacity / empty batch")
    p = px.SketchPlan(seqs)
    try:
        p.run()
        a = p.results()
        p.run()
        b = p.results()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert p.total() == len(a[0]) <= p.capacity
        off = np.zeros(len(seqs) + 1, np.uint64)
        tot = ctypes.c_uint64(0)
        small = np.zeros(1, np.uint64)
        small_p = np.zeros(1, np.uint32)
        rc_ = L.lib().phx_sketch_plan_results(p._h, small.ctypes.data, small_p.ctypes.data, 1, off.ctypes.data,
                                              ctypes.byref(tot))
        assert rc_ == L.PHX_E_CAPACITY and tot.value == len(a[0])
        assert np.array_equal(off, a[2])
    finally:
        p.close()
    assert px.sketch_batch([]) == []
    assert all(len(x) == 0 for x in px.sketch_batch([b"", b"ACGT", b"A" * 30]))
    for k, w, c in [(0, 11, True), (33, 1, False), (21, 0, False), (21, 129, False), (21, 10, True)]:
        with pytest.raises(ValueError):
            px.SketchPlan([b"ACGT" * 40], k, w, c)


def test_single_sequence_helpers(px):
    s = bytes(random.Random(9).choice(b"ACGT") for _ in range(500))
    v, p = O.sketch(s, 21, 11, True)
    got = px.sketch(s)
    assert np.array_equal(got[:, 0], v) and np.array_equal(got[:, 1], p.astype(np.uint64))
    v, p = O.sketch(s, 6, 5, False)
    got = px.sketch_alphabet(s, 6, 5, "protein")
    assert np.array_equal(got[:, 0], v) and np.array_equal(got[:, 1], p.astype(np.uint64))
