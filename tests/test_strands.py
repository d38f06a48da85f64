"""align_read's strand handling, host side (no GPU): reverse_complement
(phraya-core/src/types.rs:91-107) against upstream's own cases, and which
orientations phx_align_reads_bounds schedules for extension
(executor.rs:632-698)."""
from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

import phraya_amd as px
from phraya_amd import _lib as L


@pytest.mark.parametrize("seq,want", [
    (b"ACGT", b"ACGT"), (b"AAAA", b"TTTT"), (b"GATTACA", b"TGTAATC"),  # types.rs:755-760
    (b"N", b"N"), (b"acgt", b"acgt"), (b"ANCXT", b"AXGNT"),            # types.rs:768-774
    (b"", b"")])
def test_reverse_complement_known_answers(seq, want):
    assert px.reverse_complement(seq) == want


def test_reverse_complement_is_an_involution_and_reverses():
    # types.rs:762-766, and on every byte value: complement of the reversed input.
    s = b"GATTACAGGCCTTAACGNNTA"
    assert px.reverse_complement(px.reverse_complement(s)) == s
    allb = bytes(range(256))
    table = {ord("A"): ord("T"), ord("T"): ord("A"), ord("C"): ord("G"), ord("G"): ord("C"),
             ord("a"): ord("t"), ord("t"): ord("a"), ord("c"): ord("g"), ord("g"): ord("c")}
    assert px.reverse_complement(allb) == bytes(table.get(b, b) for b in reversed(allb))
    rng = random.Random(3)
    for _ in range(50):
        r = bytes(rng.choice(b"ACGTacgtN") for _ in range(rng.randint(0, 300)))
        assert px.reverse_complement(px.reverse_complement(r)) == r


def bounds(reads, is_dna=True, strategy=2, target_len=1000):
    keep, arr = [], (L.CReadStrands * max(len(reads), 1))()
    for i, (qlen, nf, nr) in enumerate(reads):
        q = np.zeros(max(qlen, 1), np.uint8)
        f = np.arange(max(nf, 1), dtype=np.uint64)
        r = np.arange(max(nr, 1), dtype=np.uint64)
        keep += [q, f, r]
        arr[i] = L.CReadStrands(q.ctypes.data, qlen, f.ctypes.data if nf else None, nf,
                                r.ctypes.data if nr else None, nr, 0)
    e, a, n = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = L.lib().phx_align_reads_bounds(strategy, int(is_dna), target_len, arr, len(reads),
                                        ctypes.byref(e), ctypes.byref(a), ctypes.byref(n))
    assert rc == L.PHX_OK
    return e.value, a.value


def test_orientations_extended():
    # executor.rs:676-698: forward only when only it has chains, reverse only
    # when only the reverse has, both otherwise (chains on both, or none).
    assert bounds([(100, 3, 0)]) == (1, 3)
    assert bounds([(100, 0, 2)]) == (1, 2)
    assert bounds([(100, 2, 5)]) == (2, 7)
    assert bounds([(100, 0, 0)]) == (2, 2)          # two fallback (0, 0) anchors
    # Strategy caps the anchors per orientation (Fast 1, Balanced 2).
    assert bounds([(100, 2, 5)], strategy=0) == (2, 2)
    assert bounds([(100, 3, 5)], strategy=1) == (2, 4)
    # Protein (:632-646): reverse chains ignored, so a read with forward
    # chains is extended forward only, and one without as both orientations
    # (the reverse with an empty query).
    assert bounds([(100, 3, 4)], is_dna=False) == (1, 3)
    assert bounds([(100, 0, 4)], is_dna=False) == (2, 2)
    assert bounds([(10, 1, 0), (20, 0, 1), (30, 1, 1)]) == (4, 4)


def test_bounds_reject_malformed_reads():
    # A null query with a length, or null chains with a count, is PHX_E_ARG.
    for qlen, nf, nr, null_q, null_f, null_r in [(5, 0, 0, True, False, False),
                                                 (5, 2, 0, False, True, False),
                                                 (5, 0, 2, False, False, True)]:
        q = np.zeros(5, np.uint8)
        f = np.zeros(2, np.uint64)
        arr = (L.CReadStrands * 1)()
        arr[0] = L.CReadStrands(None if null_q else q.ctypes.data, qlen,
                                None if null_f else f.ctypes.data, nf,
                                None if null_r else f.ctypes.data, nr, 0)
        e, a, n = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        rc = L.lib().phx_align_reads_bounds(2, 1, 100, arr, 1, ctypes.byref(e), ctypes.byref(a),
                                            ctypes.byref(n))
        assert rc == L.PHX_E_ARG
    # Null reverse chains are fine for protein: they are never read.
    arr = (L.CReadStrands * 1)()
    q = np.zeros(5, np.uint8)
    arr[0] = L.CReadStrands(q.ctypes.data, 5, None, 0, None, 3, 0)
    e, a, n = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    assert L.lib().phx_align_reads_bounds(2, 0, 100, arr, 1, ctypes.byref(e), ctypes.byref(a),
                                          ctypes.byref(n)) == L.PHX_OK
    assert (e.value, a.value) == (2, 2)


def test_bench_sketcher_selects_window_minima():
    # synth.minimizer_sketch (the synthetic workloads' sketcher): every
    # window of w k-mers has its leftmost smallest hash selected, and every
    # selected position is that minimum of some window, in order, once.
    from phraya_amd import synth
    rng = random.Random(8)
    for k, w in [(5, 3), (21, 11), (7, 1)]:
        seq = bytes(rng.choice(b"ACGT") for _ in range(300))
        sk = synth.minimizer_sketch(seq, k, w)
        n = len(seq) - k + 1
        code = {65: 0, 67: 1, 84: 2, 71: 3}
        mask = (1 << 64) - 1

        def h(i):
            z = 0
            for c in seq[i:i + k]:
                z = (z << 2) | code[c]
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
            return z ^ (z >> 31)
        hs = [h(i) for i in range(n)]
        want = []
        for s in range(n - w + 1):
            m = min(range(s, s + w), key=lambda i: (hs[i], i))
            if not want or want[-1] != m:
                want.append(m)
        assert [int(p) for p in sk[:, 1]] == want
        assert [int(v) for v in sk[:, 0]] == [hs[p] for p in want]
