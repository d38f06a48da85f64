"""Contig-mode anchoring (phx_contig_anchors, host only): properties checked
against brute force on small pairs."""
from __future__ import annotations

import random

import pytest

import phraya_amd as px

K, W = 15, 10
CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}


def mix32(x: int) -> int:
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def kmer_code(s: bytes, i: int, k: int = K):
    c = 0
    for b in s[i:i + k]:
        if b not in CODE:
            return None
        c = (c << 2) | CODE[b]
    return c


def brute_minimizers(s: bytes, k: int = K, w: int = W) -> dict[int, list[int]]:
    """code -> positions of (k, w) minimizers: every window of w consecutive
    valid k-mers picks its leftmost smallest mix32 hash."""
    codes = [kmer_code(s, i, k) for i in range(len(s) - k + 1)]
    picked = set()
    for a in range(len(codes) - w + 1):
        win = codes[a:a + w]
        if any(c is None for c in win):
            continue
        best = min(range(w), key=lambda j: (mix32(win[j]), j))
        picked.add(a + best)
    out: dict[int, list[int]] = {}
    for p in sorted(picked):
        out.setdefault(codes[p], []).append(p)
    return out


def make_pair(seed: int, n: int = 24000, rate: float = 0.01, with_n: bool = False):
    rng = random.Random(seed)
    t = bytearray(rng.choice(b"ACGT") for _ in range(n))
    q = bytearray()
    i = 0
    while i < len(t):
        r = rng.random()
        if r < rate * 0.1:
            if rng.random() < 0.5:
                q += bytes(rng.choice(b"ACGT") for _ in range(rng.randint(1, 50)))
            else:
                i += rng.randint(1, 50)
            continue
        q.append(rng.choice([c for c in b"ACGT" if c != t[i]]) if r < rate else t[i])
        i += 1
    if with_n:
        for _ in range(30):
            q[rng.randrange(len(q))] = ord("N")
    return bytes(q), bytes(t)


@pytest.mark.parametrize("seed,with_n", [(1, False), (2, True), (3, False)])
def test_anchor_properties(seed, with_n):
    q, t = make_pair(seed, with_n=with_n)
    params = px.ContigParams(k=K, w=W, gap_min=800, gap_max=4000, seed=99)
    anchors = px.contig_anchors(q, t, params, pair_index=seed)
    assert len(anchors) >= 3
    mq, mt = brute_minimizers(q), brute_minimizers(t)
    prev_q_end, prev_t_end = 0, 0
    for a, b in anchors:
        assert q[a:a + K] == t[b:b + K], "anchor is not an exact k-mer match"
        code = kmer_code(q, a)
        assert mq.get(code) == [a] and mt.get(code) == [b], "anchor is not a unique shared minimizer"
        assert a >= prev_q_end + params.gap_min and b >= prev_t_end, "spacing / order"
        prev_q_end, prev_t_end = a + K, b + K


def test_anchors_deterministic_and_seeded():
    q, t = make_pair(4)
    p = px.ContigParams(gap_min=500, gap_max=5000, seed=11)
    a1 = px.contig_anchors(q, t, p, 0)
    assert a1 == px.contig_anchors(q, t, p, 0)
    assert a1 != px.contig_anchors(q, t, p, 1)  # spacing draws depend on the pair index
    assert a1 != px.contig_anchors(q, t, px.ContigParams(gap_min=500, gap_max=5000, seed=12), 0)


def test_no_anchors_without_shared_kmers():
    rng = random.Random(5)
    q = bytes(rng.choice(b"ACGT") for _ in range(5000))
    t = bytes(rng.choice(b"ACGT") for _ in range(5000))
    assert px.contig_anchors(q, t, px.ContigParams(gap_min=100, gap_max=200)) == []
    assert px.contig_anchors(b"N" * 3000, b"N" * 3000) == []
    with pytest.raises(ValueError):
        px.contig_anchors(q, t, px.ContigParams(k=20))
