"""The seeding / chaining oracle (orc_seed_and_chain, orc_seed_occurrence_cap)
against upstream's own tests (phraya-align/src/chaining.rs:254-346,
phraya-align/src/seeding.rs:99-171), the score's closed form on hand-worked
gaps, and a separately written Python restatement on random seed sets.

Upstream's chaining tests hand chain_seeds a seed list; here each seed is
made from a minimizer value of its own (target sketch entry (value,
target_pos), query sketch entry (value, query_pos)), so seeding yields
exactly that list."""
from __future__ import annotations

import math
import random

import numpy as np
import pytest

from oracle import oracle as O

P = O.chain_params


def chains_of(seeds, params=None):
    """seeds: [(query_pos, target_pos)] -> list of (score, target_start,
    [(q, t), ...])."""
    tmins = [(0x1000 + i, t) for i, (q, t) in enumerate(seeds)]
    qmins = [(0x1000 + i, q) for i, (q, t) in enumerate(seeds)]
    ns, ch, off, sd = O.seed_and_chain(tmins, [qmins], params=params or P())
    assert int(ns[0]) == len(seeds)
    out = []
    for c in ch[int(off[0]):int(off[1])]:
        s = sd[int(c["seed_off"]):int(c["seed_off"]) + int(c["n_seeds"])]
        out.append((int(c["score"]), int(c["target_start"]),
                    [(int(x["query_pos"]), int(x["target_pos"])) for x in s]))
    return out


def diagonal(target_start, query_start, count, spacing):
    return [(query_start + i * spacing, target_start + query_start + i * spacing) for i in range(count)]


def test_empty_and_single_seed():
    # chaining.rs:280-300
    assert chains_of([]) == []
    assert chains_of([(0, 1000)], P(min_chain_score=25, seed_len=21)) == []
    c = chains_of([(0, 1000)], P(min_chain_score=10, seed_len=21))
    assert len(c) == 1 and len(c[0][2]) == 1 and c[0][1] == 1000 and c[0][0] == 21


def test_clean_diagonal_collapses():
    # chaining.rs:302-311: 5 seeds 40 apart, no drift: one chain of all 5,
    # score 21 + 4 x min(21, 40) = 105.
    c = chains_of(diagonal(1000, 0, 5, 40))
    assert len(c) == 1 and len(c[0][2]) == 5 and c[0][1] == 1000 and c[0][0] == 105


def test_short_indel_still_chains():
    # chaining.rs:313-330: a 5 bp deletion between seeds 80 and 160.  The
    # link 80 -> 160 has dq = 80, dr = 85: beyond max_dist_x + seed_len = 121?
    # No: 85 <= 121, dd = 5, score min(21, 80) - round(0.21 * 5 + 0.5 log2 6)
    # = 21 - round(1.05 + 1.2925) = 21 - 2 = 19.
    s = diagonal(1000, 0, 3, 40) + diagonal(1005, 160, 3, 40)
    c = chains_of(s)
    assert len(c) == 1 and len(c[0][2]) == 6
    assert c[0][0] == 21 + 21 + 21 + 19 + 21 + 21


def test_repeat_copies_do_not_merge():
    # chaining.rs:332-346 (first part)
    c = chains_of(diagonal(1000, 0, 4, 40) + diagonal(6000, 0, 4, 40))
    assert len(c) == 2 and sorted(x[1] for x in c) == [1000, 6000]


def test_longer_chain_wins_and_band_limits():
    c = chains_of(diagonal(1000, 0, 6, 40) + [(500, 9000)])
    assert len(c[0][2]) == 6 and c[0][1] == 1000
    # out_of_band_drift_prevents_chaining: drift 200 > max_band 100.
    c = chains_of([(0, 1000), (40, 1240)], P(min_chain_score=10))
    assert len(c) == 2


def test_score_closed_form():
    # chain_score (chaining.rs:96-117) on hand-worked gaps: one predecessor,
    # so f = 21 + score(link); the chain keeps both seeds when the link
    # scores above -21 + ... and the reported score is f of the end seed.
    for dq, dr in [(10, 11), (10, 13), (30, 40), (50, 20), (5, 5), (100, 110)]:
        c = chains_of([(0, 500), (dq, 500 + dr)], P(min_chain_score=0))
        dd = abs(dr - dq)
        link = min(21, min(dq, dr)) - (round_half_away(0.21 * dd + 0.5 * math.log2(dd + 1)) if dd else 0)
        best = max(21, 21 + link)
        assert max(x[0] for x in c) == best, (dq, dr)


def round_half_away(x):
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def test_seeding_mask_and_order():
    # seeding.rs:107-125: value 7 at 5 target positions, value 9 at one.
    tmins = [(7, 10), (7, 20), (7, 30), (7, 40), (7, 50), (9, 100)]
    q = [(7, 0), (9, 5)]
    params = P(min_chain_score=10 ** 9)  # count seeds only
    ns, _, _, _ = O.seed_and_chain(tmins, [q], params=params)
    assert int(ns[0]) == 6
    ns, _, _, _ = O.seed_and_chain(tmins, [q], max_occ=4, params=params)
    assert int(ns[0]) == 1


def test_occurrence_cap():
    # seeding.rs:127-171
    assert O.seed_occurrence_cap([(v, v) for v in range(1000)], 256) == 256
    t = [(v, v) for v in range(999)] + [(9999, 0)] * 10_000
    assert O.seed_occurrence_cap(t, 256) == 256
    t = [(v, 0) for v in range(500) for _ in range(100)]
    assert O.seed_occurrence_cap(t, 256) == 800
    assert O.seed_occurrence_cap([], 256) == 256


def py_chain(seeds, p):
    """chain_seeds restated in Python from chaining.rs:131-253, for random
    cross-checks of the C++ restatement (seeds: (q, t) in find_seeds order)."""
    s = sorted(seeds, key=lambda x: (x[1], x[0]))  # Python's sort is stable too
    n = len(s)
    f, pred, dom = [0] * n, [None] * n, [0] * n
    for i in range(n):
        best, bp, skip = p["seed_len"], None, 0
        for j in range(i - 1, max(0, i - p["max_iter"]) - 1, -1):
            if s[i][1] - s[j][1] >= p["max_dist_x"] + p["seed_len"]:
                break
            dq, dr = s[i][0] - s[j][0], s[i][1] - s[j][1]
            if dq <= 0 or dr <= 0 or dr > p["max_dist_x"] + p["seed_len"]:
                continue
            dd = abs(dr - dq)
            if dd > p["max_band"]:
                continue
            sc = min(p["seed_len"], min(dr, dq))
            if dd > 0:
                sc -= round_half_away(p["gap_penalty"] * dd + 0.5 * math.log2(dd + 1.0))
            tot = sc + f[j]
            if tot > best:
                best, bp = tot, j
                skip = max(0, skip - 1)
            else:
                dom[j] += 1
                if dom[j] > p["max_skip"]:
                    skip += 1
                    if skip > p["max_skip"]:
                        break
        f[i], pred[i] = best, bp
    order = sorted(range(n), key=lambda i: -f[i])
    used = [False] * n
    chains = []
    for e in order:
        if used[e] or f[e] < p["min_chain_score"]:
            continue
        rev, cur = [], e
        while cur is not None and not used[cur]:
            used[cur] = True
            rev.append(s[cur])
            cur = pred[cur]
        if not rev:
            continue
        rev.reverse()
        rev.sort(key=lambda x: x[0])
        chains.append((f[e], max(0, rev[0][1] - rev[0][0]), rev))
    chains.sort(key=lambda c: -c[0])
    return chains


@pytest.mark.parametrize("seed", range(6))
def test_random_seed_sets_match_a_second_restatement(seed):
    rng = random.Random(seed)
    p = dict(seed_len=rng.choice([11, 21]), max_band=rng.choice([5, 100]), max_dist_x=rng.choice([30, 100]),
             max_iter=rng.choice([3, 5000]), max_skip=rng.choice([1, 25]), gap_penalty=0.21,
             min_chain_score=rng.choice([0, 25]))
    for _ in range(40):
        # Several loci (diagonals with drift and indels), duplicated target
        # positions, and noise; query positions may repeat (two values at one
        # position of the read).
        seeds = []
        for _ in range(rng.randint(0, 4)):
            t0, q = rng.randrange(0, 5000), 0
            for _ in range(rng.randint(1, 12)):
                q += rng.randint(0, 30)
                seeds.append((q, max(0, t0 + q + rng.randint(-8, 8))))
        seeds += [(rng.randrange(0, 300), rng.randrange(0, 5000)) for _ in range(rng.randint(0, 10))]
        rng.shuffle(seeds)
        # find_seeds order: stable sort by query position of the sketch order.
        order = sorted(range(len(seeds)), key=lambda i: seeds[i][0])
        want = py_chain([seeds[i] for i in order], p)
        got = chains_of(seeds, P(**p))
        assert got == [(a, b, c) for a, b, c in want]
