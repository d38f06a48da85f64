"""Request server (phx_serve_*): every served result equals the oracle's and the
same pair aligned alone, whatever else is in flight."""
from __future__ import annotations

import random
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-serve] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def px():
    import phraya_amd
    return phraya_amd


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


def collect(srv, n, limit_s=120.0):
    got = {}
    t0 = time.time()
    while len(got) < n:
        for r in srv.poll(4096, timeout_us=20000):
            assert r.tag not in got, f"tag {r.tag} delivered twice"
            got[r.tag] = r
        assert time.time() - t0 < limit_s, f"only {len(got)} of {n} responses"
    return got


def oracle_check(reqs, got, costs, engine=O.AFFINE):
    """reqs: list of (tag, query, target, seed_q, seed_t)."""
    bad = []
    for tag, q, t, sq, st in reqs:
        r = got[tag]
        if engine == O.AFFINE:
            o = O.wfa_extend_affine(q, t, (sq, st), costs)
        else:
            o = O.wfa_extend(q, t, (sq, st))
        if o.status != O.OK:  # InvalidInput (1) / AlignmentFailed (2)
            if r.status != o.status:
                bad.append((tag, "status", r.status, o.status))
            continue
        if r.status != 0:
            bad.append((tag, "status", r.status, 0))
            continue
        for f in ("edit_distance", "query_start", "query_end", "target_start", "target_end"):
            if getattr(r, f) != getattr(o, f):
                bad.append((tag, f, getattr(r, f), getattr(o, f)))
        if r.cigar != o.cigar:
            bad.append((tag, "cigar", r.cigar[:60], o.cigar[:60]))
    assert not bad, f"{len(bad)} mismatches vs oracle, first: {bad[:4]}"


def make_genome(rng, n=400_000, n_count=40):
    g = bytearray(rand_seq(rng, n))
    for _ in range(n_count):  # a few N bases: windows over them take the byte path
        g[rng.randrange(n)] = ord("N")
    return bytes(g)


def test_serve_mixed_stream_matches_oracle_and_alone(px):
    rng = random.Random(2024)
    genome = make_genome(rng)
    costs = (4, 6, 2)
    reqs = []  # (tag, query, target, seed_q, seed_t, ref_start or None)
    tag = 0
    for kind, count in (("short", 160), ("mid", 24), ("long", 6), ("divergent", 10), ("bytes", 8)):
        for _ in range(count):
            L = {"short": 150, "mid": rng.randint(900, 3000), "long": rng.randint(4000, 9000),
                 "divergent": rng.randint(600, 950), "bytes": rng.randint(100, 400)}[kind]
            W = L + max(20, L // 10)
            pos = rng.randrange(len(genome) - W - 1)
            window = genome[pos:pos + W]
            rate = {"short": 0.01, "mid": 0.05, "long": 0.03, "divergent": 0.18, "bytes": 0.02}[kind]
            q = mutate(rng, window[:L].replace(b"N", b"A"), rate)
            if kind == "bytes":
                q = q[:10] + b"n" + q[11:]
            reqs.append((tag, q, window, 0, 0, pos))
            tag += 1
    # A few explicit targets and bad seeds.
    for _ in range(6):
        q = rand_seq(rng, 200)
        t = mutate(rng, q, 0.04) + rand_seq(rng, 30)
        reqs.append((tag, q, t, 3, 2, None))
        tag += 1
    reqs.append((tag, b"ACGTACGT", b"ACGTACGT", 9, 0, None))  # seed past the query
    tag += 1
    order = list(range(len(reqs)))
    rng.shuffle(order)
    say(f"submitting {len(reqs)} requests in bursts")
    with px.Server(px.ENGINE_AFFINE, px.AffineCosts(*costs), reference=genome,
                   hbm_budget_mib=4096) as srv:
        for chunk in range(0, len(order), 37):
            for i in order[chunk:chunk + 37]:
                tg, q, t, sq, st, pos = reqs[i]
                if pos is None:
                    srv.submit(tg, q, target=t, seed=px.SeedAnchor(sq, st))
                else:
                    srv.submit(tg, q, ref_start=pos, ref_len=len(t), seed=px.SeedAnchor(sq, st))
            time.sleep(0.002)
        got = collect(srv, len(reqs))
        st = srv.stats()
    say(f"stats {st}")
    assert st["completed"] == len(reqs) and st["arena_used_bytes"] == 0
    oracle_check([(r[0], r[1], r[2], r[3], r[4]) for r in reqs], got, costs)
    # The same pair alone, through the per-pair entry point.
    for i in rng.sample(range(len(reqs)), 12):
        tg, q, t, sq, st_, _ = reqs[i]
        try:
            alone = px.wfa_extend_affine(q, t, px.SeedAnchor(sq, st_), px.AffineCosts(*costs))
        except px.InvalidInput:
            assert got[tg].status == 1
            continue
        assert got[tg].alignment() == alone


def test_serve_small_budget_waits_and_rejects(px):
    """A budget that holds only a few long requests at a time: the rest wait
    for memory and still finish bit-exact; one too big for the whole budget
    reports PHX_E_NOMEM without disturbing the others."""
    rng = random.Random(77)
    genome = make_genome(rng, 200_000, 0)
    costs = (4, 6, 2)
    reqs = []
    for tag in range(12):
        L = rng.randint(2500, 5000)
        pos = rng.randrange(len(genome) - 2 * L)
        window = genome[pos:pos + L + L // 10]
        reqs.append((tag, mutate(rng, window[:L], 0.05), window, 0, 0, pos))
    big_q = rand_seq(rng, 60_000)
    with px.Server(px.ENGINE_AFFINE, px.AffineCosts(*costs), reference=genome,
                   hbm_budget_mib=96) as srv:
        for tg, q, t, _, _, pos in reqs:
            srv.submit(tg, q, ref_start=pos, ref_len=len(t))
        srv.submit(99, big_q, target=mutate(rng, big_q, 0.08))
        got = collect(srv, len(reqs) + 1)
        st = srv.stats()
    say(f"stats {st}")
    assert got[99].status == 34  # PHX_E_NOMEM
    assert st["arena_peak_bytes"] <= 96 << 20
    oracle_check([(r[0], r[1], r[2], r[3], r[4]) for r in reqs], got, costs)


def test_serve_linear_engine(px):
    rng = random.Random(5)
    genome = make_genome(rng, 100_000, 5)
    reqs = []
    for tag in range(80):
        L = rng.choice([150, 300, 1200, 2600])
        pos = rng.randrange(len(genome) - L - 40)
        window = genome[pos:pos + L + 20]
        reqs.append((tag, mutate(rng, window[:L].replace(b"N", b"C"), 0.03), window, 0, 0, pos))
    with px.Server(px.ENGINE_LINEAR, reference=genome, hbm_budget_mib=1024) as srv:
        for tg, q, t, _, _, pos in reqs:
            srv.submit(tg, q, ref_start=pos, ref_len=len(t))
        got = collect(srv, len(reqs))
    oracle_check([(r[0], r[1], r[2], r[3], r[4]) for r in reqs], got, None, engine=O.LINEAR)


def test_serve_rejects_window_outside_reference(px):
    genome = b"ACGT" * 1000
    with px.Server(px.ENGINE_AFFINE, px.AffineCosts(4, 6, 2), reference=genome,
                   hbm_budget_mib=64) as srv:
        with pytest.raises(ValueError):
            srv.submit(1, b"ACGT", ref_start=3990, ref_len=20)
        srv.submit(2, b"ACGTACGT", ref_start=0, ref_len=8)
        got = collect(srv, 1)
    assert got[2].cigar == "8M" and got[2].status == 0
