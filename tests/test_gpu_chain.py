"""Seeding and chaining on the GPU (phx_chain_plan_*, phx_seed_and_chain_batch
through the C ABI) against the oracle (orc_seed_and_chain): seed counts,
every chain's score, target start and seeds, bit-exact."""
from __future__ import annotations

import ctypes
import random
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-chain] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def px():
    import phraya_amd
    return phraya_amd


def compare(px, target_mins, queries, max_occ=2 ** 64 - 1, label="", **kw):
    say(f"{label}: {len(target_mins)} target minimizers, {len(queries)} queries, max_occ {max_occ}")
    got = px.seed_and_chain(target_mins, queries, max_occ, px.default_chain_params(**kw))
    ns, ch, off, sd = O.seed_and_chain(target_mins, queries, max_occ, O.chain_params(**kw))
    assert np.array_equal(got.n_seeds, ns), label
    assert np.array_equal(got.chain_off, off), label
    for f in ("score", "target_start", "seed_off", "n_seeds"):
        assert np.array_equal(got.chains[f], ch[f]), (label, f)
    for f in O.SEED_DTYPE.names:
        assert np.array_equal(got.seeds[f], sd[f]), (label, f)
    return got


def diagonal(target_start, query_start, count, spacing, v0):
    t = [(v0 + i, target_start + query_start + i * spacing) for i in range(count)]
    q = [(v0 + i, query_start + i * spacing) for i in range(count)]
    return t, q


def test_upstream_cases(px):
    # chaining.rs:280-346 and seeding.rs:107-125 as sketches (one value per seed).
    t1, q1 = diagonal(1000, 0, 5, 40, 0x100)            # clean diagonal
    t2a, q2a = diagonal(1000, 0, 3, 40, 0x200)          # short indel
    t2b, q2b = diagonal(1005, 160, 3, 40, 0x300)
    t3a, q3a = diagonal(1000, 0, 4, 40, 0x400)          # repeat copies
    t3b, q3b = diagonal(6000, 0, 4, 40, 0x500)
    t4, q4 = diagonal(1000, 0, 6, 40, 0x600)            # long chain + noise
    tn, qn = [(0x700, 9000)], [(0x700, 500)]
    tb, qb = [(0x800, 1000), (0x801, 1240)], [(0x800, 0), (0x801, 40)]   # out of band
    target = t1 + t2a + t2b + t3a + t3b + t4 + tn + tb + [(7, 10), (7, 20), (7, 30), (7, 40), (7, 50), (9, 100)]
    queries = [q1, q2a + q2b, q3a + q3b, q4 + qn, qb, [], [(7, 0), (9, 5)], [(0xdead, 3)]]
    compare(px, target, queries, label="upstream")
    compare(px, target, queries, max_occ=4, label="upstream masked", min_chain_score=10)
    got = compare(px, target, queries, max_occ=0, label="everything masked", min_chain_score=0)
    assert int(got.n_seeds.sum()) == 0 and len(got.chains) == 0


def test_empty_index_and_queries(px):
    compare(px, [], [[(1, 2)], []], label="empty index")
    compare(px, [(1, 5)], [], label="no queries")


def random_case(rng, nq, genome_len=50_000, k=21):
    """Target sketch of a random genome-like layout (a value per position,
    some values repeated across copies), and query sketches sampled from
    diagonals with drift, indels, wrong-orientation noise and repeats."""
    tpos = sorted(rng.sample(range(genome_len), genome_len // 12))
    values = {}
    target = []
    for p in tpos:
        v = rng.randrange(1, 1 << 40) if rng.random() > 0.05 else rng.choice([11, 12, 13, 14])  # repeats
        target.append((v, p))
        values.setdefault(v, []).append(p)
    queries = []
    by_pos = {p: v for v, p in target}
    for _ in range(nq):
        start = rng.randrange(0, genome_len - 400)
        L = rng.choice([100, 150, 300])
        q = []
        shift = 0
        for p in tpos:
            if start <= p < start + L:
                if rng.random() < 0.02:
                    shift += rng.randint(-6, 6)
                if rng.random() < 0.85:
                    q.append((by_pos[p], max(0, p - start + shift)))
        for _ in range(rng.randint(0, 4)):  # noise minimizers, some present in the target
            q.append((rng.choice([rng.randrange(1, 1 << 40), rng.choice(target)[0]]), rng.randrange(0, L)))
        if rng.random() < 0.2:
            rng.shuffle(q)
        queries.append(q)
    return target, queries


@pytest.mark.parametrize("seed", [1, 2])
def test_random_batches(px, seed):
    rng = random.Random(seed)
    target, queries = random_case(rng, 2500)  # three tiles of 1024 queries
    compare(px, target, queries, label=f"random {seed}")
    compare(px, target, queries, max_occ=3, label=f"random {seed} masked")
    compare(px, target, queries, label=f"random {seed} small params", seed_len=11, max_band=5,
            max_dist_x=30, max_iter=3, max_skip=1, min_chain_score=0)


def test_parameters_outside_the_32_bit_path(px):
    # A seed length or gap penalties beyond 2^20 take the 64-bit path for
    # every read; the results must not change character.
    rng = random.Random(4)
    target, queries = random_case(rng, 300)
    compare(px, target, queries, label="seed_len 3e6", seed_len=3_000_000, min_chain_score=0)
    compare(px, target, queries, label="gap penalty 4e6", gap_penalty=4.0e6, min_chain_score=0)
    compare(px, target, queries, label="negative gap penalty", gap_penalty=-0.5)


def test_dense_repeats_and_occurrence_cap(px):
    # Every query minimizer hits many target copies: long seed lists, ties
    # in target position, the skip heuristic and the iteration cap at work.
    rng = random.Random(5)
    target = [(rng.choice([1, 2, 3, 4, 5]), rng.randrange(0, 3000)) for _ in range(600)]
    queries = [[(rng.choice([1, 2, 3, 4, 5]), rng.randrange(0, 150)) for _ in range(rng.randint(0, 30))]
               for _ in range(200)]
    compare(px, target, queries, label="dense", max_iter=50, max_skip=3, min_chain_score=0)
    ix = px.SeedIndex(target)
    try:
        assert ix.occurrence_cap(16) == O.seed_occurrence_cap(target, 16)
        assert ix.occurrence_cap(10_000) == 10_000
    finally:
        ix.close()


def test_plan_reruns_and_one_shot_entry(px):
    from phraya_amd import _lib as L
    rng = random.Random(9)
    target, queries = random_case(rng, 300)
    want = compare(px, target, queries, label="reference run")
    ix = px.SeedIndex(target)
    plan = px.ChainPlan(ix, queries)
    try:
        for _ in range(3):
            plan.run()
            r = plan.results()
            assert np.array_equal(r.chains, want.chains) and np.array_equal(r.seeds, want.seeds)
        assert plan.last_kernel_ms() > 0
        # phx_seed_and_chain_batch with exact and short capacities.
        qv = np.concatenate([np.array([m[0] for m in q], np.uint64) for q in queries])
        qp = np.concatenate([np.array([m[1] for m in q], np.uint32) for q in queries])
        qo = np.concatenate([[0], np.cumsum([len(q) for q in queries])]).astype(np.uint64)
        ch = np.zeros(len(want.chains), L.CHAIN_DTYPE)
        sd = np.zeros(len(want.seeds), L.SEED_DTYPE)
        off = np.zeros(len(queries) + 1, np.uint64)
        ns = np.zeros(len(queries), np.uint64)
        prm = px.default_chain_params()
        out = L.CChainResults(ch.ctypes.data, len(ch), sd.ctypes.data, len(sd), off.ctypes.data, ns.ctypes.data, 0, 0)
        rc = L.lib().phx_seed_and_chain_batch(px.default_context().handle, ix._h, 2 ** 64 - 1, ctypes.byref(prm),
                                              qv.ctypes.data, qp.ctypes.data, qo.ctypes.data, len(queries),
                                              ctypes.byref(out))
        assert rc == L.PHX_OK and np.array_equal(ch, want.chains) and np.array_equal(sd, want.seeds)
        out.chains_cap = len(ch) - 1
        rc = L.lib().phx_seed_and_chain_batch(px.default_context().handle, ix._h, 2 ** 64 - 1, ctypes.byref(prm),
                                              qv.ctypes.data, qp.ctypes.data, qo.ctypes.data, len(queries),
                                              ctypes.byref(out))
        assert rc == L.PHX_E_CAPACITY and out.n_chains == len(ch)
    finally:
        plan.close()
        ix.close()
