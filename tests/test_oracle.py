"""Pins the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

* Known answers transcribed from phraya-align's tests (tests/golden/kat.jsonl).
* Closed forms: Levenshtein DP for linear WFA and Myers distances, Gotoh DP
  for the affine score, CIGAR replay, cost-of-CIGAR == score, fitting end =
  largest target end reaching the optimum.
* Brute force over every ACGT pair with |q|, |t| <= 4.
* A second, independently written restatement (tests/pyref.py) must produce
  the identical CIGAR (tie-breaking included) on random pairs, and a planted
  tie-rule mistake in that restatement must be caught by the cross-check.
"""
from __future__ import annotations

import itertools
import random

import numpy as np
import pytest

import dp
import kat
import pyref
from oracle import oracle as O

ACGT = b"ACGT"


def rand_seq(rng, n, alphabet=ACGT):
    return bytes(rng.choice(alphabet) for _ in range(n))


def mutate(rng, s: bytes, rate: float, alphabet=ACGT) -> bytes:
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
        if r < rate:
            out.append(rng.choice([c for c in alphabet if c != s[i]]))
        else:
            out.append(s[i])
        i += 1
    return bytes(out)


def rand_pair(rng, qmax=60, fitting_frac=0.3):
    qn = rng.randint(0, qmax)
    q = rand_seq(rng, qn)
    t = mutate(rng, q, rng.choice([0.0, 0.02, 0.1, 0.3]))
    if rng.random() < fitting_frac:
        t = t + rand_seq(rng, qn // 2 + rng.randint(11, 30))
    elif rng.random() < 0.3:
        t = t + rand_seq(rng, rng.randint(0, 10))
    return q, t


# --------------------------------------------------------------------------- KATs

@pytest.mark.parametrize("k", kat.load(), ids=lambda k: k["src"].split("/")[-1])
def test_kat_oracle(k):
    kat.check(k, kat.OracleImpl())


@pytest.mark.parametrize("k", kat.load(), ids=lambda k: k["src"].split("/")[-1])
def test_kat_pyref(k):
    kat.check(k, kat.PyRefImpl())


# ---------------------------------------------------------------------- LCP tiers

def test_lcp_battery():
    """count_matching_prefix tiers agree with scalar (wfa_simd.rs:3831-3889)."""
    cases = []
    for n in range(40):
        a = bytes((i * 7 + 3) % 251 for i in range(n))
        cases.append((a, a))
        for p in range(n):
            b = bytearray(a)
            b[p] ^= 0x5A
            cases.append((a, bytes(b)))
        cases.append((a, a[: n // 2]))
        cases.append((a[: n // 3], a))
    cases.append((b"", b"ACGT"))
    cases.append((b"ACGT", b""))
    for a, b in cases:
        want = 0
        while want < min(len(a), len(b)) and a[want] == b[want]:
            want += 1
        for tier in ("sse2", "u64", "scalar"):
            assert O.lcp(a, b, tier) == want, (tier, a, b)


# ------------------------------------------------------------------ closed forms

def _fitting(qn, tn):
    return tn > qn + qn // 2 + 10


def test_linear_matches_levenshtein():
    rng = random.Random(11)
    for _ in range(400):
        q, t = rand_pair(rng)
        r = O.wfa_extend(q, t)
        assert r.status == 0
        if not q:
            assert r.cigar == "" and r.edit_distance == 0
            continue
        D = dp.levenshtein(q, t)
        if _fitting(len(q), len(t)):
            last = D[len(q)]
            best = int(last.min())
            assert r.edit_distance == best
            # fitting end: the largest target end reaching the optimum
            assert r.target_end == int(np.nonzero(last == best)[0].max())
        else:
            assert r.edit_distance == int(D[len(q), len(t)])
            assert r.target_end == len(t)
        qc, tc = dp.replay(r.cigar, q, t)
        assert (qc, tc) == (len(q), r.target_end)
        assert dp.unit_cost(r.cigar) == r.edit_distance
        assert dp.compacted(r.cigar)


@pytest.mark.parametrize("costs", [(2, 4, 1), (4, 6, 2), (1, 0, 1), (3, 5, 2), (1, 1, 1)])
def test_affine_matches_gotoh(costs):
    x, o, e = costs
    rng = random.Random(hash(costs) & 0xFFFF)
    for _ in range(150):
        q, t = rand_pair(rng, qmax=40)
        r = O.wfa_extend_affine(q, t, costs=costs)
        assert r.status == 0
        if not q:
            assert r.cigar == ""
            continue
        G = dp.gotoh(q, t, x, o, e)
        if _fitting(len(q), len(t)):
            best = int(G[len(q)].min())
            assert r.target_end == int(np.nonzero(G[len(q)] == best)[0].max())
        else:
            best = int(G[len(q), len(t)])
            assert r.target_end == len(t)
        assert r.score == best
        assert dp.affine_cost(r.cigar, x, o, e) == best
        qc, tc = dp.replay(r.cigar, q, t)
        assert (qc, tc) == (len(q), r.target_end)
        assert r.edit_distance == dp.unit_cost(r.cigar)
        assert dp.compacted(r.cigar)


@pytest.mark.parametrize("costs", [None, (4, 6, 2), (2, 4, 1)])
def test_global_engines_match_dp(costs):
    """The global branches alone (fill_wfa, fill_wfa_affine_generic with
    fitting = false), used for gaps pinned at both ends: the global DP optimum
    on every shape, including the long targets for which the wrappers switch
    to fitting and empty queries; identical to the wrappers wherever those
    take the global branch themselves."""
    rng = random.Random(77 if costs is None else sum(costs))
    for it in range(200):
        q, t = rand_pair(rng, qmax=40, fitting_frac=0.5)
        if it % 25 == 0:
            q = b""
        if costs is None:
            r = O.wfa_global(q, t)
            best = int(dp.levenshtein(q, t)[len(q), len(t)])
            assert dp.unit_cost(r.cigar) == best
            wrapped = O.wfa_extend(q, t)
        else:
            x, o, e = costs
            r = O.wfa_affine_global(q, t, costs)
            best = int(dp.gotoh(q, t, x, o, e)[len(q), len(t)]) if (q or t) else 0
            assert dp.affine_cost(r.cigar, x, o, e) == best
            assert r.edit_distance == dp.unit_cost(r.cigar)
            wrapped = O.wfa_extend_affine(q, t, costs=costs)
        assert r.status == 0 and r.score == best
        assert r.target_end == len(t)
        assert dp.replay(r.cigar, q, t) == (len(q), len(t))
        assert dp.compacted(r.cigar)
        if q and not _fitting(len(q), len(t)):
            assert (r.cigar, r.score, r.edit_distance) == (wrapped.cigar, wrapped.score,
                                                           wrapped.edit_distance)


def test_myers_matches_levenshtein():
    rng = random.Random(5)
    for _ in range(400):
        q, t = rand_pair(rng, qmax=150)
        ed, cigar = O.myers_edit_distance(q, t)
        D = dp.levenshtein(q, t)
        assert ed == int(D[len(q), len(t)])
        if q or t:
            assert dp.replay(cigar, q, t) == (len(q), len(t))
            assert dp.unit_cost(cigar) == ed
        r = O.myers_extend(q, t)
        if not q:
            assert (r.edit_distance, r.cigar, r.target_end) == (0, "", 0)
            continue
        if _fitting(len(q), len(t)):
            # First strict minimum among the columns scanned before the
            # abandon margin m/10 + 1 trips (wfa_simd.rs:3799-3810), read off the
            # DP's bottom row.
            col = D[len(q), 1:]
            margin = len(q) // 10 + 1
            best_seen, best, bc = len(q), len(q), 0
            for j, s in enumerate(col):
                best_seen = min(best_seen, int(s))
                if s < best:
                    best, bc = int(s), j + 1
                if s > best_seen + margin:
                    break
            assert (r.edit_distance, r.target_end) == (best, bc)
        else:
            assert r.edit_distance == int(D[len(q), len(t)]) and r.target_end == len(t)
        assert dp.replay(r.cigar, q, t) == (len(q), r.target_end)


# -------------------------------------------------------------------- brute force

def _all(n):
    return [bytes(p) for p in itertools.product(ACGT, repeat=n)]


@pytest.mark.slow
def test_brute_force_tiny():
    """Every ACGT pair with |q|, |t| <= 4: distances and scores equal the DP
    optimum, CIGARs replay, fitting ends are the largest optimal target end."""
    for qn in range(0, 5):
        Qs = _all(qn)
        for tn in range(0, 5):
            Ts = _all(tn)
            pairs = [(q, t) for q in Qs for t in Ts]
            q_seq = np.frombuffer(b"".join(q for q, _ in pairs) or b"\0", np.uint8)
            t_seq = np.frombuffer(b"".join(t for _, t in pairs) or b"\0", np.uint8)
            P = len(pairs)
            q_off = np.arange(P + 1, dtype=np.uint64) * qn
            t_off = np.arange(P + 1, dtype=np.uint64) * tn
            Qa = np.frombuffer(b"".join(q for q, _ in pairs), np.uint8).reshape(P, qn) if qn else np.zeros((P, 0), np.uint8)
            Ta = np.frombuffer(b"".join(t for _, t in pairs), np.uint8).reshape(P, tn) if tn else np.zeros((P, 0), np.uint8)
            D = dp.levenshtein_many(Qa, Ta)
            lin, lc = O.batch(O.LINEAR, q_seq, q_off, t_seq, t_off)
            myg, mc = O.batch(O.MYERS_GLOBAL, q_seq, q_off, t_seq, t_off)
            for (x, o, e) in ((2, 4, 1), (4, 6, 2)):
                G = dp.gotoh_many(Qa, Ta, x, o, e)
                aff, ac = O.batch(O.AFFINE, q_seq, q_off, t_seq, t_off, costs=(x, o, e))
                for i, (q, t) in enumerate(pairs):
                    if qn == 0:
                        assert ac[i] == "" and aff["score"][i] == 0
                        continue
                    best = int(G[i, qn, tn])  # tn <= qn + qn/2 + 10 always here: global
                    assert aff["score"][i] == best
                    assert dp.affine_cost(ac[i], x, o, e) == best
                    assert dp.replay(ac[i], q, t) == (qn, tn)
            for i, (q, t) in enumerate(pairs):
                assert myg["edit_distance"][i] == D[i, qn, tn]
                if qn or tn:
                    assert dp.replay(mc[i], q, t) == (qn, tn)
                if qn == 0:
                    assert lc[i] == "" and lin["edit_distance"][i] == 0
                    continue
                assert lin["edit_distance"][i] == D[i, qn, tn]
                assert dp.replay(lc[i], q, t) == (qn, tn)


# -------------------------------------------------- second restatement cross-check

def test_oracle_equals_pyref_linear_and_myers():
    rng = random.Random(21)
    for _ in range(300):
        q, t = rand_pair(rng, qmax=70)
        r = O.wfa_extend(q, t)
        p = pyref.wfa_fitting(q, t) if (q or t) else ("", 0, 0)
        assert (r.cigar, r.edit_distance, r.target_end) == (p[0], p[1], p[2]), (q, t)
        ed, cg = O.myers_edit_distance(q, t)
        assert (ed, cg) == pyref.myers_global(q, t), (q, t)
        r = O.myers_extend(q, t)
        assert (r.edit_distance, r.cigar, r.target_end) == pyref.myers_fitting(q, t), (q, t)


@pytest.mark.parametrize("costs", [(2, 4, 1), (4, 6, 2), (1, 0, 1)])
def test_oracle_equals_pyref_affine(costs):
    rng = random.Random(31 + costs[0])
    for _ in range(120):
        q, t = rand_pair(rng, qmax=45)
        r = O.wfa_extend_affine(q, t, costs=costs)
        p = pyref.affine_fitting(q, t, *costs)
        assert (r.cigar, r.edit_distance, r.target_end, r.score) == p, (q, t)


def test_cap_semantics_match_pyref():
    rng = random.Random(41)
    for _ in range(150):
        q, t = rand_pair(rng, qmax=40)
        cap = rng.randint(0, 12)
        r = O.wfa_extend(q, t, cap=cap)
        p = pyref.wfa_fitting(q, t, cap) if q else ("", 0, 0)
        if p is None:
            assert r.status == O.ALIGNMENT_FAILED
        else:
            assert r.status == 0 and (r.cigar, r.edit_distance, r.target_end) == p
        ra = O.wfa_extend_affine(q, t, costs=(4, 6, 2), cap=cap * 4)
        pa = pyref.affine_fitting(q, t, 4, 6, 2, cap * 4)
        if pa is None:
            assert ra.status == O.ALIGNMENT_FAILED
        else:
            assert ra.status == 0 and (ra.cigar, ra.edit_distance, ra.target_end, ra.score) == pa


def test_crosscheck_detects_planted_tie_rule_mistake(monkeypatch):
    """Plant a wrong tie rule (first maximum instead of last) in the Python
    restatement: the oracle cross-check must notice on random pairs."""
    orig = pyref._linear_frames

    def first_max_frames(q, t, fitting, explore):
        qn, tn = len(q), len(t)
        frames = [{0: (pyref._ext(q, t, 0, 0), 0)}]
        for s in range(1, explore + 1):
            prev, cur = frames[-1], {}
            for k in range(-min(s, tn), min(s, qn) + 1):
                c = []
                if k in prev:
                    c.append((prev[k][0] + 1, 1))
                if k > -tn and (k - 1) in prev:
                    c.append((prev[k - 1][0] + 1, 2))
                if k < qn and (k + 1) in prev:
                    c.append((prev[k + 1][0], 3))
                if not c:
                    continue
                pos, op = max(c, key=lambda v: v[0])  # first maximum: the planted mistake
                if 0 <= pos <= qn and 0 <= pos - k <= tn:
                    cur[k] = (pyref._ext(q, t, pos, k), op, pos)
            frames.append(cur)
            v = cur.get(qn - tn)
            if not fitting and v is not None and v[0] >= qn:
                return frames, s, tn
        return orig(q, t, fitting, explore)

    monkeypatch.setattr(pyref, "_linear_frames", first_max_frames)
    rng = random.Random(3)
    mismatches = 0
    for _ in range(200):
        q = rand_seq(rng, rng.randint(5, 30))
        t = mutate(rng, q, 0.3)
        if _fitting(len(q), len(t)):
            continue
        r = O.wfa_extend(q, t)
        p = pyref.wfa_fitting(q, t)
        if (r.cigar, r.edit_distance) != (p[0], p[1]):
            mismatches += 1
    assert mismatches > 0
