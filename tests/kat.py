"""Known-answer vectors from phraya-align's own tests (tests/golden/kat.jsonl)
and an evaluator that applies them to any implementation of the API.

Each fixture line names the upstream test it was transcribed from (`src`).
"""
from __future__ import annotations

import json
import os

import dp

HERE = os.path.dirname(os.path.abspath(__file__))
KAT_PATH = os.path.join(HERE, "golden", "kat.jsonl")


def load():
    with open(KAT_PATH) as f:
        return [json.loads(line) for line in f if line.strip()]


class Impl:
    """Adapter: each method returns a dict with status ('ok', 'InvalidInput',
    'AlignmentFailed') and the Alignment fields."""

    def wfa_extend(self, q, t, seed, cap=None): raise NotImplementedError
    def wfa_extend_affine(self, q, t, seed, costs): raise NotImplementedError
    def myers_extend(self, q, t, seed): raise NotImplementedError
    def myers_edit_distance(self, q, t): raise NotImplementedError
    def lcp(self, a, b): raise NotImplementedError


def check(k: dict, impl: Impl) -> None:
    fn = k["fn"]
    ex = k["expect"]
    if fn == "lcp":
        assert impl.lcp(k["a"].encode(), k["b"].encode()) == ex["value"], k["src"]
        return
    q, t = k["q"].encode(), k["t"].encode()
    if fn in ("myers_vs_wfa", "myers_extend_vs_wfa"):
        w = impl.wfa_extend(q, t, (0, 0))
        if fn == "myers_vs_wfa":
            ed, cg = impl.myers_edit_distance(q, t)
            m = {"edit_distance": ed, "cigar": cg}
        else:
            m = impl.myers_extend(q, t, (0, 0))
        for f in ex["same"]:
            assert m[f] == w[f], (k["src"], f, m[f], w[f])
        return
    if fn == "myers_edit_distance":
        ed, cg = impl.myers_edit_distance(q, t)
        r = {"status": "ok", "edit_distance": ed, "cigar": cg}
    elif fn == "wfa_extend":
        r = impl.wfa_extend(q, t, tuple(k["seed"]))
    elif fn == "wfa_extend_affine":
        r = impl.wfa_extend_affine(q, t, tuple(k["seed"]), tuple(k["costs"]))
    elif fn == "fill_wfa_fitting_cap":
        r = impl.wfa_extend(q, t, (0, 0), k["cap"])
    else:
        raise AssertionError(f"unknown KAT fn {fn}")
    src = k["src"]
    if "error" in ex:
        assert r["status"] == ex["error"], (src, r)
        return
    if ex.get("not_fake_zero"):
        if r["status"] == "ok":
            assert r["edit_distance"] > 0 or r["cigar"] != "", src
        return
    assert r["status"] == "ok", (src, r)
    for f in ("cigar", "edit_distance", "query_start", "query_end", "target_start", "target_end"):
        if f in ex:
            assert r[f] == ex[f], (src, f, r[f], ex[f])
    if "cigar_contains" in ex:
        assert ex["cigar_contains"] in r["cigar"], (src, r["cigar"])
    if ex.get("cigar_nonempty"):
        assert r["cigar"], src
    if "edit_distance_min" in ex:
        assert r["edit_distance"] >= ex["edit_distance_min"], src
    if "edit_distance_max" in ex:
        assert r["edit_distance"] <= ex["edit_distance_max"], src
    if "target_end_min" in ex:
        assert r["target_end"] >= ex["target_end_min"], src
    if ex.get("replay"):
        qc, tc = dp.replay(r["cigar"], q, t)
        assert qc == len(q) and tc == r["target_end"], src
    if "indel_runs" in ex:
        assert sum(1 for _, op in dp.runs(r["cigar"]) if op in "ID") == ex["indel_runs"], (src, r["cigar"])
    if "affine_score" in ex:
        x, o, e = k["costs"]
        assert dp.affine_cost(r["cigar"], x, o, e) == ex["affine_score"], (src, r["cigar"])
        if "score" in r:
            assert r["score"] == ex["affine_score"], src


class OracleImpl(Impl):
    """The C++ oracle (oracle/liboracle.so)."""

    def __init__(self):
        from oracle import oracle as O
        self.O = O

    def _d(self, r):
        st = {0: "ok", 1: "InvalidInput", 2: "AlignmentFailed"}[r.status]
        return dict(status=st, cigar=r.cigar, edit_distance=r.edit_distance,
                    query_start=r.query_start, query_end=r.query_end,
                    target_start=r.target_start, target_end=r.target_end, score=r.score)

    def wfa_extend(self, q, t, seed, cap=None):
        return self._d(self.O.wfa_extend(q, t, seed, -1 if cap is None else cap))

    def wfa_extend_affine(self, q, t, seed, costs):
        return self._d(self.O.wfa_extend_affine(q, t, seed, costs))

    def myers_extend(self, q, t, seed):
        return self._d(self.O.myers_extend(q, t, seed))

    def myers_edit_distance(self, q, t):
        return self.O.myers_edit_distance(q, t)

    def lcp(self, a, b):
        return self.O.lcp(a, b)


class PyRefImpl(Impl):
    """The pure-Python restatement (tests/pyref.py)."""

    def __init__(self):
        import pyref
        self.P = pyref

    @staticmethod
    def _wrap(res, seed, qlen, qsuf):
        if res is None:
            return {"status": "AlignmentFailed"}
        cigar, ed, te = res[0], res[1], res[2]
        d = dict(status="ok", cigar=cigar, edit_distance=ed, query_start=seed[0],
                 query_end=seed[0] + qsuf, target_start=seed[1], target_end=seed[1] + te)
        if len(res) > 3:
            d["score"] = res[3]
        return d

    def wfa_extend(self, q, t, seed, cap=None):
        if seed[0] > len(q) or seed[1] > len(t):
            return {"status": "InvalidInput"}
        qs, ts = q[seed[0]:], t[seed[1]:]
        if not qs and not ts:
            return self._wrap(("", 0, 0), seed, len(q), 0)
        return self._wrap(self.P.wfa_fitting(qs, ts, cap), seed, len(q), len(qs))

    def wfa_extend_affine(self, q, t, seed, costs):
        if seed[0] > len(q) or seed[1] > len(t):
            return {"status": "InvalidInput"}
        qs, ts = q[seed[0]:], t[seed[1]:]
        return self._wrap(self.P.affine_fitting(qs, ts, *costs), seed, len(q), len(qs))

    def myers_extend(self, q, t, seed):
        if seed[0] > len(q) or seed[1] > len(t):
            return {"status": "InvalidInput"}
        qs, ts = q[seed[0]:], t[seed[1]:]
        ed, cg, te = self.P.myers_fitting(qs, ts)
        return self._wrap((cg, ed, te), seed, len(q), len(qs))

    def myers_edit_distance(self, q, t):
        return self.P.myers_global(q, t)

    def lcp(self, a, b):
        n = 0
        while n < min(len(a), len(b)) and a[n] == b[n]:
            n += 1
        return n
