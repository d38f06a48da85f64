"""Independent closed-form checkers (test infrastructure only).

These do not restate phraya's wavefront or bit-vector code: they are the
textbook O(n*m) recurrences the engines must agree with.
  * levenshtein(q, t): unit-cost DP matrix D (the quantity phraya's fill_scalar
    oracle computes, phraya-align/src/wfa_simd.rs:1135-1156).
  * gotoh(q, t, x, o, e): three-state affine DP where an L-base gap costs
    o + L*e and a gap may open from any state (the cost model of
    phraya-align/src/wfa_simd.rs:517-519 and its Gotoh test oracle :3118-3170).
  * replay(cigar, q, t): walks a CIGAR and checks M = true match, X = true
    mismatch, D consumes query only, I consumes target only.
  * affine_cost(cigar): sum of x per X and o + L*e per I/D run.
Vectorised variants run over many equal-shape pairs at once (brute force).
"""
from __future__ import annotations

import re

import numpy as np

BIG = 1 << 40

_RUN = re.compile(r"(\d+)([MXID])")


def runs(cigar: str) -> list[tuple[int, str]]:
    out = [(int(n), op) for n, op in _RUN.findall(cigar)]
    assert "".join(f"{n}{op}" for n, op in out) == cigar, f"malformed CIGAR {cigar!r}"
    return out


def levenshtein(q: bytes, t: bytes) -> np.ndarray:
    qn, tn = len(q), len(t)
    D = np.zeros((qn + 1, tn + 1), dtype=np.int64)
    D[:, 0] = np.arange(qn + 1)
    D[0, :] = np.arange(tn + 1)
    tq = np.frombuffer(t, dtype=np.uint8) if tn else np.zeros(0, np.uint8)
    for i in range(1, qn + 1):
        sub = (tq != q[i - 1]).astype(np.int64)
        row = np.minimum(D[i - 1, 1:] + 1, D[i - 1, :-1] + sub)
        # left-to-right dependency: D[i][j] = min(row[j-1], D[i][j-1] + 1)
        cur = np.empty(tn + 1, dtype=np.int64)
        cur[0] = i
        # prefix-min trick: D[i][j] = min_k<=j (row'[k] + (j-k)) with row'[0] = i
        vals = np.concatenate(([i], row)) - np.arange(tn + 1)
        cur = np.minimum.accumulate(vals) + np.arange(tn + 1)
        D[i] = cur
    return D


def gotoh(q: bytes, t: bytes, x: int, o: int, e: int) -> np.ndarray:
    """Returns the (qn+1, tn+1) matrix of min over the three states."""
    qn, tn = len(q), len(t)
    M = np.full((qn + 1, tn + 1), BIG, dtype=np.int64)
    Iq = np.full((qn + 1, tn + 1), BIG, dtype=np.int64)  # gap consuming query ('D')
    It = np.full((qn + 1, tn + 1), BIG, dtype=np.int64)  # gap consuming target ('I')
    M[0, 0] = 0
    for i in range(1, qn + 1):
        Iq[i, 0] = o + e * i
    for j in range(1, tn + 1):
        It[0, j] = o + e * j
    for i in range(0, qn + 1):
        for j in range(0, tn + 1):
            if i == 0 and j == 0:
                continue
            if i > 0 and j > 0:
                best_prev = min(M[i - 1, j - 1], Iq[i - 1, j - 1], It[i - 1, j - 1])
                M[i, j] = best_prev + (0 if q[i - 1] == t[j - 1] else x)
            if i > 0 and j > 0:
                Iq[i, j] = min(min(M[i - 1, j], It[i - 1, j]) + o + e, Iq[i - 1, j] + e)
            if j > 0 and i > 0:
                It[i, j] = min(min(M[i, j - 1], Iq[i, j - 1]) + o + e, It[i, j - 1] + e)
            if i > 0 and j == 0:
                pass
            if j > 0 and i == 0:
                pass
    return np.minimum(np.minimum(M, Iq), It)


def replay(cigar: str, q: bytes, t: bytes) -> tuple[int, int]:
    """Checks op semantics; returns (query consumed, target consumed)."""
    qi = ti = 0
    for n, op in runs(cigar):
        assert n > 0
        if op in "MX":
            for _ in range(n):
                assert qi < len(q) and ti < len(t), "CIGAR runs past a sequence"
                if op == "M":
                    assert q[qi] == t[ti], f"M at q[{qi}]/t[{ti}] is not a match"
                else:
                    assert q[qi] != t[ti], f"X at q[{qi}]/t[{ti}] is not a mismatch"
                qi += 1
                ti += 1
        elif op == "D":
            qi += n
        else:
            ti += n
    return qi, ti


def affine_cost(cigar: str, x: int, o: int, e: int) -> int:
    c = 0
    for n, op in runs(cigar):
        if op == "X":
            c += n * x
        elif op in "ID":
            c += o + n * e
    return c


def unit_cost(cigar: str) -> int:
    return sum(n for n, op in runs(cigar) if op != "M")


def compacted(cigar: str) -> bool:
    r = runs(cigar)
    return all(r[i][1] != r[i + 1][1] for i in range(len(r) - 1))


# ------------------------------------------------------------------ vectorised (brute force)

def levenshtein_many(Q: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Q: (P, qn) uint8, T: (P, tn) uint8 -> (P, qn+1, tn+1) DP matrices."""
    P, qn = Q.shape
    tn = T.shape[1]
    D = np.zeros((P, qn + 1, tn + 1), dtype=np.int64)
    D[:, :, 0] = np.arange(qn + 1)
    D[:, 0, :] = np.arange(tn + 1)
    for i in range(1, qn + 1):
        for j in range(1, tn + 1):
            sub = (Q[:, i - 1] != T[:, j - 1]).astype(np.int64)
            D[:, i, j] = np.minimum(np.minimum(D[:, i - 1, j] + 1, D[:, i, j - 1] + 1),
                                    D[:, i - 1, j - 1] + sub)
    return D


def gotoh_many(Q: np.ndarray, T: np.ndarray, x: int, o: int, e: int) -> np.ndarray:
    P, qn = Q.shape
    tn = T.shape[1]
    M = np.full((P, qn + 1, tn + 1), BIG, dtype=np.int64)
    Iq = np.full_like(M, BIG)
    It = np.full_like(M, BIG)
    M[:, 0, 0] = 0
    for i in range(1, qn + 1):
        Iq[:, i, 0] = o + e * i
    for j in range(1, tn + 1):
        It[:, 0, j] = o + e * j
    for i in range(1, qn + 1):
        for j in range(1, tn + 1):
            bp = np.minimum(np.minimum(M[:, i - 1, j - 1], Iq[:, i - 1, j - 1]), It[:, i - 1, j - 1])
            M[:, i, j] = bp + np.where(Q[:, i - 1] == T[:, j - 1], 0, x)
            Iq[:, i, j] = np.minimum(np.minimum(M[:, i - 1, j], It[:, i - 1, j]) + o + e,
                                     Iq[:, i - 1, j] + e)
            It[:, i, j] = np.minimum(np.minimum(M[:, i, j - 1], Iq[:, i, j - 1]) + o + e,
                                     It[:, i, j - 1] + e)
    return np.minimum(np.minimum(M, Iq), It)
