"""Second, independent restatement of phraya-align's engines in plain Python.

Test infrastructure only, for small inputs.  Written separately from the C++
oracle (oracle/phraya_oracle.cpp) from the same upstream source, so that the
two can be cross-checked on random pairs: a coding slip in either shows up as
a disagreement, including in tie-breaking.

Upstream references (phraya-align/src/wfa_simd.rs): linear :231-488 and
:940-1130, affine :581-935, Myers :3507-3814, compaction :1200-1222.
"""
from __future__ import annotations

UNSET = None


def compact(ops: list[str]) -> str:
    out = []
    i = 0
    while i < len(ops):
        j = i
        while j < len(ops) and ops[j] == ops[i]:
            j += 1
        out.append(f"{j - i}{ops[i]}")
        i = j
    return "".join(out)


def lev_from_cigar(cigar: str) -> int:
    total, num = 0, ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
        else:
            if ch != "M":
                total += int(num or 0)
            num = ""
    return total


def _ext(q: bytes, t: bytes, p: int, k: int) -> int:
    j = p - k
    if p < 0 or p >= len(q) or j < 0 or j >= len(t):
        return p
    while p < len(q) and j < len(t) and q[p] == t[j]:
        p += 1
        j += 1
    return p


# ----------------------------------------------------------------------------- linear

def _linear_frames(q: bytes, t: bytes, fitting: bool, explore: int):
    """Returns (frames, s, t_end) or (frames, None, None) when exhausted.
    frames[s] = dict k -> (pos, op)."""
    qn, tn = len(q), len(t)
    frames = [{0: (_ext(q, t, 0, 0), 0)}]

    def end_of(fr):
        if not fitting:
            v = fr.get(qn - tn)
            return tn if v is not None and v[0] >= qn else None
        for k in range(qn - tn, qn + 1):
            te = qn - k
            if 0 <= te <= tn and k in fr and fr[k][0] >= qn:
                return te
        return None

    te = end_of(frames[0])
    if te is not None:
        return frames, 0, te
    for s in range(1, explore + 1):
        prev = frames[-1]
        cur = {}
        for k in range(-min(s, tn), min(s, qn) + 1):
            cands = []
            if k in prev:
                cands.append((prev[k][0] + 1, 1))
            if k > -tn and (k - 1) in prev:
                cands.append((prev[k - 1][0] + 1, 2))
            if k < qn and (k + 1) in prev:
                cands.append((prev[k + 1][0], 3))
            if not cands:
                continue
            best = cands[0]
            for c in cands[1:]:
                if c[0] >= best[0]:  # last maximum wins
                    best = c
            pos, op = best
            j = pos - k
            if pos < 0 or pos > qn or j < 0 or j > tn:
                continue
            cur[k] = (_ext(q, t, pos, k), op, pos)
        frames.append(cur)
        te = end_of(cur)
        if te is not None:
            return frames, s, te
    return frames, None, None


def _linear_traceback(frames, q, t, s, t_end) -> str:
    qi, ti = len(q), t_end
    k = qi - ti
    ops = []
    while qi > 0 or ti > 0:
        if s == 0:
            while qi > 0 and ti > 0:
                ops.append("M")
                qi -= 1
                ti -= 1
            break
        _, op, pre = frames[s][k]
        for _ in range(max(qi - pre, 0)):
            ops.append("M")
            qi -= 1
            ti -= 1
        if op == 1:
            ops.append("X"); qi -= 1; ti -= 1
        elif op == 2:
            ops.append("D"); qi -= 1; k -= 1
        else:
            ops.append("I"); ti -= 1; k += 1
        s -= 1
    return compact(ops[::-1])


def wfa_fitting(q: bytes, t: bytes, cap=None):
    """fill_wfa_fitting_impl -> (cigar, ed, t_consumed) or None."""
    qn, tn = len(q), len(t)
    if qn == 0:
        return "", 0, 0
    fitting = tn > qn + qn // 2 + 10
    if not fitting:
        frames, s, te = _linear_frames(q, t, False, qn + tn)
        if s is None:
            return "", 0, tn
        if cap is not None and s > cap:
            return None
        return _linear_traceback(frames, q, t, s, te), s, tn
    explore = qn + tn if cap is None else min(cap, qn + tn)
    frames, s, te = _linear_frames(q, t, True, explore)
    if s is None:
        return None
    return _linear_traceback(frames, q, t, s, te), s, te


# ----------------------------------------------------------------------------- affine

def affine_fitting(q: bytes, t: bytes, x: int, o: int, e: int, cap=None):
    """fill_wfa_affine_fitting_impl -> (cigar, ed, t_consumed, score) or None."""
    qn, tn = len(q), len(t)
    if qn == 0:
        return "", 0, 0, 0
    fitting = tn > qn + qn // 2 + 10

    def valid(p, k):
        return p is not None and 0 <= p <= qn and 0 <= p - k <= tn

    # frame: dict k -> [m, ins, del, m_op, ins_op, del_op]
    H = [{0: [_ext(q, t, 0, 0), None, None, 0, 0, 0]}]

    def get(s, k, layer):
        if s < 0 or s >= len(H):
            return None
        c = H[s].get(k)
        return None if c is None else c[layer]

    def endpoint(fr):
        def chk(k):
            c = fr.get(k)
            if c is None:
                return None
            for layer in range(3):
                if c[layer] == qn:
                    return layer
            return None
        if not fitting:
            lay = chk(qn - tn)
            return None if lay is None else (tn, qn - tn, lay)
        for k in range(qn - tn, qn + 1):
            te = qn - k
            if 0 <= te <= tn:
                lay = chk(k)
                if lay is not None:
                    return te, k, lay
        return None

    end = endpoint(H[0])
    s_end = 0
    if end is None:
        hard = (qn + tn) * max(x, 1) + o + e + 4
        explore = hard if cap is None else min(cap, hard)
        for s in range(1, explore + 1):
            cur = {}
            so, se, sx = s - o - e, s - e, s - x
            for k in range(-tn, qn + 1):
                ins, iop = None, 0
                if k - 1 >= -tn:
                    p = get(so, k - 1, 0)
                    if p is not None and valid(p + 1, k):
                        ins, iop = p + 1, 0
                    p = get(se, k - 1, 1)
                    if p is not None and valid(p + 1, k) and (ins is None or p + 1 > ins):
                        ins, iop = p + 1, 1
                dele, dop = None, 0
                if k + 1 <= qn:
                    p = get(so, k + 1, 0)
                    if p is not None and valid(p, k):
                        dele, dop = p, 0
                    p = get(se, k + 1, 2)
                    if p is not None and valid(p, k) and (dele is None or p > dele):
                        dele, dop = p, 1
                best, mop = None, 0
                p = get(sx, k, 0)
                if p is not None and valid(p + 1, k):
                    best, mop = p + 1, 0
                if ins is not None and valid(ins, k) and (best is None or ins > best):
                    best, mop = ins, 1
                if dele is not None and valid(dele, k) and (best is None or dele > best):
                    best, mop = dele, 2
                m = None if best is None else _ext(q, t, best, k)
                if m is not None or ins is not None or dele is not None:
                    cur[k] = [m, ins, dele, mop, iop, dop]
            H.append(cur)
            end = endpoint(cur)
            if end is not None:
                s_end = s
                break
        if end is None:
            return None
    t_end, k, layer = end
    qi, ti, s = qn, t_end, s_end
    ops = []
    while qi > 0 or ti > 0:
        if s == 0:
            while qi > 0 and ti > 0:
                ops.append("M"); qi -= 1; ti -= 1
            break
        c = H[s][k]
        if layer == 0:
            if c[3] == 0:
                raw = H[s - x][k][0] + 1
                for _ in range(max(qi - raw, 0)):
                    ops.append("M"); qi -= 1; ti -= 1
                ops.append("X"); qi -= 1; ti -= 1
                s -= x
            else:
                raw = c[c[3]]
                for _ in range(max(qi - raw, 0)):
                    ops.append("M"); qi -= 1; ti -= 1
                layer = c[3]
        elif layer == 1:
            ops.append("D"); qi -= 1
            if c[4] == 0:
                s -= o + e; layer = 0
            else:
                s -= e
            k -= 1
        else:
            ops.append("I"); ti -= 1
            if c[5] == 0:
                s -= o + e; layer = 0
            else:
                s -= e
            k += 1
    cigar = compact(ops[::-1])
    return cigar, lev_from_cigar(cigar), t_end, s_end


# ----------------------------------------------------------------------------- Myers

def _myers_cols(q: bytes, t: bytes, margin):
    m = len(q)
    W = 64
    nb = (m + W - 1) // W
    M64 = (1 << 64) - 1
    last_bits = W if m % W == 0 else m % W
    last_mask = M64 if last_bits == W else (1 << last_bits) - 1
    score_bit = (m - 1) % W
    peq = [dict() for _ in range(nb)]
    for i, b in enumerate(q):
        peq[i // W][b] = peq[i // W].get(b, 0) | (1 << (i % W))
    vp = [M64] * nb
    vp[-1] = last_mask
    vn = [0] * nb
    score = best = m
    cols = []
    for tb in t:
        carry, hpc, hnc = 0, 1, 0
        lhp = lhn = 0
        nvp, nvn = [0] * nb, [0] * nb
        for b in range(nb):
            eq = peq[b].get(tb, 0)
            xh = eq | vn[b]
            ssum = vp[b] + (xh & vp[b]) + carry
            carry = ssum >> 64
            ssum &= M64
            d0 = (ssum ^ vp[b]) | xh
            hn = vp[b] & d0
            hp = (vn[b] | ~(vp[b] | d0)) & M64
            if b == nb - 1 and last_bits < W:
                d0 &= last_mask; hn &= last_mask; hp &= last_mask
            if b == nb - 1:
                lhp, lhn = hp, hn
            xx = ((hp << 1) & M64) | hpc
            hpc = hp >> 63
            hs = ((hn << 1) & M64) | hnc
            hnc = hn >> 63
            nvn[b] = xx & d0
            v2 = (hs | ~(d0 | xx)) & M64
            nvp[b] = v2 & last_mask if (b == nb - 1 and last_bits < W) else v2
        if (lhp >> score_bit) & 1:
            score += 1
        if (lhn >> score_bit) & 1:
            score = max(score - 1, 0)
        vp, vn = nvp, nvn
        cols.append((vp, vn, score))
        best = min(best, score)
        if margin is not None and score > best + margin:
            break
    return cols


def _col(cols, m, j):
    if j == 0:
        return list(range(m + 1))
    vp, vn, sc = cols[j - 1]
    c = [0] * (m + 1)
    c[m] = sc
    for i in range(m - 1, -1, -1):
        c[i] = c[i + 1] - ((vp[i // 64] >> (i % 64)) & 1) + ((vn[i // 64] >> (i % 64)) & 1)
    return c


def _myers_bt(q, t, cols, end):
    m = len(q)
    qi, ti = m, end
    ops = []
    while qi > 0 or ti > 0:
        if qi == 0:
            ops.append("I"); ti -= 1; continue
        if ti == 0:
            ops.append("D"); qi -= 1; continue
        cur, prev = _col(cols, m, ti), _col(cols, m, ti - 1)
        c = cur[qi]
        if c == cur[qi - 1] + 1:
            ops.append("D"); qi -= 1
        elif c == prev[qi] + 1:
            ops.append("I"); ti -= 1
        else:
            ops.append("M" if q[qi - 1] == t[ti - 1] else "X"); qi -= 1; ti -= 1
    return compact(ops[::-1])


def myers_global(q: bytes, t: bytes):
    m, n = len(q), len(t)
    if m == 0:
        return n, (f"{n}I" if n else "")
    if n == 0:
        return m, f"{m}D"
    cols = _myers_cols(q, t, None)
    return cols[-1][2], _myers_bt(q, t, cols, n)


def myers_fitting(q: bytes, t: bytes):
    m, n = len(q), len(t)
    if m == 0:
        return 0, "", 0
    if n == 0:
        return m, f"{m}D", 0
    if n <= m + m // 2 + 10:
        ed, cg = myers_global(q, t)
        return ed, cg, n
    cols = _myers_cols(q, t, m // 10 + 1)
    best, bc = m, 0
    for c, (_, _, sc) in enumerate(cols):
        if sc < best:
            best, bc = sc, c + 1
    return best, _myers_bt(q, t, cols, bc), bc
