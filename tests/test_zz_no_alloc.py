"""phx_align_batch_noalloc through the C ABI: every field and CIGAR run equals
phx_align_batch's on the same batch, for two calls in a row (the second
smaller, so what the first left in the shared scratch array must not leak
into it), for every engine; batches that do not fit the scratch array are
refused with PHX_E_NOMEM and the next call is unaffected."""
from __future__ import annotations

import ctypes
import random
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


def rand_seq(rng, n, alphabet=b"ACGT"):
    return bytes(rng.choice(alphabet) for _ in range(n))


def mutate(rng, s, rate):
    out = bytearray()
    i = 0
    while i < len(s):
        r = rng.random()
        if r < rate * 0.3:
            out += rand_seq(rng, rng.randint(1, 3))
            continue
        if r < rate * 0.6:
            i += rng.randint(1, 3)
            continue
        out.append(rng.choice([c for c in b"ACGT" if c != s[i]]) if r < rate else s[i])
        i += 1
    return bytes(out)


def batch(seed, n_short, longs):
    """Short reads against windows (LDS tier), long pairs (length, error
    rate) for the HBM wave and workgroup tiers, a byte pair, an invalid seed
    and empty sequences."""
    from phraya_amd import synth
    rng = random.Random(seed)
    w = synth.workload("short150", n=n_short, seed=seed)
    qs = [w.pair(i)[0] for i in range(w.n)]
    ts = [w.pair(i)[1] for i in range(w.n)]
    seeds = [(0, 0)] * w.n
    for L, rate in longs:
        t = rand_seq(rng, int(L * 1.1))
        qs.append(mutate(rng, t[:L], rate))
        ts.append(t)
        seeds.append((0, 0))
    q = rand_seq(rng, 200)
    qs += [q[:100] + b"N" + q[101:], q, b"", b"ACGT"]
    ts += [mutate(rng, q, 0.02) + b"nn", q, b"ACG", b""]
    seeds += [(0, 0), (500, 0), (0, 0), (0, 0)]
    return qs, ts, seeds


def call(fn, ctx, pb, engine, costs):
    import phraya_amd as px
    from phraya_amd import _lib as L
    cb = px._cbatch(pb, engine, costs, None)
    cap = int(L.lib().phx_cigar_bound(ctypes.byref(cb)))
    aln = np.zeros(pb.n, L.ALN_DTYPE)
    runs = np.zeros(max(cap, 1), np.uint32)
    off = np.zeros(pb.n + 1, np.uint64)
    r = L.CResults(aln.ctypes.data, runs.ctypes.data, cap, off.ctypes.data)
    rc = fn(ctx.handle, ctypes.byref(cb), ctypes.byref(r))
    return rc, aln, runs, off


def same(a, b, label, ctx=None):
    rc_a, aln_a, runs_a, off_a = a
    rc_b, aln_b, runs_b, off_b = b
    assert rc_a == rc_b == 0, (label, rc_a, rc_b, ctx.last_error() if ctx else "")
    assert np.array_equal(aln_a, aln_b), f"{label}: per-pair fields differ"
    assert np.array_equal(off_a, off_b), f"{label}: run offsets differ"
    n = int(off_a[-1])
    assert np.array_equal(runs_a[:n], runs_b[:n]), f"{label}: CIGAR runs differ"


ENGINES = [("affine", 1, (4, 6, 2)), ("linear", 0, (0, 0, 0)), ("myers", 2, (0, 0, 0)),
           ("myers_global", 3, (0, 0, 0)), ("myers_dist", 4, (0, 0, 0))]


@pytest.mark.parametrize("name,engine,costs", ENGINES, ids=[e[0] for e in ENGINES])
def test_noalloc_matches_align_batch(name, engine, costs):
    import phraya_amd as px
    from phraya_amd import _lib as L
    ctx = px.default_context()
    lib = L.lib()
    c = px.AffineCosts(*costs) if engine == 1 else px.AffineCosts()
    # The long pairs' wavefront history (host memory model: 6, 8, 13 and 10
    # MiB) fits the 64 MiB array together with the short reads' buffers.
    for call_no, (seed, n_short, longs) in enumerate([
            (3, 3000, [(2500, 0.08), (3000, 0.08), (5000, 0.04), (9000, 0.01)]),
            (4, 400, [(2600, 0.08)])]):
        say(f"{name}: call {call_no + 1}, {n_short} short pairs, long {longs}")
        qs, ts, seeds = batch(seed, n_short, longs)
        pb = px.pack_batch(qs, ts, seeds)
        want = call(lib.phx_align_batch, ctx, pb, engine, c)
        got = call(lib.phx_align_batch_noalloc, ctx, pb, engine, c)
        same(got, want, f"{name} call {call_no + 1}", ctx)
        st = got[1]["status"]
        if engine == L.ENGINE_MYERS_GLOBAL:  # myers_edit_distance takes no seed
            assert (st == L.PHX_OK).all()
        else:
            assert (st == L.PHX_OK).sum() == len(qs) - 1 and st[len(qs) - 3] == L.PHX_INVALID_INPUT


def test_noalloc_refuses_what_does_not_fit():
    """A batch whose CIGAR slots alone exceed the scratch array, and a pair
    whose wavefront history does, fail with PHX_E_NOMEM and a clear reason;
    the next call still matches phx_align_batch."""
    import phraya_amd as px
    from phraya_amd import _lib as L
    ctx = px.default_context()
    lib = L.lib()
    c = px.AffineCosts(4, 6, 2)
    rng = random.Random(9)
    # 220k pairs x (150 + 170 + 1) run slots x 4 B > 64 MiB.
    q = rand_seq(rng, 150)
    t = q + rand_seq(rng, 20)
    big = px.pack_batch([q] * 220_000, [t] * 220_000)
    say("refusal: 220k pairs")
    rc, *_ = call(lib.phx_align_batch_noalloc, ctx, big, 1, c)
    assert rc == L.PHX_E_NOMEM
    assert "allocation-free scratch" in ctx.last_error()
    # One 60 kbp pair at 8%: its history needs far more than the array
    # (about 3.3 GiB by the memory model).
    t = rand_seq(rng, 66_000)
    huge = px.pack_batch([mutate(rng, t[:60_000], 0.08)], [t])
    say("refusal: 60 kbp pair")
    rc, *_ = call(lib.phx_align_batch_noalloc, ctx, huge, 1, c)
    assert rc == L.PHX_E_NOMEM
    assert "allocation-free scratch" in ctx.last_error()
    qs, ts, seeds = batch(5, 300, [(3000, 0.08)])
    pb = px.pack_batch(qs, ts, seeds)
    say("after the refusals: 300 short + 3 kbp")
    same(call(lib.phx_align_batch_noalloc, ctx, pb, 1, c), call(lib.phx_align_batch, ctx, pb, 1, c),
         "after refusals", ctx)
