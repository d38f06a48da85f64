"""The variant-site oracle (orc_extract_variants) against upstream's own
cases and hand-derived expectations.

Upstream's tests (phraya-align/src/executor.rs) run these scenarios through
align_task; here the CIGAR each alignment yields is given explicitly, so the
expectations follow from extract_variants_from_cigar's rules
(executor.rs:814-1044), compute_windowed_coverage (:1071-1090) and the
interval lookups (:795-811) alone."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

OPS = {"M": 0, "X": 1, "I": 2, "D": 3}


def runs_of(cigar: str) -> list[int]:
    out, num = [], ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
        else:
            out.append((int(num) << 4) | OPS[ch])
            num = ""
    return out


def extract(items, target: bytes, repeats=None, hotspots=None, radius=25):
    """items: (cigar, target_start, query, spans) per alignment."""
    runs, run_off, ts, qs, q_off, sp, sp_off = [], [0], [], b"", [0], [], [0]
    for cigar, start, q, spans in items:
        runs += runs_of(cigar)
        run_off.append(len(runs))
        ts.append(start)
        qs += q
        q_off.append(len(qs))
        sp += [list(s) for s in spans]
        sp_off.append(len(sp))
    return O.extract_variants(np.array(runs, np.uint32), run_off, ts, np.frombuffer(qs, np.uint8),
                              q_off, target, np.array(sp, np.uint64).reshape(-1, 2), sp_off,
                              None if repeats is None else np.array(repeats, np.uint64),
                              None if hotspots is None else np.array(hotspots, np.uint32), radius)


def test_single_snp():
    # executor.rs:1512-1540: query ACTACGT vs target ACCACGT, one SNP at 2.
    v, off, cov = extract([("2M1X4M", 0, b"ACTACGT", [(0, 7)])], b"ACCACGT", radius=50)
    assert len(v) == 1 and list(off) == [0, 1]
    r = v[0]
    assert (r["position"], r["type"], chr(r["ref_base"]), chr(r["alt_base"])) == (2, 0, "C", "T")
    # Window [0, min(2 + 51, 7)) = [0, 7), the variant at offset 2, coverage 1.
    assert (r["window_start"], r["window_len"], r["window_offset"]) == (0, 7, 2)
    assert list(cov) == [1] * 7


def test_deletion_and_insertion():
    # executor.rs:1625-1700: ACGACGT vs ACGTACGT lacks the T at 3 (an I run:
    # target-only bases); ACGTTACGT vs ACGTACGT has an extra T (a D run).
    v, _, _ = extract([("3M1I4M", 0, b"ACGACGT", [(0, 8)])], b"ACGTACGT")
    assert len(v) == 1
    assert (v[0]["type"], v[0]["position"], chr(v[0]["ref_base"]), chr(v[0]["alt_base"])) == (1, 3, "T", ".")
    v, _, _ = extract([("4M1D4M", 0, b"ACGTTACGT", [(0, 8)])], b"ACGTACGT")
    assert len(v) == 1
    r = v[0]
    assert (r["type"], r["position"], chr(r["ref_base"])) == (2, 4, ".")
    assert (r["ins_query_pos"], r["ins_len"]) == (4, 1)  # the inserted base is query[4] = T


def test_tandem_repeat_flags():
    # executor.rs:1587-1623: SNPs at 2 (unique) and 15 (inside the repeat
    # [10, 32)); only the latter is in a repeat.
    target = b"TTAACCGGTA" + b"ATATATATATATATATATATAT" + b"CGTACCGATT"
    q = bytearray(target)
    q[2] = ord("C") if q[2] == ord("G") else ord("G")
    q[15] = ord("T") if q[15] == ord("A") else ord("A")
    v, _, _ = extract([("2M1X12M1X26M", 0, bytes(q), [(0, 42)])], target, repeats=[(10, 32)])
    flags = {int(r["position"]): int(r["flags"]) for r in v}
    assert flags == {2: 0, 15: 1}


def test_interval_lookups_match_a_linear_scan():
    # executor.rs:1132-1169: adjacent regions [10, 20), [20, 30), gap, [50, 55);
    # every position 0..59 as an SNV (query all 'T', target all 'A').
    regions = [(10, 20), (20, 30), (50, 55)]
    v, _, _ = extract([("60X", 0, b"T" * 60, [(0, 60)])], b"A" * 60, repeats=regions,
                      hotspots=regions)
    assert len(v) == 60
    for r in v:
        p = int(r["position"])
        inside = any(s <= p < e for s, e in regions)
        assert int(r["flags"]) == (3 if inside else 0), p
    v, _, _ = extract([("5X", 0, b"TTTTT", [])], b"AAAAA")  # no intervals at all
    assert all(int(r["flags"]) == 0 for r in v)


def test_tail_gaps_and_bounds_are_skipped():
    # An I run after the query is used up (q_pos == |query|) is no deletion,
    # and a D run past the target (t_pos == |target|) no insertion
    # (executor.rs:913-916, :976-978); X bases past either end are dropped.
    v, _, _ = extract([("4M4I", 0, b"ACGT", [(0, 8)])], b"ACGTAAAA")
    assert len(v) == 0
    v, _, _ = extract([("4M4D", 0, b"ACGTAAAA", [(0, 4)])], b"ACGT")
    assert len(v) == 0
    # Alignment starting at 6 of an 8-base target: X bases at 6, 7 are kept,
    # 8 and 9 lie past the target.
    v, _, _ = extract([("4X", 6, b"TTTT", [(6, 8)])], b"AAAAAAAA")
    assert [int(r["position"]) for r in v] == [6, 7]
    # An insertion run longer than the rest of the query is cut at its end.
    v, _, _ = extract([("2M5D", 0, b"ACGT", [(0, 4)])], b"ACGGG")
    assert (int(v[0]["ins_query_pos"]), int(v[0]["ins_len"])) == (2, 2)


def test_coverage_windows():
    # executor.rs:1542-1585: a 100-base alignment on a 200-base target; the
    # window around a variant counts the read's alignments over each position
    # (0 outside them).  Two windows: the alignment [0, 100) and an
    # alternative [90, 150).
    target = b"A" * 200
    q = bytearray(b"A" * 100)
    q[95] = ord("T")
    v, _, cov = extract([("95M1X4M", 0, bytes(q), [(0, 100), (90, 150)])], target, radius=50)
    r = v[0]
    assert (r["position"], r["window_start"], r["window_len"], r["window_offset"]) == (95, 45, 101, 50)
    expect = [(1 if p < 100 else 0) + (1 if 90 <= p < 150 else 0) for p in range(45, 146)]
    assert list(cov[r["coverage_off"]:r["coverage_off"] + r["window_len"]]) == expect
    # Windows clip at the target's ends; spans are clipped to the target and
    # empty ones dropped (compute_windowed_coverage).
    v, _, cov = extract([("1X", 199, b"T", [(199, 260), (300, 400)])], target, radius=5)
    r = v[0]
    assert (r["window_start"], r["window_len"], r["window_offset"]) == (194, 6, 5)
    assert list(cov) == [0, 0, 0, 0, 0, 1]


def test_order_and_offsets():
    # Records in alignment order, then CIGAR order; var_off delimits them; an
    # alignment with no variant has an empty range.
    items = [("1X1M1X", 0, b"TAT", [(0, 3)]), ("3M", 0, b"AAA", [(0, 3)]),
             ("1M1D1I1M", 0, b"AGA", [(0, 3)])]
    v, off, _ = extract(items, b"AAA")
    assert list(off) == [0, 2, 2, 4]
    # Third alignment: the D run (query G) at t_pos 1, then the I run (target
    # A at 1) at t_pos 1 as well.
    assert [(int(r["alignment"]), int(r["position"]), int(r["type"])) for r in v] == \
        [(0, 0, 0), (0, 2, 0), (2, 1, 2), (2, 1, 1)]
