"""The .phraya and .phraya.queries writers (phx_phraya_*, phx_queries_*,
phx_cross_space_queries_* through the C ABI; host code, no GPU).

Pins: the payload bytes equal msgpack-python's encoding of the restated
structure (an independent MessagePack encoder: integer, float, string, array
and map forms at every length boundary); CoverageTrack's quantization and
runs equal the reference's own test vectors (phraya-core/src/types.rs:
1195-1332); the query files follow the reference's round-trip, order and
threshold tests (phraya-io/src/queries.rs:185-330); files decompress with
libzstd to the payload.  rmp-serde's choice of strings for unit enum
variants ("snp", "+") is not pinned by any fixture in the reference.
"""
import random

import numpy as np
import pytest

import phraya_amd as px
import phraya_amd._lib as L
from oracle import formats as F


def rec(alignment, position, typ, ref_base, alt_base=0, flags=0, ins=(0, 0), window=(0, 0, 0), off=0):
    r = np.zeros(1, L.VARIANT_DTYPE)[0]
    r["alignment"], r["position"], r["type"] = alignment, position, typ
    r["ref_base"], r["alt_base"], r["flags"] = ref_base, alt_base, flags
    r["ins_query_pos"], r["ins_len"] = ins
    r["window_start"], r["window_len"], r["window_offset"] = window
    r["coverage_off"] = off
    return r


def meta_dict(m: px.ReadMeta) -> dict:
    return dict(cigar=m.cigar, provenance=m.provenance, query=m.query, edit_distance=m.edit_distance,
                mapq=m.mapq, strand=m.strand, avg_base_quality=m.avg_base_quality, confidence=m.confidence,
                mate=m.mate, insert_stats=m.insert_stats)


def random_case(seed: int, n_reads=6, n_var=40):
    rng = random.Random(seed)
    reads = []
    for i in range(n_reads):
        q = bytes(rng.choice(b"ACGTN") for _ in range(rng.randint(0, 300)))
        mate = None
        if rng.random() < 0.4:
            mate = dict(mate_id=f"read{i}/2", insert_size=rng.choice([0, -450, 450, -40000, 2 ** 31 - 1, -2 ** 31]),
                        proper_pair=rng.random() < 0.5, is_first_in_pair=True, is_second_in_pair=False,
                        mate_mapped=rng.random() < 0.5)
        stats = rng.choice([None, (0, 0), (-1, 1), (-129, 3), (-40000, 9), (2 ** 40, 70000), (-2 ** 40, 2)])
        reads.append(px.ReadMeta(cigar=rng.choice(["", "150M", "3M1X146M", "7M2I3M" * 30]),
                                 provenance=rng.choice(["s:r", "sample1:read42", "x" * 40, "y" * 300]),
                                 query=q, edit_distance=rng.choice([0, 1, 127, 128, 70000]),
                                 mapq=rng.choice([0, 60, 255]), strand=rng.randint(0, 1),
                                 avg_base_quality=rng.choice([0.0, 31.5, 40.0]),
                                 confidence=rng.choice([1.0, 0.75, 0.1]), mate=mate, insert_stats=stats))
    wc, recs = [], []
    for _ in range(n_var):
        a = rng.randrange(n_reads)
        t = rng.randrange(3)
        wl = rng.choice([0, 1, 15, 16, 51, 101])
        off = len(wc)
        wc += [rng.choice([0, 1, 2, 127, 128, 255, 256, 65535, 65536, 2 ** 32 - 1]) for _ in range(wl)]
        qn = len(reads[a].query)
        ins = (0, 0)
        if t == 2 and qn:
            s = rng.randrange(qn)
            ins = (s, rng.randint(0, qn - s))
        recs.append(rec(a, rng.choice([0, 127, 128, 255, 256, 65535, 65536, 2 ** 32 - 1]), t,
                        rng.choice(b"ACGT.") if t != 2 else ord("."), rng.choice(b"ACGT") if t == 0 else 0,
                        rng.randrange(4), ins, (0, wl, rng.randint(0, max(0, wl - 1))), off))
    variants = np.array(recs, dtype=L.VARIANT_DTYPE) if recs else np.zeros(0, L.VARIANT_DTYPE)
    cov = np.array([rng.choice([0, 1, 2, 3, 7, 8, 13, 252, 253, 1000]) for _ in range(rng.randint(0, 500))],
                   np.uint32)
    return reads, variants, np.array(wc, np.uint32), cov


@pytest.mark.parametrize("seed", range(12))
def test_phraya_payload_matches_independent_encoder(seed):
    reads, v, wc, cov = random_case(seed)
    got = px.encode_phraya(5_000_000, "worker_0" * (seed % 5), "2026-01-01T00:00:00Z", v, wc, reads, cov)
    want = F.pack(F.phraya_file(5_000_000, "worker_0" * (seed % 5), "2026-01-01T00:00:00Z", v, wc,
                                [meta_dict(r) for r in reads], cov))
    assert got == want


def test_length_boundaries_of_every_form():
    # arrays of 65536+ (coverage runs, window values), maps of 16+ keys
    # (insertion alleles over 256 byte values), strings of 32 / 256 / 65536+.
    q = bytes(range(256)) * 2
    reads = [px.ReadMeta(cigar="M" * 70000, provenance="p" * 256, query=q, mate=dict(mate_id="m" * 32))]
    wc = np.arange(70000, dtype=np.uint32)
    v = np.array([rec(0, 9, 2, ord("."), ins=(0, 512), window=(0, 70000, 5)),
                  rec(0, 10, 2, ord("."), ins=(3, 20), window=(0, 16, 0)),
                  rec(0, 11, 0, ord("A"), ord("C"), window=(0, 65535, 1))], dtype=L.VARIANT_DTYPE)
    cov = np.tile(np.array([0, 5], np.uint32), 40000)    # 80000 runs
    got = px.encode_phraya(80000, "s" * 31, "t" * 65536, v, wc, reads, cov)
    assert got == F.pack(F.phraya_file(80000, "s" * 31, "t" * 65536, v, wc, [meta_dict(reads[0])], cov))


def test_coverage_track_reference_vectors():
    # types.rs:1195-1332: quantize() and CoverageTrack::new examples.
    for v, q in [(0, 0), (1, 0), (2, 0), (3, 5), (4, 5), (5, 5), (6, 5), (7, 5), (8, 10), (12, 10), (13, 15),
                 (17, 15), (15, 15), (252, 250), (253, 255), (1000, 255)]:
        assert F.quantize(v) == q
        doc = F.unpack(px.encode_phraya(1, "s", "t", np.zeros(0, L.VARIANT_DTYPE), [], [], [v]))
        assert doc[2] == [[[q, 1]], 1]
    cases = {(10, 10, 10, 10): [[10, 4]], (10, 5, 10, 5, 10): [[10, 1], [5, 1], [10, 1], [5, 1], [10, 1]],
             (0, 0, 5, 5, 0, 0): [[0, 2], [5, 2], [0, 2]], (10, 10, 5, 5): [[10, 2], [5, 2]], (15,): [[15, 1]],
             (): []}
    for cov, runs in cases.items():
        doc = F.unpack(px.encode_phraya(len(cov), "s", "t", np.zeros(0, L.VARIANT_DTYPE), [], [], list(cov)))
        assert doc[2] == [runs, len(cov)]
        assert doc[0] == [1, len(cov), "s", "t", 0] and doc[1] == []


def test_observation_fields_decode_as_upstream_builds_them():
    q = b"ACGTTGCAAC"
    reads = [px.ReadMeta("4M2D4M", "sample1:read42", q, edit_distance=2, mapq=60, strand=1,
                         avg_base_quality=30.0, confidence=0.8,
                         mate=dict(mate_id="read42/2", insert_size=-450, proper_pair=True, is_first_in_pair=True,
                                   mate_mapped=True), insert_stats=(450, 1))]
    wc = np.array([1, 1, 2, 1], np.uint32)
    v = np.array([rec(0, 100, 0, ord("A"), ord("G"), flags=F.IN_REPEAT, window=(98, 4, 2)),
                  rec(0, 104, 2, ord("."), ins=(4, 2), flags=F.IN_HOTSPOT, window=(102, 2, 2), off=2),
                  rec(0, 110, 1, ord("T"), ord("."))], dtype=L.VARIANT_DTYPE)
    doc = F.unpack(px.encode_phraya(1000, "worker_1", "now", v, wc, reads, np.ones(1000, np.uint32)))
    snv, ins, dele = doc[1]
    assert snv[:4] == [100, ord("A"), {ord("G"): 1}, 0.8]
    assert snv[4:10] == ["4M2D4M", 60, 2, [1, 1, 2, 1], 30.0, "sample1:read42"]
    assert snv[10:] == [True, "snp", 1.0, "-", 1, 1, 450, 1, 2, ["read42/2", True, -450, True, False, True]]
    assert ins[2] == {ord("G"): 1, ord("T"): 1} and ins[3] == 0.0 and ins[11] == "insertion" and ins[12] == 0.0
    assert dele[2] == {ord("."): 1} and dele[11] == "deletion" and dele[7] == []
    assert doc[2] == [[[0, 1000]], 1000]


def test_phraya_file_round_trip(tmp_path):
    if not L.lib().phx_zstd_available():
        pytest.skip("libzstd.so.1 not present")
    reads, v, wc, cov = random_case(99)
    path = tmp_path / "out.phraya"
    px.write_phraya(path, 1234, "worker_3", "ts", v, wc, reads, cov)
    payload = F.zstd_decompress(path.read_bytes())
    assert payload == px.encode_phraya(1234, "worker_3", "ts", v, wc, reads, cov)
    assert F.unpack(payload) == F.phraya_file(1234, "worker_3", "ts", v, wc, [meta_dict(r) for r in reads], cov)


def test_phraya_refuses_inconsistent_inputs():
    reads = [px.ReadMeta("1M", "p", b"A")]
    bad = [np.array([rec(1, 0, 0, 65, 67)], dtype=L.VARIANT_DTYPE),                  # alignment out of range
           np.array([rec(0, 0, 2, 46, ins=(0, 2))], dtype=L.VARIANT_DTYPE),           # insertion past the query
           np.array([rec(0, 0, 0, 65, 67, window=(0, 3, 0))], dtype=L.VARIANT_DTYPE),  # window past the values
           np.array([rec(0, 0, 3, 65, 67)], dtype=L.VARIANT_DTYPE)]                   # unknown type
    for v in bad:
        with pytest.raises(ValueError):
            px.encode_phraya(10, "s", "t", v, [1, 2], reads, [])


def test_queries_follow_reference_tests(tmp_path):
    # round_trip_single_query / multiple / empty / subthreshold dropping /
    # order independence (queries.rs:188-330).
    assert F.unpack(px.encode_queries({})) == {}
    one = {"query1": [(100, 0.98), (200, 0.95)]}
    assert F.unpack(px.encode_queries(one)) == {"query1": [[100, 0.98], [200, 0.95]]}
    multi = {"query1": [(100, 0.98), (200, 0.95)], "query2": [(50, 0.99)],
             "query3": [(150, 0.96), (250, 0.95), (350, 0.97)]}
    doc = F.unpack(px.encode_queries(multi))
    assert {k: len(v) for k, v in doc.items()} == {"query1": 2, "query2": 1, "query3": 3}
    drops = {"query_no_hits": [], "query_subthreshold": [(10, 0.40)], "kept": [(5, 0.951), (4, 0.2)]}
    assert F.unpack(px.encode_queries(drops)) == {"kept": [[5, 0.951]]}
    fwd = {f"q{i}": [(i * 7, 0.99), (i * 7 + 3, 0.96)] for i in range(50)}
    rev = {f"q{i}": [(i * 7 + 3, 0.96), (i * 7, 0.99)] for i in reversed(range(50))}
    assert px.encode_queries(fwd) == px.encode_queries(rev)
    big = {f"query_{i}": [(j * 11 % 1000, 0.95 + (j % 5) / 100) for j in range(i % 100 + 1)] for i in range(10000)}
    assert px.encode_queries(big) == F.pack(F.queries(big))
    ties = {"t": [(7, 0.96), (3, 0.99), (7, 0.99), (3, 0.97), (7, 0.96)], "é": [(1, 1.0)], "Z": [(2, 0.95)]}
    assert px.encode_queries(ties) == F.pack(F.queries(ties))
    if L.lib().phx_zstd_available():
        p = tmp_path / "out.phraya.queries"
        px.write_queries(p, multi)
        assert F.unpack(F.zstd_decompress(p.read_bytes())) == F.queries(multi)
    # NaN scores fail the threshold like upstream's filter; repeated ids are refused (a map key each).
    assert F.unpack(px.encode_queries({"a": [(1, float("nan")), (2, 1.0)]})) == {"a": [[2, 1.0]]}
    hits = (L.CQueryHits * 2)(L.CQueryHits(b"a", None, None, 0), L.CQueryHits(b"a", None, None, 0))
    assert L.lib().phx_queries_encode(hits, 2, None, 0, None) == L.PHX_E_ARG


def test_cross_space_queries(tmp_path):
    idx = {"r1": [("host", 10, 0.96), ("target", 5, 0.98), ("host", 3, 0.99), ("host", 10, 0.97)],
           "r2": [("host", 1, 0.5)], "r0": [("b", 1, 0.95), ("a", 2, 0.951)]}
    got = px.encode_cross_space_queries(idx)
    assert got == F.pack(F.cross_space_queries(idx))
    assert F.unpack(got) == {"r0": [["a", 2, 0.951], ["b", 1, 0.95]],
                             "r1": [["host", 3, 0.99], ["host", 10, 0.97], ["host", 10, 0.96],
                                    ["target", 5, 0.98]]}
    if L.lib().phx_zstd_available():
        p = tmp_path / "out.cross.queries"
        px.write_cross_space_queries(p, idx)
        assert F.zstd_decompress(p.read_bytes()) == got
