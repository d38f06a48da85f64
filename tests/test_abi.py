"""CPU checks of the drop-in boundary: the library loads, exports every
symbol include/phraya_hip.h declares, and its host-only helpers behave."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from phraya_amd import _lib as L
from phraya_amd import format_cigar, pack_batch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "phraya_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_what_python_binds():
    assert declared_symbols() == sorted(L.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.phx_version().decode().startswith("phraya_hip")


def test_last_impl_is_empty_before_any_per_pair_call():
    """phx_last_impl is per thread (as upstream's LAST_IMPL): a thread that
    has made no per-pair call reads ""."""
    import threading
    got = []
    t = threading.Thread(target=lambda: got.append(L.lib().phx_last_impl()))
    t.start()
    t.join()
    assert got == [b""]


def test_format_cigar_host_helper():
    lib = L.lib()
    runs = np.array([(150 << 4) | 0, (20 << 4) | 2, (1 << 4) | 1, (3 << 4) | 3], dtype=np.uint32)
    buf = ctypes.create_string_buffer(64)
    n = lib.phx_format_cigar(runs.ctypes.data, len(runs), buf, 64)
    assert buf.value.decode() == "150M20I1X3D" and n == len("150M20I1X3D")
    assert format_cigar(runs) == "150M20I1X3D"
    # truncation keeps the NUL and reports the full length
    small = ctypes.create_string_buffer(4)
    assert lib.phx_format_cigar(runs.ctypes.data, len(runs), small, 4) == 11
    assert small.value.decode() == "150"


def test_cigar_bound_counts_worst_case_runs():
    lib = L.lib()
    pb = pack_batch([b"ACGT", b"", b"AC"], [b"ACG", b"TT", b""])
    from phraya_amd import _cbatch, AffineCosts
    cb = _cbatch(pb, L.ENGINE_AFFINE, AffineCosts(), None)
    assert lib.phx_cigar_bound(ctypes.byref(cb)) == (4 + 3 + 1) + (0 + 2 + 1) + (2 + 0 + 1)
    cb = _cbatch(pb, L.ENGINE_MYERS_DIST, AffineCosts(), None)
    assert lib.phx_cigar_bound(ctypes.byref(cb)) == 0


def test_format_cigars_batch_matches_per_pair():
    lib = L.lib()
    rng = np.random.default_rng(3)
    runs_l, off = [], [0]
    for _ in range(50):
        k = int(rng.integers(0, 6))
        for _ in range(k):
            runs_l.append((int(rng.integers(1, 100000)) << 4) | int(rng.integers(0, 4)))
        off.append(len(runs_l))
    runs = np.array(runs_l or [0], np.uint32)
    run_off = np.array(off, np.uint64)
    text_off = np.zeros(51, np.uint64)
    total = lib.phx_format_cigars(runs.ctypes.data, run_off.ctypes.data, 50, None, 0, text_off.ctypes.data)
    buf = ctypes.create_string_buffer(int(total))
    assert lib.phx_format_cigars(runs.ctypes.data, run_off.ctypes.data, 50, buf, total,
                                 text_off.ctypes.data) == total
    text = buf.raw.decode()
    for i in range(50):
        want = format_cigar(runs[off[i]:off[i + 1]])
        assert text[int(text_off[i]):int(text_off[i + 1])] == want


def test_host_sequence_helpers():
    """The library's word-at-a-time ACGT test and 2-bit packer against
    per-byte references (tests/native/host_helpers_test.cpp)."""
    import subprocess
    here = os.path.join(ROOT, "tests", "native")
    subprocess.run(["make", "-s", "-C", here, "host_helpers_test"], check=True)
    out = subprocess.run([os.path.join(here, "host_helpers_test")], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
