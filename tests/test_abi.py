"""The C-ABI library loads and exports every entry point include/montmul_hip.h declares (no
device calls: this runs without a GPU)."""

import ctypes
import os
import re

import montmul_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "montmul_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = text[text.index("} mm_backend;"):]
    names = re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", body)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(montmul_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_params_without_device():
    k, ln = ctypes.c_uint32(), ctypes.c_uint32()
    assert montmul_amd.lib.mm_params_for_bits(100_000, 8, 1 << 24, ctypes.byref(k), ctypes.byref(ln)) == 0
    assert (k.value, ln.value) == (3126, 8192)
    assert montmul_amd.lib.mm_params_for_bits(1 << 30, 8, 1 << 24, ctypes.byref(k), ctypes.byref(ln)) != 0


def test_score_workload_is_seeded():
    m1, e1 = montmul_amd.score_workload(4096, 4, 512, 2026)
    m2, e2 = montmul_amd.score_workload(4096, 4, 512, 2026)
    assert m1 == m2 and e1 == e2
    for n, e in zip(m1, e1):
        assert n.bit_length() == 4096 and n & 1
        assert e == (n - 1) >> (4096 - 512)
