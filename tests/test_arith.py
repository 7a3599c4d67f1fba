"""Host check of the register transforms' lazy GF(M31^2) / GF(M61^2) arithmetic.

tests/arith/v2_arith_check.cpp includes the device header (montmul_amd/csrc/mm_v2.h), whose
arithmetic and butterflies are host+device functions, and compares every operation with exact
residues at the top of each lazy range, and the radix-8/4 butterflies with a direct DFT over
omega_8 = 2^h (1 + i) (the roots of montmul/src/field.rs:237-267). Runs on the CPU.
"""

import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "arith", "v2_arith_check.cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("arith") / "v2_arith_check")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip", SRC, "-o", exe],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return exe


def test_lazy_arithmetic_and_butterflies(checker):
    r = subprocess.run([checker, "200000"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0 and "OK" in r.stdout, r.stdout
