"""The command line on the GPU (main.rs:124-189, 300-343): bench and prp print their key=value
lines, fermat checks its planted primes and a sample against GMP powm."""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cli(*args):
    r = subprocess.run([sys.executable, "-m", "montmul_amd", *args], cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    return dict(kv.split("=", 1) for kv in r.stdout.strip().splitlines()[-1].split() if "=" in kv)


def test_cli_bench(tmp_path):
    csv = str(tmp_path / "b.csv")
    f = cli("bench", "--bits", "5000", "--slots", "16", "--steps", "200", "--csv", csv)
    assert f["L"] == "512" and float(f["bitsteps_per_s"]) > 0
    assert open(csv).read().startswith("backend,device,bits,L")


def test_cli_prp():
    r = subprocess.run([sys.executable, "-m", "montmul_amd", "prp", "--index", "87"], cwd=ROOT,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0 and "prp=true" in r.stdout, r.stdout


def test_cli_fermat_planted(tmp_path):
    primes = tmp_path / "p.txt"
    primes.write_text("%x\n%x\n" % (2**4253 - 1, 2**4423 - 1))
    f = cli("fermat", "--bits", "4253", "--slots", "32", "--planted", str(primes), "--planted-slots", "8", "--check", "4")
    assert f["planted"] == "8" and f["planted_not_prime"] == "0" and f["powm_mismatch"] == "0"


def test_cli_fermat_20000_bits():
    """The full-test workload at a reduced slot count: 20,000-bit moduli (L = 2,048), e = N - 1,
    half the slots cycling through the planted primes of tests/golden/planted_20000bit.txt, the
    others from GMP's generator seeded 7; planted primes come out prime, sampled others match
    GMP powm."""
    planted = os.path.join(ROOT, "tests", "golden", "planted_20000bit.txt")
    f = cli("fermat", "--bits", "20000", "--slots", "256", "--seed", "7", "--planted", planted,
            "--planted-slots", "128", "--check", "4")
    assert f["L"] == "2048" and f["planted"] == "128" and f["planted_not_prime"] == "0" and f["powm_mismatch"] == "0"
