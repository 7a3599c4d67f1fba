"""The request server on the GPU: Fermat PRP requests of mixed lengths over HIP pools (one
context and stream per pool), against base^(N-1) mod N and against each request run alone."""

import random

import pytest

import common
import montmul_amd as mm

pytestmark = pytest.mark.gpu


def test_server_mixed_lengths(hip):
    rng = random.Random(31)
    moduli = [common.rand_odd(rng, b) for b in (60, 500, 1000, 3000, 5000, 5000, 20000, 20000, 50000)]
    moduli += [2**4253 - 1, 2**9689 - 1, 2**19937 - 1, common.tribonacci(4202)]  # primes
    srv = mm.Server(base=3, pool_slots=3, target_ms=10, keep_raw=True)
    try:
        got = common.serve_mix(srv, lambda: hip, moduli, timeout_s=300)
        st = srv.stats()
        assert st["failed"] == 0 and st["completed"] == len(moduli)
        assert sorted({r["len"] for r in got.values()}) == [8, 32, 64, 256, 512, 1024, 2048, 4096]
    finally:
        srv.close()


def test_server_memory_budget(hip):
    """Admission by device memory on HIP pools: a budget one byte short of a slot of the longest
    length (L = 512) plus one of the shortest (L = 32), so in any order of pool creation some
    length waits and an idle pool frees its batch for it; every result is right and equal to the
    test run alone."""
    rng = random.Random(32)
    moduli = [common.rand_odd(rng, b) for b in (6000, 6000, 1000, 1000, 500, 500)]
    moduli += [2**4423 - 1, 2**607 - 1]  # primes
    budget = mm.Server.slot_bytes(512) + mm.Server.slot_bytes(32) - 1  # a deferral in any order
    srv = mm.Server(base=3, pool_slots=4, target_ms=10, keep_raw=True, hbm_bytes=budget)
    try:
        common.serve_mix(srv, lambda: hip, moduli, timeout_s=300)
        st = srv.stats()
        assert st["failed"] == 0 and st["completed"] == len(moduli)
        assert st["releases"] >= 1 and st["hbm_used"] <= budget
    finally:
        srv.close()


def test_server_staggered_up_to_l32768(hip):
    """Config 5's upper lengths with arrivals while other requests run: requests at L = 8,192
    (T(96879), an 85,169-bit prime, as a full Fermat test, and a 100,000-bit modulus admitted
    into that pool's running batch), L = 16,384 (150,000 and 200,000 bits) and L = 32,768
    (270,000 bits), plus a 20,000-bit planted prime. The long lengths run base^e mod N for short
    exponents (mm_server_submit_exp) so the test stays in seconds. Each result is checked against
    pow (prp and the value raw R^-1 mod N) and its raw limbs against the request run alone."""
    import os

    rng = random.Random(33)
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "planted_20000bit.txt")
    with open(golden) as f:
        planted = int(f.readline().strip(), 16)
    plan = [
        (0.0, common.tribonacci(96879), None),
        (0.0, common.rand_odd(rng, 270_000), common.rand_exponent(rng, 3000)),
        (1.0, common.rand_odd(rng, 100_000), common.rand_exponent(rng, 4000)),
        (0.0, common.rand_odd(rng, 150_000), common.rand_exponent(rng, 4000)),
        (1.0, common.rand_odd(rng, 200_000), common.rand_exponent(rng, 2000)),
        (0.0, planted, None),
        (0.0, common.rand_odd(rng, 5000), None),
    ]
    srv = mm.Server(base=3, pool_slots=4, target_ms=20, keep_raw=True)
    try:
        got = common.serve_staggered(srv, lambda: hip, plan, timeout_s=240)
        assert sorted({r["len"] for r in got}) == [512, 2048, 8192, 16384, 32768]
        assert got[0]["prp"] == 1 and got[5]["prp"] == 1  # the primes
        st = srv.stats()
        assert st["failed"] == 0 and st["completed"] == len(plan)
    finally:
        srv.close()
