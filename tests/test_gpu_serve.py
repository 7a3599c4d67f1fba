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
