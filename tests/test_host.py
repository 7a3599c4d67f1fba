"""The host API (ModBatch, fermat, Params; montmul/src/batch.rs) on CPU, driven over the
oracle's backend tables: the shared cases of montmul/tests/cpu.rs and tests/gmp.rs at sizes
that finish in seconds."""

import pytest

import common
import montmul_amd as mm
from montmul_amd import MontmulError, ModBatch, fermat
from oracle_backends import Cpu, Gmp

BACKENDS = {
    "cpu": Cpu,
    "cpu_wrap": lambda: Cpu(wraparound=True),
    "gmp": Gmp,
}


@pytest.mark.parametrize("name", list(BACKENDS))
def test_random_modmul(name):
    make = BACKENDS[name]
    for bits, cases in [(64, 60), (100, 60), (1000, 30), (10_000, 6)]:
        for base in (3, 4):
            common.random_modmul(make, bits, cases, bits, base)


@pytest.mark.parametrize("name", list(BACKENDS))
def test_edges(name):
    for base in (3, 4):
        common.edges(BACKENDS[name], base, max_len_log2=11)


@pytest.mark.parametrize("name", list(BACKENDS))
def test_independence(name):
    common.batch_independence(BACKENDS[name], 3000, [1, 7, 16])


@pytest.mark.parametrize("name", list(BACKENDS))
def test_powmod_mul(name):
    common.powmod_and_mul(BACKENDS[name], 2000, 9)


@pytest.mark.parametrize("name", list(BACKENDS))
def test_api(name):
    common.api_contract(BACKENDS[name])


@pytest.mark.parametrize("name", list(BACKENDS))
def test_carries(name):
    common.carry_stress(BACKENDS[name], 1 << 11)


@pytest.mark.parametrize("name", list(BACKENDS))
def test_fermat_tribonacci(name):
    common.end_to_end(BACKENDS[name], common.A092835_SMALL + [100, 200, 1001])


def test_known_prps():
    common.known_prps(Gmp, common.A092835_SMALL, [8, 9, 10, 12, 13, 20, 86, 88, 803])


def test_is_one_accepts_upper_lazy_representative():  # tests/cpu.rs:83-95
    n = common.rand_odd(__import__("random").Random(1), 1000)
    b = ModBatch(Cpu(), [n], 3)
    r = (1 << (32 * b.params().k)) % n
    b.load_raw(0, r + n)
    assert b.is_one(0)
    b.load_raw(0, r)
    assert b.is_one(0)
    b.load_raw(0, r + 1)
    assert not b.is_one(0)


def test_load_raw_rejects_out_of_range():  # tests/cpu.rs:97-103
    n = 1_000_003
    b = ModBatch(Cpu(), [n], 3)
    with pytest.raises(MontmulError, match=r"raw value outside \[0, 2N\)"):
        b.load_raw(0, 2 * n)


def test_fermat_chunk_zero_is_error():  # tests/cpu.rs:105-109
    with pytest.raises(MontmulError):
        fermat(Cpu(), [1_000_003], 3, 0)


def test_long_run():
    common.long_run(lambda: Cpu(wraparound=True), 3000, 4, 300)


def test_params_rules():
    from montmul_amd import Params
    lengths = Cpu().supported_lengths()
    assert lengths[0] == 8 and lengths[-1] == 1 << 24
    assert Params.for_bits(100_000, lengths) == Params(3126, 8192)
    assert Params.for_bits(20_000, lengths) == Params(626, 2048)
    assert Params.for_bits(4096, lengths) == Params(129, 512)
    b = ModBatch(Cpu(wraparound=True), [(1 << 99_999) | 1], 3)
    assert b.params() == Params(4096, 8192)


def test_server_on_oracle():
    """The request server over length pools, with the oracle's Cpu backend (wraparound) for
    every pool: mixed lengths and a pool smaller than the load, so slots are reused; every result
    equals base^(N-1) = 1 mod N and its raw limbs equal a run alone."""
    import random

    from oracle_backends import Cpu

    rng = random.Random(12)
    moduli = [common.rand_odd(rng, b) for b in (60, 64, 100, 500, 1000, 1000, 2000, 3000)]
    moduli += [2**127 - 1, 2**521 - 1, 2**607 - 1, common.tribonacci(87)]  # primes
    srv = mm.Server(base=3, pool_slots=2, target_ms=5, keep_raw=True, backend_factory=lambda: Cpu(wraparound=True))
    try:
        common.serve_mix(srv, lambda: Cpu(wraparound=True), moduli)
        st = srv.stats()
        assert st["completed"] == len(moduli) and st["failed"] == 0 and st["pools"] >= 3
    finally:
        srv.close()


def test_server_memory_budget_on_oracle():
    """Admission by device memory: a budget with room for one slot of the longer length only.
    Pools are cut to what fits, the other length waits, and an idle pool frees its batch for it;
    every result is still right, and the reserved bytes never exceed the budget at the end."""
    import random

    from oracle_backends import Cpu

    rng = random.Random(14)
    long_len, short_len = 256, 64
    budget = mm.Server.slot_bytes(long_len) + mm.Server.slot_bytes(short_len) // 2
    moduli = [common.rand_odd(rng, 3000) for _ in range(3)] + [common.rand_odd(rng, 1000) for _ in range(4)]
    moduli += [2**607 - 1, 2**1279 - 1]
    srv = mm.Server(base=3, pool_slots=2, target_ms=5, keep_raw=True, hbm_bytes=budget,
                    backend_factory=lambda: Cpu(wraparound=True))
    try:
        common.serve_mix(srv, lambda: Cpu(wraparound=True), moduli)
        st = srv.stats()
        assert st["completed"] == len(moduli) and st["failed"] == 0
        assert st["releases"] >= 1 and st["hbm_used"] <= budget
    finally:
        srv.close()
    # a length whose single slot exceeds the budget is refused, not queued forever
    srv = mm.Server(base=3, pool_slots=2, hbm_bytes=mm.Server.slot_bytes(short_len),
                    backend_factory=lambda: Cpu(wraparound=True))
    try:
        i = srv.submit(moduli[0])
        got = []
        for _ in range(100):
            got += srv.poll(8, 0.1)
            if got:
                break
        assert got and got[0]["id"] == i and got[0]["prp"] == -1
    finally:
        srv.close()


def test_cli_arguments():
    """The command line's option rules (main.rs:32-66): unknown, repeated and value-less keys
    are errors; values parse as numbers where needed."""
    from montmul_amd.__main__ import Args

    a = Args("bench", ["--bits", "1000", "--slots", "4", "--steps", "10", "--bare"])
    assert a.num("bits") == 1000 and a.flag("bare") and not a.flag("global")
    for bad in (["--bits", "1", "--bits", "2"], ["--nope", "1"], ["--bits"], ["--bits", "--slots"], ["bits"]):
        with pytest.raises(SystemExit):
            Args("bench", bad)
    with pytest.raises(SystemExit):
        Args("prp", ["--index", "x"]).num("index")
