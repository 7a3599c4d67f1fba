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


@pytest.mark.parametrize("name", ["cpu_wrap", "gmp"])
def test_check_pow_range(name):
    common.pow_range_check(BACKENDS[name], 3000, 9, 60)


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


@pytest.mark.parametrize("name", ["cpu", "gmp"])
def test_is_one_many_on_oracle(name):
    """Backends without a device-side test: is_one_many reads each slot."""
    common.is_one_many_check(lambda: Cpu() if name == "cpu" else Gmp(), 1000, prime=2**607 - 1)


def test_mix_is_one_many_on_oracle():
    common.mix_is_one_many_check(lambda: Gmp(full_k=True))


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


def test_several_slots_carry_out_on_oracle():
    """Two slots carry out in one run: the error names the first and lists both."""
    common.carry_overflow_slots(lambda: Cpu(wraparound=True), 1000)
    common.carry_overflow_slots(lambda: Cpu(wraparound=True), 1000, bad=(0, 2, 3), slots=5)


class _CorruptFirstRun(mm.Backend):
    """The oracle's Cpu backend (wraparound) whose first run with every slot of `bad` active
    starts those slots from x = R - 1, so their carry checks fail in the same run."""

    def __init__(self, bad):
        import ctypes

        self.inner = Cpu(wraparound=True)
        self.bad, self.done, self.k = tuple(bad), False, {}
        t = mm.BackendTable()
        for name, _ in mm.BackendTable._fields_:
            setattr(t, name, getattr(self.inner.table, name))

        def create(selfp, setup, state):
            r = self.inner.table.create(selfp, setup, state)
            if r == 0:
                self.k[state[0]] = setup[0].k
            return r

        def run(state, bits):
            st = ctypes.c_void_p(state)
            if not self.done and all(self.inner.table.remaining(st, i) > 0 for i in self.bad):
                self.done = True
                k = self.k[state]
                buf = (ctypes.c_uint32 * k)(*([0xFFFFFFFF] * k))
                for i in self.bad:
                    self.inner.table.load(st, i, buf)
            return self.inner.table.run(st, bits)

        self._cbs = (mm._CREATE(create), mm._RUN(run))
        t.create, t.run = self._cbs
        self.table = t


def test_server_fails_every_slot_a_carry_hits():
    """Requests in slots 0 and 2 of one pool carry out in the same round: both fail, and the
    request in slot 1 still completes with the right answer and raw limbs."""
    import random
    import time

    rng = random.Random(15)
    moduli = [common.rand_odd(rng, 1000) for _ in range(3)]
    be = _CorruptFirstRun((0, 2))
    srv = mm.Server(base=3, pool_slots=4, target_ms=5, keep_raw=True, backend_factory=lambda: be)
    try:
        ids = [srv.submit(n) for n in moduli]
        got, t_end = {}, time.time() + 60
        while len(got) < 3 and time.time() < t_end:
            for r in srv.poll(16, 0.2):
                got[r["id"]] = r
        assert len(got) == 3
        assert be.done
        assert [got[i]["prp"] for i in ids] == [-1, int(pow(3, moduli[1] - 1, moduli[1]) == 1), -1]
        assert srv.stats()["failed"] == 2
        alone = ModBatch(Cpu(wraparound=True), [moduli[1]], 3)
        alone.load(0, 1)
        alone.set_exponent(0, moduli[1] - 1)
        alone.run(1 << 20)
        assert srv.raw(ids[1]) == alone.read_raw(0)
    finally:
        srv.close()


def test_server_staggered_on_oracle():
    """Requests arriving while others run (the oracle's Cpu backend per pool), with a given
    exponent for one of them: every result equals pow and its raw limbs a run alone."""
    import random

    rng = random.Random(16)
    plan = [(0.0, common.rand_odd(rng, 3000), None), (0.0, common.rand_odd(rng, 1000), common.rand_exponent(rng, 500)),
            (0.3, common.rand_odd(rng, 3000), None), (0.1, 2**607 - 1, None), (0.0, common.rand_odd(rng, 200), 0)]
    srv = mm.Server(base=3, pool_slots=4, target_ms=5, keep_raw=True, backend_factory=lambda: Cpu(wraparound=True))
    try:
        got = common.serve_staggered(srv, lambda: Cpu(wraparound=True), plan, timeout_s=120)
        assert got[3]["prp"] == 1 and got[4]["bits"] == 0 and got[4]["prp"] == 1
    finally:
        srv.close()


def _two_cpu():
    return mm.Multi([Cpu(wraparound=True), Cpu(wraparound=True)])


def test_multi_on_oracle():
    """Multi over two oracle backends (tests/gpu.rs:347-357 with the oracle for the GPU parts):
    shards of 1, 5 and 16 slots give the same raw limbs as one backend; chunked powers and mul,
    within and across shards; Fermat tests on tribonacci numbers."""
    for slots in (1, 5, 16):
        common.assert_same_raw(_two_cpu, lambda: Cpu(wraparound=True), 3000, slots, 40)
    common.powmod_and_mul(_two_cpu, 2000, 9)
    common.end_to_end(_two_cpu, [6, 7, 11, 87, 98, 215, 802])
    two = _two_cpu()
    assert len(two.parts) == 2 and mm.lib.mm_multi_parts(two.table) == 2


def test_multi_cross_shard_mul():
    """A mul whose operands sit on different shards runs on the host (multi.rs:128-138): the
    result is a b mod N, in [0, N)."""
    import random

    rng = random.Random(17)
    n = common.rand_odd(rng, 1500)
    b = ModBatch(mm.Multi([Cpu(wraparound=True)] * 3), [n] * 6, 3)
    xs = [rng.randrange(n) for _ in range(6)]
    for i, x in enumerate(xs):
        b.load(i, x)
    b.mul(5, 0, 3)  # shards 0, 1 and 2
    assert b.read(5) == xs[0] * xs[3] % n
    b.mul(1, 1, 1)  # within shard 0
    assert b.read(1) == xs[1] * xs[1] % n


def test_multi_carry_names_the_global_slot():
    """tests/gpu.rs:368-373: slot 1 is slot 0 of shard 1, and the error says "shard 1, slot 1";
    with two failing slots on different shards every one is listed (global numbers)."""
    import random

    from montmul_amd import CarryError

    rng = random.Random(12)
    b = ModBatch(_two_cpu(), [common.rand_odd(rng, 1000) for _ in range(2)], 4)
    for i in range(2):
        b.load(i, 1)
        b.set_exponent(i, 1)
    common.load_all_ones(b, 1)
    with pytest.raises(CarryError, match="shard 1, slot 1") as ei:
        b.run(1)
    assert ei.value.slot == 1
    common.carry_overflow_slots(_two_cpu, 1000, bad=(1, 3), slots=4)


def test_multi_rejects_parts_with_other_limits():
    with pytest.raises(MontmulError, match="part 1 has other limits than part 0"):
        mm.Multi([Cpu(wraparound=True), Cpu()])
