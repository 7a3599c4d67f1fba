"""GPU parity: the HIP backend against the oracle and the closed form, through the C ABI.

Mirrors montmul/tests/gpu.rs at sizes the oracle finishes in seconds. Every comparison is bit
for bit on raw Montgomery limbs (all integer arithmetic: no tolerance applies).
"""

import json
import os
import random

import pytest

import common
import montmul_amd as mm
from montmul_amd import CarryError, Hip, ModBatch, MontmulError
from oracle_backends import Cpu, Gmp

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


def test_device(hip):
    say("device", hip.device_name(), "CUs", hip.cu_count(), "LDS", hip.max_lds())
    assert Hip.device_count() >= 1
    assert hip.cu_count() > 0 and hip.max_lds() >= 65536
    assert hip.full_k()


@pytest.mark.parametrize("mode", ["lds", "global"])
def test_golden_steps(hip, hip_global, mode):
    """Every golden step vector (tests/golden/steps.json), one step each."""
    be = hip if mode == "lds" else hip_global
    with open(os.path.join(GOLD, "steps.json")) as f:
        steps = json.load(f)
    groups = {}
    for v in steps:
        groups.setdefault((v["L"], v["c"]), []).append(v)
    for (L, c), vs in sorted(groups.items()):
        say("golden", mode, "L", L, "c", c, "vectors", len(vs))
        b = ModBatch(be, [int(v["n"], 16) for v in vs], max(c, 1))
        assert b.params().k == vs[0]["k"] and b.params().len == L
        for i, v in enumerate(vs):
            b.load_raw(i, int(v["x"], 16))
            b.set_exponent(i, 1)
        b.run(1)
        for i, v in enumerate(vs):
            assert b.read_raw(i) == int(v["out"], 16), f"L {L} c {c} bits {v['bits']} vector {i}"


def test_golden_chains(hip):
    with open(os.path.join(GOLD, "chains.json")) as f:
        chains = json.load(f)
    for ch in chains:
        say("chain bits", ch["bits"])
        b = ModBatch(hip, [int(ch["n"], 16)], 3)
        b.load(0, 1)
        b.set_exponent(0, int(ch["e"], 16))
        for raw in ch["raw"]:
            b.run(1)
            assert b.read_raw(0) == int(raw, 16)
        assert b.read(0) == int(ch["pow"], 16)


def test_gpu_matches_cpu_raw(hip):
    """GPU (wraparound, k = L/2) against the Cpu restatement (tests/gpu.rs:169-177, 315-330)."""
    for bits, slots, steps in [(64, 8, 60), (200, 8, 60), (1000, 8, 60), (5000, 8, 40), (87_914, 4, 12),
                               (100_000, 4, 12)]:
        say("same raw", bits, slots, steps)
        common.assert_same_raw(lambda: hip, lambda: Cpu(wraparound=True), bits, slots, steps)


def test_global_mode_matches_lds_mode(hip, hip_global):
    """tests/gpu.rs:155-167: same raw limbs in both modes, and the modes are as planned."""
    n = common.rand_odd(random.Random(11), 87_914)
    assert not Hip.plan_of(ModBatch(hip, [n], 3))["global"]
    assert Hip.plan_of(ModBatch(hip_global, [n], 3))["global"]
    say("global == lds at 87,914 bits")
    common.assert_same_raw(lambda: hip, lambda: hip_global, 87_914, 16, 60)
    common.assert_same_raw(lambda: hip, lambda: hip_global, 200, 8, 60)


def test_large_lengths_global(hip):
    """L = 16,384 and 32,768 do not fit LDS and run in global mode (tests/gpu.rs:325-329)."""
    for bits in (200_000, 400_000):
        b = ModBatch(hip, [common.rand_odd(random.Random(bits), bits)], 3)
        assert Hip.plan_of(b)["global"]
        say("large L", b.params().len)
        common.assert_same_raw(lambda: hip, lambda: Cpu(wraparound=True), bits, 2, 6)


def test_radix4_global_tiles_every_length(monkeypatch, hip_global):
    """The radix-4 kernel in global mode (forced with MONTMUL_KERNEL=v1) runs its transforms'
    narrow passes in an LDS tile of up to 2,048 elements (lib.rs:455-511): by default from
    L = 65,536 on, and at any length with tile = 2,048 (one tile for the whole vector up to
    n = 2,048, 2 to 32 tiles above). L = 8 .. 131,072 (odd and even log2 n), tiles forced: raw
    limbs equal the oracle's Gmp closed form with k = L/2 after 12 steps on 2 slots."""
    monkeypatch.setenv("MONTMUL_KERNEL", "v1")
    tiled = Hip(force_global=True, tile=2048)
    try:
        for bits in (64, 200, 500, 1000, 2000, 4000, 8000, 12_000, 20_000, 40_000, 90_000, 150_000, 300_000, 1_000_000,
                     2_000_000):
            moduli = [common.rand_odd(random.Random(bits), bits)]
            b, d = ModBatch(tiled, moduli, 3), ModBatch(hip_global, moduli, 3)
            p, n = Hip.plan_of(b), b.params().len // 2
            assert p["kernel"] == 4 and p["global"] and p["tile"] == min(n, 2048), (bits, p)
            assert Hip.plan_of(d)["tile"] == (2048 if n >= 32768 else 0), (bits, Hip.plan_of(d))
            b.free()
            d.free()
            say("radix-4 tiles, L =", 2 * n)
            common.assert_same_raw(lambda: tiled, lambda: Gmp(full_k=True), bits, 2, 12, seed=bits)
    finally:
        tiled.close()


def test_random_modmul(hip):
    for bits, cases in [(64, 100), (1000, 60), (10_000, 16), (100_000, 4)]:
        for base in (3, 4):
            say("random_modmul", bits, base)
            common.random_modmul(lambda: hip, bits, cases, bits, base)


def test_edges(hip):
    for base in (3, 4):
        say("edges", base)
        common.edges(lambda: hip, base, max_len_log2=15)


def test_independence(hip):
    common.batch_independence(lambda: hip, 3000, [1, 7, 64, 1024])


def test_powmod_mul(hip):
    common.powmod_and_mul(lambda: hip, 2000, 9)
    common.powmod_and_mul(lambda: hip, 87_914, 2)


def test_api(hip):
    common.api_contract(lambda: hip)


def test_carries(hip):
    for length in (512, 2048, 8192, 65_536):
        say("carry stress L", length)
        common.carry_stress(lambda: hip, length)


def test_fermat_tribonacci(hip):
    common.end_to_end(lambda: hip, [6, 7, 11, 87, 98, 215, 802, 1001, 2000, 4000])


def test_tribonacci_prps(hip):
    common.known_prps(lambda: hip, [4202, 18699], [4203, 18700])


def test_tribonacci_prp_96879(hip):
    """T(96879), 85,169 bits, is a known PRP (tests/gpu.rs:148-153, where it is ignored for its
    run time): one slot, 85,168 sequential steps at L = 8,192."""
    t = common.tribonacci(96879)
    say("T(96879) bits", t.bit_length())
    assert t.bit_length() == 85_169
    common.known_prps(lambda: hip, [96879], [])


def carry_overflow(make, bits):
    """x = R - 1 is outside [0, 2N); T = 4 (R - 1)^2 carries out of 2k limbs and the error names
    slot 1 (tests/gpu.rs:269-305)."""
    rng = random.Random(12)
    moduli = [common.rand_odd(rng, bits), common.rand_odd(rng, bits)]
    b = ModBatch(make(), moduli, 4)
    k = b.params().k
    for i in range(2):
        b.load(i, 1)
        b.set_exponent(i, 1)
    import ctypes
    buf = (ctypes.c_uint32 * k)(*([0xFFFFFFFF] * k))
    assert mm.lib.mm_load(ctypes.c_void_p(b.state), 1, buf) == 0
    with pytest.raises(CarryError, match="slot 1") as ei:
        b.run(1)
    assert ei.value.slot == 1


def test_carry_overflow_names_the_slot(hip, hip_global):
    carry_overflow(lambda: hip, 1000)
    carry_overflow(lambda: hip, 87_914)
    carry_overflow(lambda: hip_global, 87_914)


def test_carry_overflow_in_several_slots(hip, hip_global):
    """Two (then three) slots carry out in one run: the error names the first, failed_slots()
    lists every one, and all error words are cleared for the next run."""
    common.carry_overflow_slots(lambda: hip, 1000)
    common.carry_overflow_slots(lambda: hip_global, 100_000)
    common.carry_overflow_slots(lambda: hip_global, 100_000, bad=(0, 2, 3), slots=5)


def test_without_wraparound():
    """tests/gpu.rs:332-345 at smaller sizes."""
    h = Hip(wraparound=False)
    hg = Hip(wraparound=False, force_global=True)
    try:
        assert not h.full_k()
        common.assert_same_raw(lambda: h, Cpu, 100_000, 4, 12)
        common.assert_same_raw(lambda: hg, Cpu, 87_914, 4, 12)
        common.powmod_and_mul(lambda: h, 2000, 9)
        common.random_modmul(lambda: h, 1000, 40, 1000, 3)
        common.edges(lambda: h, 3, max_len_log2=13)
        common.carry_stress(lambda: h, 1 << 13)
    finally:
        h.close()
        hg.close()


def test_long_run_split_into_launches():
    """Runs split into launches of about max_launch_ms (tests/gpu.rs:179-184, shorter)."""
    h = Hip(max_launch_ms=20)
    try:
        common.long_run(lambda: h, 87_914, 8, 600)
    finally:
        h.close()


def test_long_run_200000_steps():
    """tests/gpu.rs:179-184 at its own length: 87,914-bit moduli and 200,000-bit exponents, one
    run(200000) split into launches of about 200 ms (a plain first chunk, then graph replays),
    against GMP powm on every slot (8 slots, checked on 8 host threads)."""
    h = Hip(max_launch_ms=200)
    try:
        rng = random.Random(179)
        moduli = [common.rand_odd(rng, 87_914) for _ in range(8)]
        exps = [common.rand_exponent(rng, 200_000) for _ in range(8)]
        b = ModBatch(h, moduli, 3)
        for i, e in enumerate(exps):
            b.load(i, 1)
            b.set_exponent(i, e)
        say("long run: 8 slots x 200,000 steps")
        b.run(200_000)
        ms, launches = Hip.last_run_kernel_ms(b)
        say("long run:", launches, "launches,", round(ms), "ms; checking against powm")
        assert launches > 1
        assert b.check_pow_range(3, 0, exps, threads=8) is None
        b.free()
    finally:
        h.close()


def test_graph_replayed_chunks_match_eager():
    """With max_launch_ms, the whole chunks of a run after the first are replays of one captured
    graph. Several runs in a row, with host reads and eager runs of a second backend between
    them: every slot's raw limbs equal the single-launch (eager) backend's after every run."""
    rng = random.Random(21)
    split, eager = Hip(max_launch_ms=1), Hip()
    try:
        for bits in (20_000, 100_000):
            moduli = [common.rand_odd(rng, bits) for _ in range(16)]
            exps = [rng.getrandbits(700) | (1 << 699) for _ in range(16)]
            a, b = ModBatch(split, moduli, 3), ModBatch(eager, moduli, 3)
            for bt in (a, b):
                for i, e in enumerate(exps):
                    bt.load(i, 1)
                    bt.set_exponent(i, e)
            launches = 0
            for chunk in (150, 250, 300):
                a.run(chunk)
                launches = max(launches, Hip.last_run_kernel_ms(a)[1])
                b.run(chunk)
                for i in range(16):
                    assert a.read_raw(i) == b.read_raw(i), f"{bits} bits, slot {i}, after run({chunk})"
            assert launches > 2  # chunks replayed as a graph
            for i, n in enumerate(moduli):
                assert a.remaining(i) == 0 and a.read(i) == common.powmod(3, exps[i], n)
            a.free()
            b.free()
    finally:
        split.close()
        eager.close()


def test_autotuned_plans_match_default():
    """The autotuner times every plan that fits (radix-8 register kernel or radix-4 kernel,
    global or LDS-resident, exchange tile, slots per CU) on the batch's own slots and keeps the
    fastest: the chosen plan is
    reported as tuned, a second batch of the same shape reuses it, and the results are the
    default planner's, bit for bit."""
    tuned = Hip(autotune=True)
    try:
        for bits, slots in ((3000, 64), (20_000, 64), (100_000, 32)):
            rng = random.Random(bits)
            moduli = [common.rand_odd(rng, bits) for _ in range(slots)]
            b1, b2 = ModBatch(tuned, moduli, 3), ModBatch(tuned, moduli, 3)
            p1, p2 = Hip.plan_of(b1), Hip.plan_of(b2)
            assert p1["tuned"] and p1 == p2, (bits, p1, p2)
            assert p1["kernel"] in (4, 8)
            if b1.params().len < 1024:
                assert p1["kernel"] == 4
            b1.free()
            b2.free()
            common.assert_same_raw(lambda: tuned, lambda: Hip(), bits, 16, 96, seed=bits)
    finally:
        tuned.close()


def test_tile_and_slots_per_cu_plans_match_default():
    """The plans the autotuner adds, fixed through the options: the 4,096-element exchange tile
    at L = 8,192 (one chunk per exchange) and one slot per CU (LDS padded), alone and together,
    give the default plan's raw limbs; a tile the batch's kernel cannot take is an error."""
    for tile, spc in ((4096, 0), (0, 1), (4096, 1)):
        say("plan", tile, spc)
        h = Hip(tile=tile, slots_per_cu=spc)
        try:
            rng = random.Random(tile + spc)
            b = ModBatch(h, [common.rand_odd(rng, 100_000) for _ in range(4)], 3)
            p = Hip.plan_of(b)
            assert p["global"] and p["kernel"] == 8 and p["tile"] == (tile or 2048) and p["slots_per_cu"] == spc, p
            assert p["lds_bytes"] > (80 * 1024 if spc else 0)
            b.free()
            common.assert_same_raw(lambda: h, lambda: Hip(), 100_000, 6, 48, seed=tile + spc)
            if spc and not tile:  # one slot per CU at a radix-4 length too
                common.assert_same_raw(lambda: h, lambda: Hip(), 3000, 5, 40, seed=7)
        finally:
            h.close()
    h = Hip(tile=4096)
    try:
        with pytest.raises(MontmulError, match="is not a plan for L = 2048"):
            ModBatch(h, [common.rand_odd(random.Random(1), 20_000)], 3)
    finally:
        h.close()


@pytest.mark.parametrize("bits, n", [(9000, 512), (20_000, 1024), (40_000, 2048)])
def test_half_vector_tile_matches_whole_vector_tile(bits, n):
    """At L = 1,024 .. 4,096 the register kernel's default exchange tile in global mode is half
    the vector (two chunks per exchange, T kept in registers); the whole-vector tile, asked for
    by its size, is the other plan built there. Both give the same raw limbs on every slot."""
    say("tile", n // 2, "against", n, "bits", bits)
    whole = Hip(tile=n)
    try:
        b = ModBatch(Hip(), [common.rand_odd(random.Random(n), bits)], 3)
        p = Hip.plan_of(b)
        assert p["global"] and p["kernel"] == 8 and p["tile"] == n // 2, p
        b.free()
        b = ModBatch(whole, [common.rand_odd(random.Random(n), bits)], 3)
        p = Hip.plan_of(b)
        assert p["tile"] == n and p["lds_bytes"] > 24 * n, p
        b.free()
        common.assert_same_raw(lambda: whole, lambda: Hip(), bits, 64, 96, seed=n)
    finally:
        whole.close()


@pytest.mark.parametrize("bits, mersenne", [(3000, 3217), (20_000, 19937), (100_000, 0)])
def test_is_one_many_on_the_device(hip_global, bits, mersenne):
    """is_one for many slots by one kernel and one copy (the Hip backend's is_one_many), against
    is_one and the definition: radix-4 (L = 256) and register kernels (L = 2,048, 8,192); Fermat
    runs on composites and on a Mersenne prime of the same length admitted by set_modulus
    (L = 256, 2,048)."""
    say("is_one_many", bits)
    common.is_one_many_check(lambda: hip_global, bits, prime=(1 << mersenne) - 1 if mersenne else None)


def test_two_batches_interleaved(hip):
    """Two live batches with different L, run alternately (tests/gpu.rs:227-258)."""
    rng = random.Random(13)
    batches = []
    for bits in (3000, 50_000):
        moduli = [common.rand_odd(rng, bits) for _ in range(4)]
        exps = [rng.getrandbits(300) | 1 for _ in range(4)]
        b = ModBatch(hip, moduli, 3)
        for i, e in enumerate(exps):
            b.load(i, 1)
            b.set_exponent(i, e)
        batches.append((b, moduli, exps))
    assert batches[0][0].params().len != batches[1][0].params().len
    while any(b.remaining(i) > 0 for b, _, _ in batches for i in range(4)):
        for b, _, _ in batches:
            b.run(7)
    for b, moduli, exps in batches:
        for i in range(4):
            assert b.read(i) == common.powmod(3, exps[i], moduli[i]), f"slot {i}"


def test_oversized_batch_is_error(hip):
    """create fails with a message naming bytes before allocating (tests/gpu.rs:186-196). The
    setup is passed straight to the backend: its limb arrays are not read before the memory
    check, and 2^31 slots at L = 8,192 need far more than any device holds."""
    import ctypes
    dummy = (ctypes.c_uint32 * 4096)()
    st = mm.BatchSetup(4096, 8192, 3, 2**31, dummy, dummy, dummy)
    out = ctypes.c_void_p()
    rc = hip.table.create(hip.table.self, ctypes.byref(st), ctypes.byref(out))
    assert rc != 0 and not out.value
    assert "bytes" in hip.table.last_error().decode()


def test_no_leak_on_drop(hip):
    rng = random.Random(14)
    moduli = [common.rand_odd(rng, 50_000) for _ in range(64)]

    def once():
        b = ModBatch(hip, moduli, 3)
        for i in range(len(moduli)):
            b.load(i, 1)
            b.set_exponent(i, 12345)
        b.run(5)
        b.mul(0, 0, 0)
        b.free()

    once()
    before = hip.free_memory()
    for _ in range(20):
        once()
    after = hip.free_memory()
    assert after + (16 << 20) >= before, f"free device memory {before} -> {after}"


def test_headline_shape_closed_form(hip, hip_global):
    """100,000-bit moduli (L = 8,192, k = 4,096), 256 slots: every slot's raw limbs equal the
    closed-form REDC chain with k = L/2 (GMP), in both modes."""
    moduli, exps = mm.score_workload(100_000, 256, 16, 2026)
    ref = ModBatch(Gmp(full_k=True), moduli, 3)
    for i, e in enumerate(exps):
        ref.load(i, 1)
        ref.set_exponent(i, e)
    ref.run(16)
    auto = Hip()  # the default planner: global mode for the register-transform kernel
    for be, name in ((hip, "lds"), (hip_global, "global"), (auto, "default")):
        say("headline shape", name)
        b = ModBatch(be, moduli, 3)
        assert b.params().k == 4096 and b.params().len == 8192
        assert Hip.plan_of(b)["global"] == (name != "lds"), name
        for i, e in enumerate(exps):
            b.load(i, 1)
            b.set_exponent(i, e)
        b.run(16)
        bad = [i for i in range(256) if b.read_raw(i) != ref.read_raw(i)]
        assert not bad, f"{name}: slots {bad[:8]} differ"
        b.free()
    auto.close()


def test_headline_all_4096_slots_against_powm(hip_global):
    """The headline batch itself: 4,096 moduli of 100,000 bits from the score workload, the top
    64 bits of N - 1 as exponents, one run(64) in global mode; every slot equals 3^e mod N
    (GMP powm on 16 host threads, in parts of 1,024 slots)."""
    moduli, exps = mm.score_workload(100_000, 4096, 64, 2026)
    b = ModBatch(hip_global, moduli, 3)
    for i, e in enumerate(exps):
        b.load(i, 1)
        b.set_exponent(i, e)
    b.run(64)
    for first in range(0, 4096, 1024):
        say("headline slots", first)
        assert b.check_pow_range(3, first, exps[first:first + 1024], threads=16) is None
    common.pow_range_check(lambda: hip_global, 3000, 9, 60)
    b.free()


def test_full_fermat_20000_bits_every_slot(hip, hip_global):
    """The full-test workload itself (config 3 of the survey's measurement table): 2,048 slots of
    20,000-bit moduli (L = 2,048), the first 1,024 cycling through the planted primes of
    tests/golden/planted_20000bit.txt, the rest from GMP's generator seeded 7, e = N - 1. Global
    and LDS-resident plans give the same raw limbs on every slot; every planted prime reads 1,
    and every other slot equals 3^(N-1) mod N (GMP powm on 16 host threads)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "planted_20000bit.txt")
    planted = [int(line, 16) for line in open(path) if line.strip()]
    rand, _ = mm.score_workload(20_000, 1024, 20_000, 7)
    moduli = [planted[i % len(planted)] for i in range(1024)] + rand
    exps = [n - 1 for n in moduli]
    raws = []
    for be, name in ((hip_global, "global"), (hip, "lds")):
        say("full Fermat, 2,048 slots,", name)
        b = ModBatch(be, moduli, 3)
        assert b.params().len == 2048 and Hip.plan_of(b)["global"] == (name == "global")
        for i, e in enumerate(exps):
            b.load(i, 1)
            b.set_exponent(i, e)
        b.run(20_000)
        raws.append([b.read_raw(i) for i in range(len(moduli))])
        if name == "global":
            assert all(b.is_one(i) for i in range(1024))
            assert b.check_pow_range(3, 1024, exps[1024:], threads=16) is None
        b.free()
    bad = [i for i in range(len(moduli)) if raws[0][i] != raws[1][i]]
    assert not bad, f"slots {bad[:8]} differ between the plans"


def test_cpu_case_4096_bits(hip):
    """The reference's CPU-runnable case: 64 moduli of 4,096 bits (L = 512, k = 256) from the
    score workload seeded 2026, e = (N - 1) >> 3,584; arm, run(16), re-arm, run(512). Every slot:
    read == 3^e mod N and raw == the closed-form chain (GMP, k = L/2)."""
    moduli, exps = mm.score_workload(4096, 64, 512, 2026)
    assert all(e == (n - 1) >> 3584 for n, e in zip(moduli, exps))
    b = ModBatch(hip, moduli, 3)
    assert (b.params().k, b.params().len) == (256, 512)
    ref = ModBatch(Gmp(full_k=True), moduli, 3)
    for bb in (b, ref):
        for i, e in enumerate(exps):
            bb.load(i, 1)
            bb.set_exponent(i, e)
    b.run(16)
    for i, e in enumerate(exps):
        b.load(i, 1)
        b.set_exponent(i, e)
    b.run(512)
    ref.run(512)
    for i, (n, e) in enumerate(zip(moduli, exps)):
        assert b.remaining(i) == 0
        assert b.read(i) == common.powmod(3, e, n), f"slot {i}"
        assert b.read_raw(i) == ref.read_raw(i), f"slot {i} raw"


def test_set_modulus(hip, hip_global):
    """Admission of a new modulus into a running batch, both modes and both kernels' lengths."""
    for be in (hip, hip_global):
        for bits in (60, 1000, 5000, 100_000):
            common.set_modulus_admission(lambda: be, bits)
    common.set_modulus_admission(lambda: Hip(wraparound=False), 1000)


@pytest.mark.parametrize("bits", [1000, 20_000, 100_000, 300_000])
def test_run_bare_is_the_squaring_alone(hip_global, bits):
    """Gpu::run_bare (gpu/mod.rs:109-111; lib.rs:949-961): each step is the squaring alone, so x
    stays as loaded, the position advances, and the product left in HBM is T = c x^2 exactly
    (c = 3 on a 1 bit, 1 on a 0 bit): the radix-8 register kernel (L = 1,024 .. 8,192) and the
    radix-4 kernel (L = 64 and 32,768) in global mode."""
    rng = random.Random(bits)
    moduli = [common.rand_odd(rng, bits) for _ in range(3)]
    b = ModBatch(hip_global, moduli, 3)
    xs = [rng.randrange(2 * n) for n in moduli]
    for i, x in enumerate(xs):
        b.load_raw(i, x)
        b.set_exponent(i, 0b101)
    for step, c in enumerate((3, 1, 3)):
        Hip.run_bare(b, 1)
        for i, x in enumerate(xs):
            assert b.remaining(i) == 2 - step
            assert b.read_raw(i) == x, f"slot {i}: x changed"
            assert Hip.read_product(b, i) == c * x * x, f"bits {bits} slot {i} step {step}"


def test_overhead_of_the_reduction():
    """score.rs:138-160 on a reduced score workload (100,000 bits, 256 slots, 64-bit exponents):
    run over run_bare, the whole step against the squaring alone, is above 1."""
    from montmul_amd import score

    cfg = score.ScoreConfig(bits=100_000, slots=256, ebits=64, repeats=3)
    ratio = score.gpu_overhead(lambda: Hip(), cfg)
    say("overhead run / run_bare", round(ratio, 3))
    assert 1.2 < ratio < 4.0


def test_l32768_register_kernel_against_oracles(hip_global):
    """L = 32,768 (n = 2^14) runs the register-transform step as two sub-transforms of n/2 in
    turn. Raw limbs after every slot's chain equal the Gmp closed form (k = L/2) on 16 slots x 64
    steps, and the Cpu restatement on 2 slots x 12 steps; x = 2N - 1 and N = 2^b - 1 edges."""
    for bits in (262_200, 500_000):
        b = ModBatch(hip_global, [common.rand_odd(random.Random(bits), bits)], 3)
        assert b.params().len == 32768
        plan = Hip.plan_of(b)
        assert plan["kernel"] == 8 and plan["global"] and plan["threads"] == 1024, plan
    say("L=32768 vs Gmp, 16 slots x 64 steps")
    common.assert_same_raw(lambda: hip_global, lambda: Gmp(full_k=True), 300_000, 16, 64)
    say("L=32768 vs Cpu, 2 slots x 12 steps")
    common.assert_same_raw(lambda: hip_global, lambda: Cpu(wraparound=True), 400_000, 2, 12)
    bits = 32 * 16384 - 4  # the top of the bucket: N = 2^b - 1, x = 2N - 1
    n = (1 << bits) - 1
    b = ModBatch(hip_global, [n, common.rand_odd(random.Random(7), bits)], 3)
    g = ModBatch(Gmp(full_k=True), [n, common.rand_odd(random.Random(7), bits)], 3)
    for bb in (b, g):
        bb.load_raw(0, 2 * n - 1)
        bb.load_raw(1, 12345)
        for i in range(2):
            bb.set_exponent(i, 0xF0F1)
        bb.run(16)
    for i in range(2):
        assert b.read_raw(i) == g.read_raw(i), f"edge slot {i}"


def test_multi_two_contexts_on_one_device():
    """tests/gpu.rs:347-357: Multi over two HIP contexts (streams) on device 0, one host thread
    per shard: raw limbs equal to the Cpu restatement for 1, 5 and 16 slots, powers and mul
    (within and across shards), Fermat tests; then tests/gpu.rs:368-373, the carry-out of
    slot 1 (slot 0 of shard 1) named "shard 1, slot 1"."""
    parts = [Hip(), Hip()]
    try:
        two = lambda: mm.Multi(parts)
        for slots in (1, 5, 16):
            say("multi: 87,914 bits,", slots, "slots")
            common.assert_same_raw(two, lambda: Cpu(wraparound=True), 87_914, slots, 50)
        common.assert_same_raw(two, lambda: Cpu(wraparound=True), 1000, 7, 100)
        common.powmod_and_mul(two, 2000, 9)
        common.end_to_end(two, [6, 7, 11, 87, 98, 215, 802, 1001, 2000, 4000])
        rng = random.Random(12)
        b = ModBatch(two(), [common.rand_odd(rng, 1000) for _ in range(2)], 4)
        for i in range(2):
            b.load(i, 1)
            b.set_exponent(i, 1)
        common.load_all_ones(b, 1)
        with pytest.raises(CarryError, match="shard 1, slot 1"):
            b.run(1)
    finally:
        for p in parts:
            p.close()
