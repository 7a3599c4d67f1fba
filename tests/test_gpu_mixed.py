"""Mixed-length batches on the GPU: slots of several transform lengths stepped by the
mixed-length kernel (one persistent launch per chunk, workgroups taking per-slot descriptors
from a device queue). Every slot must equal pow(x, 2^nbits) 3^e mod N and carry the same raw
limbs as its length's plain ModBatch (the per-length step kernels). No reference counterpart:
the reference's ModBatch holds one length (batch.rs:64-77)."""

import pytest

import common
from montmul_amd import Hip, MixBatch

pytestmark = pytest.mark.gpu

# Sizes per transform length, L = 8 .. 16,384 (register-transform lengths 1,024 .. 16,384).
SMALL = [64, 200, 1000, 2000]                        # L = 8, 16, 64, 128
V2 = [6000, 12000, 20000, 40000, 60000, 100000]      # L = 512 .. 8,192
LARGEST = [150000, 200000]                           # L = 16,384: 1,024-thread workgroups


def test_mixed_all_lengths_1024_threads(hip_global):
    """A group with L = 16,384 runs 1,024-thread workgroups: every shorter register-transform
    length runs with idle waves."""
    lengths = common.mixed_lengths(lambda: hip_global, SMALL + V2 + LARGEST + [3000, 8000, 130000], 64, 25)
    assert lengths[-1] == 16384 and 256 in lengths and 8192 in lengths


def test_mixed_up_to_8192_512_threads(hip):
    """Without L = 16,384 the workgroups have 512 threads; batches planned LDS-resident get
    their HBM buffers for the mixed launch on first use."""
    lengths = common.mixed_lengths(lambda: hip, SMALL + V2 + [9000, 70000], 48, 16, seed=10)
    assert lengths[-1] == 8192


def test_mixed_with_radix4_global_length(hip_global):
    """L = 32,768 (above the register-transform range) runs the radix-4 step in global mode
    inside the mixed launch."""
    lengths = common.mixed_lengths(lambda: hip_global, [300000, 5000, 100000, 16000], 16, 7, seed=11)
    assert lengths == [512, 1024, 8192, 32768]


def test_mixed_without_wraparound():
    """The reduction by the full product N-hat (no half path) in the mixed launch."""
    h = Hip(wraparound=False)
    try:
        common.mixed_lengths(lambda: h, [700, 3000, 20000, 60000], 32, 11, seed=12)
    finally:
        h.close()


def test_mixed_launch_splitting():
    """max_launch_ms splits a run into several mixed launches (steps per launch from the
    previous launch's time)."""
    h = Hip(max_launch_ms=1)
    try:
        common.mixed_lengths(lambda: h, [5000, 20000, 100000, 200000], 80, 80, seed=13)
    finally:
        h.close()


def test_mixed_fermat_planted(hip_global):
    """Full Fermat tests of known primes and composites of different lengths in one batch."""
    ms = [2**4253 - 1, 2**4423 - 1, 2**9689 - 1, common.tribonacci(4202)]
    ms = [m for m in ms if m % 2 == 1]
    comp = [m * 3 + 2 for m in ms]
    comp = [c if c % 2 else c + 1 for c in comp]
    moduli = ms + comp
    b = MixBatch(hip_global, moduli, 3)
    for i, n in enumerate(moduli):
        b.load(i, 1)
        b.set_exponent(i, n - 1)
    while any(b.remaining(i) for i in range(len(moduli))):
        b.run(4096)
    for i, n in enumerate(moduli):
        assert b.is_one(i) == (common.powmod(3, n - 1, n) == 1), f"slot {i}"
    assert all(b.is_one(i) for i in range(len(ms)))
    b.free()


def test_mixed_full_slot_groups(hip_global):
    """Enough slots per length that the short register-transform lengths fill whole slot
    groups (several slots per workgroup) and leave partial ones, with exponents of different
    lengths, so groups go idle at different steps."""
    sizes = [6000] * 17 + [12000] * 9 + [20000] * 10 + [40000] * 5 + [100000] * 5 + [150000]
    lengths = common.mixed_lengths(lambda: hip_global, sizes, 48, 20, seed=14)
    assert lengths == [512, 1024, 2048, 4096, 8192, 16384]


def test_mixed_radix4_in_1024_thread_launch(hip_global):
    """With L = 16,384 in the group the launch has 1,024-thread workgroups: the radix-4 step runs
    there both as one-wave slot groups (L <= 512) and over the whole workgroup in global mode
    (L = 32,768)."""
    lengths = common.mixed_lengths(lambda: hip_global, [300000, 150000, 3000, 3000, 7000, 60, 900], 24, 9, seed=15)
    assert lengths == [8, 64, 256, 512, 16384, 32768]


def test_mixed_carry_out_lists_every_slot(hip_global):
    """Slots of two lengths carry out in the same mixed launch: the error names the first, and
    failed_slots() lists both (by their index in the mixed batch); the other slots are right."""
    import ctypes
    import random

    import montmul_amd as mm
    from montmul_amd import CarryError

    rng = random.Random(40)
    moduli = [common.rand_odd(rng, b) for b in (1000, 1000, 20000, 20000, 100000)]
    mix = MixBatch(hip_global, moduli, 4)
    assert mix.groups() == 3
    for i in range(len(moduli)):
        mix.load(i, 1)
        mix.set_exponent(i, 1)
    # slot 1 is index 1 of group 0 (L = 64), slot 3 is index 1 of group 1 (L = 2,048)
    for g, j in ((0, 1), (1, 1)):
        st = mm.lib.mm_mixbatch_group_state(mix._h, g)
        k = mix.params(1 if g == 0 else 3).k
        buf = (ctypes.c_uint32 * k)(*([0xFFFFFFFF] * k))
        assert hip_global.table.load(ctypes.c_void_p(st), j, buf) == 0
    with pytest.raises(CarryError) as ei:
        mix.run(1)
    assert ei.value.slot == 1 and ei.value.slots == [1, 3]
    assert mix.failed_slots() == [1, 3]
    for i in (0, 2, 4):
        assert mix.read(i) == 4 % moduli[i]


def test_mixed_1024_threads_with_radix4_global_length(hip_global):
    """A 1,024-thread launch (L = 16,384 present) that also steps L = 32,768 slots: their batch
    runs the register step on its own, but the radix-4 step in global mode inside the mixed
    launch (whole-workgroup carries, a called function in the 1,024-thread build)."""
    lengths = common.mixed_lengths(lambda: hip_global, [150000, 300000, 5000, 60000], 24, 9, seed=12)
    assert lengths == [512, 4096, 16384, 32768]


def test_mixed_raw_limbs_against_the_closed_form(hip_global):
    """One persistent launch over L = 256 .. 16,384 (radix-4 slot groups, register kernels in
    slot groups, wide and full workgroups), every slot's raw limbs against the oracle's Gmp
    closed form with k = L/2 (gmp.rs:45-61) per length, not against another HIP batch."""
    from oracle_backends import Gmp

    sizes = [3000, 6000, 9000, 20000, 40000, 60000, 90000, 120000, 150000, 200000] * 2
    lengths = common.mixed_lengths(lambda: hip_global, sizes, 24, 7, seed=21, ref=lambda: Gmp(full_k=True))
    assert lengths == [256, 512, 1024, 2048, 4096, 8192, 16384]


def test_mixed_is_one_many(hip_global):
    """Fermat tests over six lengths in one mixed batch, retired by is_one_many on the device."""
    common.mix_is_one_many_check(lambda: hip_global)
