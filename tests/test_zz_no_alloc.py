"""The allocation-free product (mm_mul_noalloc) against mm_mul, through the C ABI.

Two calls in a row, the second on a smaller input (100,000-bit moduli at L = 8,192, then
5,000-bit moduli at L = 512), each bit for bit against mm_mul on an identical batch and against
x_a x_b R^-1 mod N; a carry-out names the slot as mm_mul's does; lengths whose scratch exceeds
the fixed array are refused with an error."""

import ctypes
import random

import pytest

import common
import montmul_amd as mm
from montmul_amd import Hip, ModBatch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def noalloc(b, out, a, c):
    return mm.lib.mm_mul_noalloc(ctypes.c_void_p(b.state), out, a, c)


def pair(hip, moduli, xs):
    """Two batches with the same moduli and loaded values."""
    out = []
    for _ in range(2):
        b = ModBatch(hip, moduli, 3)
        for i, x in enumerate(xs):
            b.load(i, x)
        out.append(b)
    return out


@pytest.mark.parametrize("bits,slots", [(100_000, 3), (5000, 4)])
def test_mul_noalloc_matches_mul(bits, slots):
    rng = random.Random(bits)
    hip = Hip()
    try:
        n = common.rand_odd(rng, bits)
        moduli = [n] * slots
        xs = [rng.randrange(n) for _ in range(slots)]
        ref, got = pair(hip, moduli, xs)
        say("mul_noalloc", bits, "bits, L", ref.params().len)
        # two products in a row on the same batch: the second reads the first's output
        for out, a, c in ((0, 1, 2), (2, 0, 1)):
            ref.mul(out, a, c)
            assert noalloc(got, out, a, c) == mm.MM_OK, mm.lib.mm_last_error()
            assert got.read_raw(out) == ref.read_raw(out), f"slot {out}"
        for i in range(slots):
            assert got.read_raw(i) == ref.read_raw(i)
        # the values themselves: x0' = x1 x2, x2' = x0' x1 (Montgomery: loaded values x R)
        assert got.read(0) == xs[1] * xs[2] % n
        assert got.read(2) == xs[1] * xs[2] * xs[1] % n
    finally:
        hip.close()


def test_mul_noalloc_twice_then_smaller():
    """The spec's order: a large product, then a smaller one on another batch, both exact."""
    rng = random.Random(77)
    hip = Hip()
    try:
        for bits in (100_000, 5000):
            n = common.rand_odd(rng, bits)
            xs = [rng.randrange(n) for _ in range(2)]
            ref, got = pair(hip, [n, n], xs)
            ref.mul(1, 0, 1)
            assert noalloc(got, 1, 0, 1) == mm.MM_OK, mm.lib.mm_last_error()
            assert got.read_raw(1) == ref.read_raw(1)
            assert got.read(1) == xs[0] * xs[1] % n
            got.free()
            ref.free()
    finally:
        hip.close()


def test_mul_noalloc_carry_names_the_slot():
    """x = R - 1 (outside [0, 2N)) in both factors: the product carries out, as in mm_mul."""
    rng = random.Random(78)
    hip = Hip()
    try:
        n = common.rand_odd(rng, 5000)
        b = ModBatch(hip, [n, n], 3)
        k = b.params().k
        buf = (ctypes.c_uint32 * k)(*([0xFFFFFFFF] * k))
        for i in range(2):
            assert mm.lib.mm_load(ctypes.c_void_p(b.state), i, buf) == 0
        assert noalloc(b, 1, 0, 0) == 1
        assert b"slot 1" in mm.lib.mm_last_error()
        # the error word was cleared: a valid product afterwards succeeds
        b.load(0, 5)
        b.load(1, 7)
        assert noalloc(b, 1, 0, 1) == mm.MM_OK, mm.lib.mm_last_error()
        assert b.read(1) == 35 % n
    finally:
        hip.close()


def test_mul_noalloc_refuses_lengths_beyond_its_scratch():
    """L = 262,144 needs 48 n + 12 k = 7.5 MiB of scratch, above the 4 MiB array."""
    rng = random.Random(79)
    hip = Hip()
    try:
        n = common.rand_odd(rng, 2_100_000)
        b = ModBatch(hip, [n], 3)
        assert b.params().len == 262_144
        b.load(0, 3)
        assert noalloc(b, 0, 0, 0) == mm.MM_EFAULT
        assert b"fixed scratch" in mm.lib.mm_last_error()
    finally:
        hip.close()
