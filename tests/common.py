"""Backend-generic correctness cases, restating the reference's shared test library
(montmul/tests/common/mod.rs). Every check compares against Python integers (pow, modular
inverse, exact products), which play the role of GMP through `rug` in the reference.

`make` is a callable returning a fresh backend (an object with a `table` and the Limits
methods); `ModBatch` and `fermat` are the host API under test.
"""

import random

import pytest

from montmul_amd import MixBatch, MontmulError, ModBatch, Params, fermat


def powmod(b, e, m):
    """b^e mod m (e = -1: the inverse). Moduli above 2,048 bits go through GMP (the oracle's
    og_powm / og_invert, as the reference's tests use rug's pow_mod and invert), where Python's
    integers take seconds per call; smaller ones use Python's pow."""
    if m.bit_length() <= 2048:
        return pow(b, e, m)
    from oracle_backends import gmp_invert, gmp_powm

    if e == -1:
        return gmp_invert(b % m, m)
    return gmp_powm(b % m, e, m)


def mulmod(a, b, m):
    """a b mod m, through GMP above 2,048 bits (Python's division is quadratic there)."""
    if m.bit_length() <= 2048:
        return a * b % m
    from oracle_backends import gmp_mulmod

    return gmp_mulmod(a % m, b % m, m)


def rand_odd(rng, bits):
    """Random odd integer of exactly `bits` bits (common/mod.rs:10-15)."""
    return rng.getrandbits(bits) | (1 << (bits - 1)) | 1


def expected(x, c, n, k):
    """x^2 c R^-1 mod N, the expected step result (common/mod.rs:17-21)."""
    rinv = powmod(1 << (32 * k), -1, n)
    return mulmod(powmod(x, 2, n), c * rinv, n)


def tribonacci(n):
    """T(n), T(0) = T(1) = 0, T(2) = 1 (OEIS A000073; seq.rs:6-15)."""
    a, b, c = 0, 0, 1
    for _ in range(n):
        a, b, c = b, c, a + b + c
    return a


# OEIS A092835: n - 1 for prime T(n) (common/mod.rs:284-288).
A092835_SMALL = [6, 7, 11, 87, 98, 215, 802]
A092835_LARGE = [4202, 18699, 96879]


def random_modmul(make, bits, cases, seed, base):
    """Random odd N, random x in [0, 2N); one step with c = base, then one with c = 1
    (common/mod.rs:93-120)."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, bits) for _ in range(cases)]
    batch = ModBatch(make(), moduli, base)
    k = batch.params().k
    xs = [rng.randrange(2 * n) for n in moduli]
    for i, x in enumerate(xs):
        batch.load_raw(i, x)
        batch.set_exponent(i, 2)
    prev = list(xs)
    for c in (base, 1):
        batch.run(1)
        for i, n in enumerate(moduli):
            raw = batch.read_raw(i)
            assert raw < 2 * n, f"lazy range, bits {bits} slot {i}"
            assert raw % n == expected(prev[i], c, n, k), f"bits {bits} slot {i} c {c}"
            prev[i] = raw


def edge_sizes(max_len_log2=16):
    sizes = [28, 60, 124, 252, 1020, 4092, 32764]
    for m in range(3, max_len_log2 + 1):
        kmax = (1 << m) // 2
        sizes.append(32 * kmax - 4)           # top of the bucket
        sizes.append(32 * (kmax // 2) - 3)    # bottom of the bucket
    return sizes


def edges(make, base, max_len_log2=16, seed=1):
    """x = 0, N, 2N - 1 with N random and N = 2^b - 1, for bucket edges (common/mod.rs:122-159)."""
    rng = random.Random(seed)
    for bits in edge_sizes(max_len_log2):
        if bits < 2:
            continue
        moduli = [rand_odd(rng, bits), (1 << bits) - 1]
        batch = ModBatch(make(), moduli, base)
        k = batch.params().k
        for which in range(3):
            for i, n in enumerate(moduli):
                x = [0, n, 2 * n - 1][which]
                batch.load_raw(i, x)
                batch.set_exponent(i, 1)
            batch.run(1)
            for i, n in enumerate(moduli):
                x = [0, n, 2 * n - 1][which]
                raw = batch.read_raw(i)
                assert raw < 2 * n, f"lazy range bits {bits}"
                assert raw % n == expected(x, base, n, k), f"bits {bits} case {which}"


def batch_independence(make, bits, sizes, seed=2):
    """One modulus gives identical raw limbs in every slot position and batch size
    (common/mod.rs:161-188)."""
    rng = random.Random(seed)
    target = rand_odd(rng, bits)
    e = rng.getrandbits(64) | 1
    reference = None
    for size in sizes:
        moduli = [rand_odd(rng, bits) for _ in range(size)]
        pos = min(size * 5 // 7, size - 1)
        for p in (0, pos, size - 1):
            moduli[p] = target
        batch = ModBatch(make(), moduli, 3)
        for i in range(size):
            batch.load(i, 1)
            batch.set_exponent(i, e)
        batch.run(1000)
        for p in (0, pos, size - 1):
            raw = batch.read_raw(p)
            if reference is None:
                reference = raw
            assert raw == reference, f"size {size} slot {p}"
            assert batch.read(p) == powmod(3, e, target)


def powmod_and_mul(make, bits, count, seed=3):
    """Chunked exponentiation against pow, then `mul` (common/mod.rs:190-220)."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, bits) for _ in range(count)]
    batch = ModBatch(make(), moduli, 3)
    exps = []
    for i in range(count):
        e = rng.getrandbits(32 + i * 7) | 1
        batch.load(i, 1)
        batch.set_exponent(i, e)
        exps.append(e)
    while any(batch.remaining(i) > 0 for i in range(count)):
        batch.run(37)
    for i, n in enumerate(moduli):
        assert batch.read(i) == powmod(3, exps[i], n), f"slot {i}"
    n = rand_odd(rng, bits)
    batch = ModBatch(make(), [n, n, n], 3)
    a, b = rng.randrange(n), rng.randrange(n)
    batch.load(0, a)
    batch.load(1, b)
    batch.mul(2, 0, 1)
    assert batch.read(2) == a * b % n
    batch.mul(0, 0, 0)
    assert batch.read(0) == a * a % n


def rand_exponent(rng, bits):
    return rng.getrandbits(bits) | (1 << (bits - 1))


def assert_same_raw(make_a, make_b, bits, slots, steps, seed=4):
    """Two backends on the same moduli and exponents end with equal raw limbs
    (common/mod.rs:229-247): the parity primitive."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, bits) for _ in range(slots)]
    exps = [rand_exponent(rng, steps) for _ in range(slots)]
    x = ModBatch(make_a(), moduli, 3)
    y = ModBatch(make_b(), moduli, 3)
    for i, e in enumerate(exps):
        for b in (x, y):
            b.load(i, 1)
            b.set_exponent(i, e)
    x.run(steps)
    y.run(steps)
    for i in range(slots):
        assert x.read_raw(i) == y.read_raw(i), f"bits {bits} slot {i}"


def long_run(make, bits, slots, ebits, seed=5):
    """One run(ebits) from load(1) gives 3^e mod N (common/mod.rs:249-264)."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, bits) for _ in range(slots)]
    exps = [rand_exponent(rng, ebits) for _ in range(slots)]
    batch = ModBatch(make(), moduli, 3)
    for i, e in enumerate(exps):
        batch.load(i, 1)
        batch.set_exponent(i, e)
    batch.run(ebits)
    for i, n in enumerate(moduli):
        assert batch.remaining(i) == 0, f"slot {i}"
        assert batch.read(i) == powmod(3, exps[i], n), f"bits {bits} slot {i}"


def pow_range_check(make, bits, slots, ebits, seed=15):
    """check_pow_range: -1 (None) when every slot of a finished run equals 3^e mod N, in one
    part or several; the lowest wrong slot when one is changed; a slot with bits left fails."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, bits) for _ in range(slots)]
    exps = [rand_exponent(rng, ebits) for _ in range(slots)]
    batch = ModBatch(make(), moduli, 3)
    for i, e in enumerate(exps):
        batch.load(i, 1)
        batch.set_exponent(i, e)
    batch.run(ebits)
    assert batch.check_pow_range(3, 0, exps, threads=3) is None
    assert batch.check_pow_range(3, 2, exps[2:5]) is None
    wrong = [slots - 2, slots - 1]
    for i in wrong:
        batch.load(i, powmod(3, exps[i], moduli[i]) ^ 2)
    assert batch.check_pow_range(3, 0, exps, threads=4) == slots - 2
    assert batch.check_pow_range(3, slots - 1, exps[-1:]) == slots - 1
    batch.load(0, 1)
    batch.set_exponent(0, 5)
    assert batch.check_pow_range(3, 0, exps[:1]) == 0  # bits left
    with pytest.raises(MontmulError, match="slot range past the batch"):
        batch.check_pow_range(3, slots - 1, exps[:2])


def end_to_end(make, ns):
    """Fermat base 3 on T(n), grouped by transform length (common/mod.rs:266-282)."""
    moduli = [t for t in (tribonacci(n) for n in ns) if t % 2 == 1 and t > 3]
    lengths = make().supported_lengths()
    groups = {}
    for m in moduli:
        groups.setdefault(Params.for_bits(m.bit_length(), lengths).len, []).append(m)
    for ms in groups.values():
        got = fermat(make(), ms, 3, 500)
        for m, g in zip(ms, got):
            assert g == (powmod(3, m - 1, m) == 1), f"T with {m.bit_length()} bits"


def known_prps(make, prime_ns, composite_ns):
    """Known tribonacci primes report PRP, composites do not (common/mod.rs:290-300)."""
    for n, want in [(p, True) for p in prime_ns] + [(c, False) for c in composite_ns]:
        t = tribonacci(n)
        if t % 2 == 0:
            continue
        assert fermat(make(), [t], 3, 1 << 20)[0] == want, f"T({n})"


def api_contract(make, seed=6):
    """Parameter rules, rejection of bad input, canonical read, is_one, chunked runs,
    frozen slots and re-arming (common/mod.rs:302-374)."""
    import pytest

    lengths = make().supported_lengths()
    assert all(l & (l - 1) == 0 for l in lengths)
    for bad, base, what in [([], 3, "empty batch"), ([10], 3, "even modulus"), ([1], 3, "modulus 1"),
                            ([-7], 3, "negative modulus"), ([11], 5, "base above lazy range"),
                            ([11], 0, "base 0"), ([101, (1 << 5000) + 1], 3, "mixed transform lengths")]:
        with pytest.raises(MontmulError):
            ModBatch(make(), bad, base)
    for bits in [2, 28, 29, 60, 61, 1000, 87_914]:
        p = Params.for_bits(bits, lengths)
        assert p.k * 32 >= bits + 4 and (p.k - 1) * 32 < bits + 4
        assert p.len >= 2 * p.k and (p.len == lengths[0] or p.len // 2 < 2 * p.k)
    if make().coefficient_bound() < 2**128 - 1:
        with pytest.raises(MontmulError):
            Params(1 << 30, 1 << 31).check_exact(4, make().coefficient_bound())
    rng = random.Random(seed)
    for n in [3, 5, 7, 9, 15, 101, 65_537, 4_294_967_291, 2**32 - 1]:
        b = ModBatch(make(), [n], 3)
        for x in [0, 1, n, n * 5 + 2, -4]:
            b.load(0, x)
            assert b.read(0) == x % n, f"read after load, n {n}"
            assert b.is_one(0) == (x % n == 1), f"is_one, n {n}"
            e = rng.getrandbits(40) + 1
            b.load(0, x)
            b.set_exponent(0, e)
            b.run(1 << 20)
            assert b.remaining(0) == 0
            sq = 1 << e.bit_length()
            want = mulmod(powmod(x % n, sq, n), powmod(3, e, n), n)
            assert b.read(0) == want, f"pow, n {n}"
    moduli = [rand_odd(rng, 777) for _ in range(5)]
    exps = [rng.getrandbits(10 + 60 * i) | 1 for i in range(5)]
    a = ModBatch(make(), moduli, 3)
    b = ModBatch(make(), moduli, 3)
    for i in range(5):
        for m in (a, b):
            m.load(i, 1)
            m.set_exponent(i, exps[i])
    a.run(0)
    assert a.remaining(4) == exps[4].bit_length()
    a.run(1000)
    for chunk in [1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 1000]:
        b.run(chunk)
    for i in range(5):
        assert a.read(i) == powmod(3, exps[i], moduli[i])
        assert a.read_raw(i) == b.read_raw(i), f"chunked run slot {i}"
    a.run(100)
    assert a.read(0) == powmod(3, exps[0], moduli[0]), "frozen after finish"
    a.load(0, 2)
    a.set_exponent(0, 5)
    a.run(10)
    assert a.read(0) == (256 * 243) % moduli[0]


def carry_stress(make, max_len):
    """x = 2N - 1 with N = 2^b - 1 and N = 2^b - 2^(b/2) + 1, 32 squarings, every bucket
    (common/mod.rs:376-398)."""
    length = 8
    while length <= max_len:
        bits = 32 * (length // 2) - 4
        n = (1 << bits) - 1
        n2 = (1 << bits) - (1 << (bits // 2)) + 1
        b = ModBatch(make(), [n, n2], 3)
        e = 0xFFFFFFFF
        for i, m in enumerate((n, n2)):
            b.load_raw(i, 2 * m - 1)
            b.set_exponent(i, e)
        b.run(32)
        k = b.params().k
        for i, m in enumerate((n, n2)):
            rinv = powmod(1 << (32 * k), -1, m)
            x = (m - rinv) % m
            want = mulmod(powmod(x, 1 << 32, m), powmod(3, e, m), m)
            assert b.read(i) == want, f"len {length} slot {i}"
        length *= 2

def load_all_ones(b, i):
    """Writes x = R - 1 (every limb 0xFFFFFFFF, outside [0, 2N)) into slot i through the
    backend table's own load, bypassing ModBatch's range check."""
    import ctypes

    k = b.params().k
    buf = (ctypes.c_uint32 * k)(*([0xFFFFFFFF] * k))
    assert b.backend.table.load(ctypes.c_void_p(b.state), i, buf) == 0


def carry_overflow_slots(make, bits, bad=(1, 3), slots=4):
    """Several slots carry out in one run: x = R - 1 in the slots `bad` of a base-4 batch, so
    T = 4 (R - 1)^2 overflows 2k limbs (tests/gpu.rs:269-305 with more than one failing slot).
    The error names the first, failed_slots() lists all of them, every error word is cleared
    (a second run on valid values succeeds with no failed slot), and the other slots are right."""
    from montmul_amd import CarryError

    rng = random.Random(13)
    moduli = [rand_odd(rng, bits) for _ in range(slots)]
    b = ModBatch(make(), moduli, 4)
    for i in range(slots):
        b.load(i, 1)
        b.set_exponent(i, 1)
    for i in bad:
        load_all_ones(b, i)
    try:
        b.run(1)
        raise AssertionError("no carry error")
    except CarryError as e:
        assert e.slot == bad[0] and f"slot {bad[0]}" in str(e)
        assert e.slots == sorted(bad), e.slots
    assert b.failed_slots() == sorted(bad)
    for i in range(slots):
        if i not in bad:
            assert b.read(i) == 4 % moduli[i], f"slot {i}"
    for i in bad:
        b.load(i, 5)
        b.set_exponent(i, 1)
    b.run(1)
    assert b.failed_slots() == []
    for i in bad:
        assert b.read(i) == 25 * 4 % moduli[i]


# ---- Transform layer (common/mod.rs:23-91) -------------------------------------------------
# `t` is an adapter over a transform backend with upload(coeffs, len, count) -> buf,
# forward/inverse(buf, len, count), pointwise(out, a, b, scales, len, count) and
# download_words(buf) -> sequence of 64-bit words, two (low, high) per coefficient.


def coeffs_to_int(words):
    """sum_q c_q 2^(32 q) for coefficients given as (low, high) word pairs, by a linear carry."""
    limbs, carry = [], 0
    for q in range(len(words) // 2):
        carry += int(words[2 * q]) | (int(words[2 * q + 1]) << 64)
        limbs.append(carry & 0xFFFFFFFF)
        carry >>= 32
    while carry:
        limbs.append(carry & 0xFFFFFFFF)
        carry >>= 32
    return int.from_bytes(b"".join(l.to_bytes(4, "little") for l in limbs), "little")


def exactness(t, max_len):
    """All-ones limbs at k = L/2 times themselves with scale 4: coefficient q equals
    4 (2^32 - 1)^2 min(q + 1, 2k - 1 - q) and the whole product is (R - 1)^2 4 (common/mod.rs:23-53)."""
    import numpy as np

    e = 3
    while (1 << e) <= max_len:
        length = 1 << e
        k, c = length // 2, 4
        coeffs = np.zeros(length, dtype=np.uint32)
        coeffs[:k] = 0xFFFFFFFF
        x, y = t.upload(coeffs, length, 1), t.upload(coeffs, length, 1)
        t.forward(x, length, 1)
        t.forward(y, length, 1)
        out = t.upload(np.zeros(length, dtype=np.uint32), length, 1)
        t.pointwise(out, x, y, [c], length, 1)
        t.inverse(out, length, 1)
        w = np.asarray(t.download_words(out), dtype=np.uint64).reshape(-1, 2)
        sq = c * (2**32 - 1) ** 2
        q = np.arange(length, dtype=np.int64)
        cnt = np.where(q < k, q + 1, np.maximum(2 * k - 1 - q, 0))
        want = [sq * int(v) for v in cnt]
        got = [int(lo) | (int(hi) << 64) for lo, hi in w]
        assert got == want, f"len {length}"
        a = (1 << (32 * k)) - 1
        assert coeffs_to_int(w.reshape(-1)) == a * a * c, f"len {length}"
        e += 1


def transform_batched(t, max_len, seed=7):
    """Three vectors per call, products checked against integer multiplication, with scales
    [1, 3, 4] and with none (common/mod.rs:55-91)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    e = 3
    while (1 << e) <= max_len:
        length, count = 1 << e, 3
        k = length // 2
        cs = []
        for _ in range(2):
            c = np.zeros(length * count, dtype=np.uint32)
            for v in range(count):
                c[v * length:v * length + k] = rng.integers(0, 2**32, size=k, dtype=np.uint64).astype(np.uint32)
            cs.append(c)
        for scales in ([1, 3, 4], None):
            x, y = t.upload(cs[0], length, count), t.upload(cs[1], length, count)
            t.forward(x, length, count)
            t.forward(y, length, count)
            out = t.upload(np.zeros(length * count, dtype=np.uint32), length, count)
            t.pointwise(out, x, y, scales, length, count)
            t.inverse(out, length, count)
            w = np.asarray(t.download_words(out), dtype=np.uint64)
            for v in range(count):
                val = coeffs_to_int(w[2 * v * length:2 * (v + 1) * length])
                a = int.from_bytes(cs[0][v * length:(v + 1) * length].tobytes(), "little")
                b = int.from_bytes(cs[1][v * length:(v + 1) * length].tobytes(), "little")
                assert val == a * b * (1 if scales is None else scales[v]), f"len {length} vector {v}"
        e += 1


def set_modulus_admission(make, bits, seed=8):
    """A slot reprogrammed for a new modulus of the same L mid-batch computes as a fresh batch
    would (admission for serving): all slots 3^e mod N after the run, the replaced slot's
    exponent cleared by the replacement, and a modulus of another L refused."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, bits) for _ in range(4)]
    batch = ModBatch(make(), moduli, 3)
    exps = [rng.getrandbits(96) | 1 for _ in moduli]
    for i, e in enumerate(exps):
        batch.load(i, 1)
        batch.set_exponent(i, e)
    batch.run(40)  # part of the way
    n2 = rand_odd(rng, bits)
    batch.set_modulus(2, n2)
    assert batch.remaining(2) == 0 and batch.modulus(2) == n2
    moduli[2] = n2
    batch.load(2, 1)
    batch.set_exponent(2, exps[2])
    batch.run(1 << 20)
    for i, (n, e) in enumerate(zip(moduli, exps)):
        assert batch.remaining(i) == 0
        assert batch.read(i) == powmod(3, e, n), f"bits {bits} slot {i}"
    try:
        batch.set_modulus(1, rand_odd(rng, 4 * bits + 64))
    except MontmulError:
        pass
    else:
        raise AssertionError("a modulus of another transform length was accepted")


def is_one_many_check(make, bits, seed=12, prime=None):
    """ModBatch.is_one_many (is_one of batch.rs:179-183 for many slots at once) agrees with
    is_one and with the definition on every representative that matters: R mod N and
    R mod N + N (1), R mod N + 1, 0 and 2N - 1 (not 1), for slots in any order and repeated.
    With `prime` (a known prime of the batch's transform length), also after Fermat runs on
    composites and on that prime, admitted into a slot by set_modulus mid-batch (its values of
    1 follow the new modulus)."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, bits) for _ in range(6)]
    b = ModBatch(make(), moduli, 3)
    k = b.params().k
    reps = []
    for i, n in enumerate(moduli):
        r = (1 << (32 * k)) % n
        v = [r, r + n, r + 1, 0, 2 * n - 1, r][i]
        b.load_raw(i, v)
        reps.append(v in (r, r + n))
    order = [5, 0, 3, 1, 1, 4, 2]
    got = b.is_one_many(order)
    assert got == [reps[i] for i in order], (got, reps)
    assert got == [b.is_one(i) for i in order]
    if prime is not None:
        p = prime
        b.set_modulus(3, p)
        for i, n in enumerate(moduli[:3] + [p]):
            b.load(i, 1)
            b.set_exponent(i, n - 1)
        b.run(1 << 20)
        got = b.is_one_many(range(4))
        assert got == [b.is_one(i) for i in range(4)]
        assert got[3], "a prime modulus admitted by set_modulus came out composite"
        assert got[:3] == [powmod(3, n - 1, n) == 1 for n in moduli[:3]]
    try:
        b.is_one_many([len(moduli)])
    except MontmulError:
        pass
    else:
        raise AssertionError("a slot past the batch was accepted")


def mix_is_one_many_check(make, seed=13):
    """MixBatch.is_one_many after full Fermat tests over several lengths: Mersenne primes
    (2^p - 1, p = 127 .. 4423) and random composites of the same sizes, slots listed out of
    order and repeated; each answer equals is_one and the primality of the modulus."""
    rng = random.Random(seed)
    exps = [127, 521, 1279, 2203, 3217, 4423]
    moduli = []
    for p in exps:
        moduli.append((1 << p) - 1)
        moduli.append(rand_odd(rng, p))
    b = MixBatch(make(), moduli, 3)
    for i, n in enumerate(moduli):
        b.load(i, 1)
        b.set_exponent(i, n - 1)
    b.run(1 << 20)
    order = list(range(len(moduli)))
    rng.shuffle(order)
    order += order[:3]
    got = b.is_one_many(order)
    assert got == [b.is_one(i) for i in order]
    assert got == [i % 2 == 0 for i in order], got  # even slots hold the primes


def serve_mix(server, make, moduli, base=3, timeout_s=120.0):
    """Submits every modulus to `server` (a montmul_amd.Server with keep_raw), waits for all,
    and checks each result: prp == [base^(N-1) = 1 mod N] and the final raw limbs equal those of
    the same test run alone in a batch of its own on `make()` (alone-vs-mix equality)."""
    import time

    ids = {server.submit(n): n for n in moduli}
    got, t_end = {}, time.time() + timeout_s
    while len(got) < len(ids) and time.time() < t_end:
        for r in server.poll(256, 0.2):
            got[r["id"]] = r
    assert len(got) == len(ids), f"{len(ids) - len(got)} requests unfinished"
    for i, n in ids.items():
        r = got[i]
        assert r["prp"] == int(powmod(base, n - 1, n) == 1), f"request {i}: prp {r['prp']}"
        assert r["bits"] == (n - 1).bit_length()
        assert r["submit_s"] <= r["admit_s"] <= r["done_s"]
        alone = ModBatch(make(), [n], base)
        alone.load(0, 1)
        alone.set_exponent(0, n - 1)
        alone.run(1 << 30)
        assert server.raw(i) == alone.read_raw(0), f"request {i}: raw differs alone and in the mix"
    return got


def mixed_lengths(make, bit_sizes, ebits, chunk, seed=9, ref=None):
    """A MixBatch over moduli of the given sizes (several transform lengths) against plain
    per-length ModBatches, on the same backend or on `ref` (a factory, e.g. the oracle's closed
    form): every slot loaded with 1 or a random x, random exponents of 0 .. ebits bits (so slots
    finish at different steps), run(chunk) until no bits remain. Every slot: read ==
    x^(2^nbits) 3^e mod N, and raw limbs equal to its length's ModBatch. No reference
    counterpart (ModBatch holds one length, batch.rs:64-77)."""
    rng = random.Random(seed)
    moduli = [rand_odd(rng, b) for b in bit_sizes]
    xs, exps = [], []
    for i, n in enumerate(moduli):
        xs.append(1 if i % 3 else rng.randrange(n))
        nb = (0, 1, ebits)[i % 3] if i < 3 else rng.randrange(ebits // 2, ebits + 1)
        exps.append(rand_exponent(rng, nb) if nb else 0)
    mix = MixBatch(make(), moduli, 3)
    lengths = sorted({mix.params(i).len for i in range(len(moduli))})
    assert mix.groups() == len(lengths)
    for i in range(len(moduli)):
        mix.load(i, xs[i])
        mix.set_exponent(i, exps[i])
    rounds = 0
    while any(mix.remaining(i) for i in range(len(moduli))):
        mix.run(chunk)
        rounds += 1
        assert rounds <= ebits // chunk + 2
    for i, n in enumerate(moduli):
        nb = exps[i].bit_length()
        assert mix.read(i) == mulmod(powmod(xs[i], 1 << nb, n), powmod(3, exps[i], n), n), f"slot {i}, {n.bit_length()} bits"
    for length in lengths:
        idx = [i for i in range(len(moduli)) if mix.params(i).len == length]
        plain = ModBatch((ref or make)(), [moduli[i] for i in idx], 3)
        for j, i in enumerate(idx):
            plain.load(j, xs[i])
            plain.set_exponent(j, exps[i])
        plain.run(ebits)
        for j, i in enumerate(idx):
            assert plain.read_raw(j) == mix.read_raw(i), f"slot {i}, L = {length}"
        plain.free()
    with_err = False
    try:
        mix.read(len(moduli))
    except MontmulError:
        with_err = True
    assert with_err
    mix.free()
    return lengths


def serve_staggered(server, make, plan, base=3, timeout_s=120.0):
    """Submits plan = [(delay_s, N, e or None)] in order, sleeping delay_s before each submit
    (polling meanwhile), so that later requests arrive while earlier ones are on the device.
    e None is a Fermat test (e = N - 1). Checks every result: prp == [base^e = 1 mod N], the value
    raw R^-1 mod N == base^e mod N (R = 2^(32 k), k = L/2), and the raw limbs equal those of the
    same request run alone. Also checks that some request was submitted after another had been
    admitted and before it completed, and was admitted before that one completed (continuous
    batching: it joined work already on the device). Returns the results in plan order."""
    import time

    ids, got = [], {}
    t_end = time.time() + timeout_s

    def drain(wait):
        for r in server.poll(256, wait):
            got[r["id"]] = r

    for delay, n, e in plan:
        t_next = time.time() + delay
        while time.time() < t_next:
            drain(min(0.05, max(0.0, t_next - time.time())))
        ids.append(server.submit(n, e))
    while len(got) < len(ids) and time.time() < t_end:
        drain(0.2)
    assert len(got) == len(ids), f"{len(ids) - len(got)} requests unfinished"
    out = []
    for i, (delay, n, e) in zip(ids, plan):
        r = got[i]
        ee = n - 1 if e is None else e
        want = powmod(base, ee, n)
        assert r["prp"] == int(want == 1), f"request {i}: prp {r['prp']}"
        assert r["bits"] == ee.bit_length()
        assert r["submit_s"] <= r["admit_s"] <= r["done_s"]
        raw = server.raw(i)
        k = r["len"] // 2
        assert raw * powmod(1 << (32 * k), -1, n) % n == want, f"request {i}: value"
        alone = ModBatch(make(), [n], base)
        alone.load(0, 1)
        alone.set_exponent(0, ee)
        alone.run(1 << 30)
        assert raw == alone.read_raw(0), f"request {i}: raw differs alone and in the mix"
        out.append(r)
    joined = [(a, b) for a in out for b in out
              if a is not b and a["admit_s"] < b["submit_s"] < a["done_s"] and b["admit_s"] < a["done_s"]]
    assert joined, "no request arrived while another was running"
    return out
