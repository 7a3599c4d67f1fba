"""Writes the golden vectors in this directory with Python integers only.

Each vector fixes one Montgomery step by its closed form, which the reference's tests rely on
(`expected` in montmul/tests/common/mod.rs:17-21; the Gmp backend's REDC, montmul/src/gmp.rs:45-61):

    T  = x^2 * c
    m  = (T mod R) * N' mod R,   N' = -N^-1 mod R,   R = 2^(32k)
    x' = (T + m * N) / R          (an exact division; x' < 2N whenever x < 2N and R >= 16N)

with k = L/2 limbs (the wraparound path's k, montmul/src/batch.rs:79-81). The inputs mirror
the reference's cases: random moduli per transform-length bucket, the bucket's largest and
smallest bit counts (tests/common/mod.rs:125-131), N = 2^b - 1 and N = 2^b - 2^(b/2) + 1
(carry stress, tests/common/mod.rs:377-381), x in {0, N, 2N - 1, random} and c in {1, 3, 4}.
Exponentiation vectors give 3^e mod N; the tribonacci vectors give T(n) and whether it is a
base-3 Fermat probable prime (OEIS A092835 lists the primes, tests/common/mod.rs:284-288).

Run from the repository root:  python tests/golden/make_golden.py
"""

import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))


def params_for_bits(bits):
    """k and L as montmul/src/batch.rs:20-25, then k widened to L/2 (batch.rs:79-81)."""
    k = (bits + 4 + 31) // 32
    length = 8
    while length < 2 * k:
        length *= 2
    return length // 2, length


def step(x, n, c, k):
    r_bits = 32 * k
    r = 1 << r_bits
    nprime = (-pow(n, -1, r)) % r
    t = x * x * c
    m = ((t & (r - 1)) * nprime) & (r - 1)
    s = t + m * n
    assert s % r == 0
    return s >> r_bits


def rand_odd(rng, bits):
    n = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
    return n


def tribonacci(n):
    """T(n) with T(0) = T(1) = 0, T(2) = 1 (OEIS A000073; montmul/src/seq.rs:6-15)."""
    a, b, c = 0, 0, 1
    for _ in range(n):
        a, b, c = b, c, a + b + c
    return a


def main():
    rng = random.Random(20260)
    steps = []
    for e in range(3, 12):
        length = 1 << e
        kmax = length // 2
        sizes = [32 * kmax - 4, max(2, 32 * (kmax // 2) - 3)]
        for bits in sizes:
            k, L = params_for_bits(bits)
            assert L == length, (bits, L, length)
            moduli = [rand_odd(rng, bits), (1 << bits) - 1]
            if bits >= 8:
                moduli.append((1 << bits) - (1 << (bits // 2)) + 1)
            # Every x and c at small L; the largest x and a random one at larger L (size).
            small = length <= 128
            for n in moduli:
                for which in (range(4) if small else (2, 3)):
                    x = [0, n, 2 * n - 1, rng.randrange(2 * n)][which]
                    for c in ((1, 3, 4) if small else ((4,) if which == 2 else (1, 3))):
                        steps.append({"bits": bits, "L": L, "k": k, "n": hex(n), "x": hex(x), "c": c,
                                      "out": hex(step(x, n, c, k))})
    # Two cases at the 100,000-bit bucket (L = 8,192).
    for _ in range(2):
        n = rand_odd(rng, 100_000)
        k, L = params_for_bits(100_000)
        x = rng.randrange(2 * n)
        steps.append({"bits": 100_000, "L": L, "k": k, "n": hex(n), "x": hex(x), "c": 3,
                      "out": hex(step(x, n, 3, k))})

    # Chains: from x = R mod N (the Montgomery form of 1), one step per exponent bit, MSB
    # first, c = 3 on one bits; the end value is 3^e * R mod N up to adding N.
    chains = []
    for bits, ebits in [(64, 40), (1000, 60), (5000, 24)]:
        n = rand_odd(rng, bits)
        k, L = params_for_bits(bits)
        e = rng.getrandbits(ebits) | (1 << (ebits - 1))
        r = 1 << (32 * k)
        x = r % n
        raws = []
        for p in range(ebits):
            bit = (e >> (ebits - 1 - p)) & 1
            x = step(x, n, 3 if bit else 1, k)
            raws.append(hex(x))
        assert x % n == pow(3, e, n) * r % n
        chains.append({"bits": bits, "L": L, "k": k, "n": hex(n), "e": hex(e), "raw": raws,
                       "pow": hex(pow(3, e, n))})

    tri = []
    for idx in [6, 7, 8, 9, 10, 11, 12, 13, 20, 86, 87, 88, 98, 100, 200, 215, 802, 803, 1001]:
        t = tribonacci(idx)
        if t % 2 == 0 or t <= 3:
            continue
        tri.append({"index": idx, "t": hex(t), "prp": pow(3, t - 1, t) == 1})

    with open(os.path.join(HERE, "steps.json"), "w") as f:
        json.dump(steps, f, separators=(",", ":"))
    with open(os.path.join(HERE, "chains.json"), "w") as f:
        json.dump(chains, f, separators=(",", ":"))
    with open(os.path.join(HERE, "tribonacci.json"), "w") as f:
        json.dump(tri, f, indent=0)
    print(f"steps {len(steps)}, chains {len(chains)}, tribonacci {len(tri)}")


if __name__ == "__main__":
    main()
