"""Pins the CPU oracle (oracle/mm_oracle.cpp) before anything is compared against it.

Checks: the golden step vectors written with Python integers (tests/golden), GMP's closed
form, generator orders and root choices (field.rs:111-120, 237-267, 296-308), CRT round
trips (field.rs:310-316), a brute-force O(n^2) DFT over GF(p^2) at small n (the DIF output in
bit-reversed order), the exactness closed form and batched transform products
(common/mod.rs:23-91), the partner positions of the half path, and known PRP answers.
"""

import json
import os
import random

import numpy as np
import pytest

import common
from montmul_amd import ModBatch, fermat
from oracle_backends import Cpu, CpuTransform, Gmp, gmp_step, olib, t_download, t_forward, t_inverse, t_pointwise, t_upload
import ctypes as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P1, P2 = 2**31 - 1, 2**61 - 1


def gmul(a, b, p):
    return ((a[0] * b[0] - a[1] * b[1]) % p, (a[0] * b[1] + a[1] * b[0]) % p)


def gpow(a, e, p):
    r = (1, 0)
    while e:
        if e & 1:
            r = gmul(r, a, p)
        a = gmul(a, a, p)
        e >>= 1
    return r


def sylow():
    g1 = (C.c_uint32 * 2)()
    g2 = (C.c_uint64 * 2)()
    olib.oc_sylow2(g1, g2)
    return (g1[0], g1[1]), (g2[0], g2[1])


def test_sylow_orders_and_smallest_choice():
    g1, g2 = sylow()
    for g, p, e, cof in ((g1, P1, 32, 2**30 - 1), (g2, P2, 62, 2**60 - 1)):
        assert gpow(g, 1 << e, p) == (1, 0)
        assert gpow(g, 1 << (e - 1), p) != (1, 0)
        # The generator is (a + i)^cofactor for the smallest a >= 1 of full order.
        a = 1
        while gpow(gpow((a, 1), cof, p), 1 << (e - 1), p) == (1, 0):
            a += 1
        assert gpow((a, 1), cof, p) == g


def test_crt_roundtrip():
    inv = olib.oc_crt_inv()
    assert inv * (P2 % P1) % P1 == 1
    out = (C.c_uint64 * 2)()
    Q = P1 * P2
    for c in [0, 1, Q - 1, Q // 3, 123456789012345678901234567 % Q]:
        olib.oc_crt(c % P1, c % P2, out)
        assert out[0] | (out[1] << 64) == c


@pytest.mark.parametrize("n", [2, 4, 8, 64, 1024])
def test_tables_roots(n):
    a = [np.zeros(n, dtype=np.uint32) if i < 2 else np.zeros(2 * n, dtype=np.uint32) for i in range(4)]
    tw1, itw1, w1, iw1 = (np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(2 * n, np.uint32), np.zeros(2 * n, np.uint32))
    tw2, itw2, w2, iw2 = (np.zeros(n, np.uint64), np.zeros(n, np.uint64), np.zeros(2 * n, np.uint64), np.zeros(2 * n, np.uint64))
    p32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint32))
    p64 = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint64))
    olib.oc_tables(n, p32(tw1), p32(itw1), p32(w1), p32(iw1), p64(tw2), p64(itw2), p64(w2), p64(iw2))
    for tw, itw, w, iw, p in ((tw1, itw1, w1, iw1, P1), (tw2, itw2, w2, iw2, P2)):
        if n >= 4:
            omega = (int(tw[2]), int(tw[3]))
            assert gpow(omega, n, p) == (1, 0) and gpow(omega, n // 2, p) != (1, 0)
        zeta = (int(w[2]), int(w[3]))
        assert gpow(zeta, n, p) == (0, 1), "zeta^n = i"
        ninv = pow(n, -1, p)
        for j in range(n):
            prod = gmul((int(w[2 * j]), int(w[2 * j + 1])), (int(iw[2 * j]), int(iw[2 * j + 1])), p)
            assert prod == (ninv, 0)
        for t in range(n // 2):
            assert gmul((int(tw[2 * t]), int(tw[2 * t + 1])), (int(itw[2 * t]), int(itw[2 * t + 1])), p) == (1, 0)


@pytest.mark.parametrize("n", [16, 256, 16384])
def test_split_stage_identity(n):
    """The GPU's register step at n = 2^14 runs a length-n DIF as one radix-2 stage and two
    DIFs of n/2: A_j = a_j + a_{j+n/2}, B_j = (a_j - a_{j+n/2}) omega^j with omega = g^(2^(E -
    log n)) (the generator g of order 2^E; the host's split_roots), and DIF_n(a) in bit-reversed
    order is DIF_{n/2}(A) followed by DIF_{n/2}(B). Checked against the oracle's own DIFs of n
    and n/2, both fields; at n = 16 the oracle's DIF is itself pinned by the brute-force DFT."""
    rng = random.Random(n)
    g1, g2 = sylow()
    h = n // 2
    bits = n.bit_length() - 1
    p32 = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint32))
    p64 = lambda x: x.ctypes.data_as(C.POINTER(C.c_uint64))
    f1 = np.array([rng.randrange(P1) for _ in range(2 * n)], dtype=np.uint32)
    f2 = np.array([rng.randrange(P2) for _ in range(2 * n)], dtype=np.uint64)
    halves = []
    for g, e, f, p, dt in ((g1, 32, f1, P1, np.uint32), (g2, 62, f2, P2, np.uint64)):
        omega = gpow(g, 1 << (e - bits), p)
        a = [(int(f[2 * j]), int(f[2 * j + 1])) for j in range(n)]
        A, B, w = [], [], (1, 0)
        for j in range(h):
            lo, hi = a[j], a[j + h]
            A.append(((lo[0] + hi[0]) % p, (lo[1] + hi[1]) % p))
            B.append(gmul(((lo[0] - hi[0]) % p, (lo[1] - hi[1]) % p), w, p))
            w = gmul(w, omega, p)
        halves.append([np.array([c for z in v for c in z], dtype=dt) for v in (A, B)])
    olib.oc_dif(p32(f1), p64(f2), n)
    for half in range(2):
        s1, s2 = halves[0][half], halves[1][half]
        olib.oc_dif(p32(s1), p64(s2), h)
        assert np.array_equal(s1, f1[half * n:(half + 1) * n]), f"M31 half {half}"
        assert np.array_equal(s2, f2[half * n:(half + 1) * n]), f"M61 half {half}"


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32])
def test_dif_is_brute_force_dft_in_bit_reversed_order(n):
    rng = random.Random(n)
    f1 = np.array([rng.randrange(P1) for _ in range(2 * n)], dtype=np.uint32)
    f2 = np.array([rng.randrange(P2) for _ in range(2 * n)], dtype=np.uint64)
    x1 = [(int(f1[2 * j]), int(f1[2 * j + 1])) for j in range(n)]
    x2 = [(int(f2[2 * j]), int(f2[2 * j + 1])) for j in range(n)]
    olib.oc_dif(f1.ctypes.data_as(C.POINTER(C.c_uint32)), f2.ctypes.data_as(C.POINTER(C.c_uint64)), n)
    g1, g2 = sylow()
    bits = n.bit_length() - 1
    rev = lambda p: int(format(p, f"0{bits}b")[::-1], 2) if bits else 0
    for g, e, x, out, p in ((g1, 32, x1, f1, P1), (g2, 62, x2, f2, P2)):
        omega = gpow(g, 1 << (e - bits), p)
        for pos in range(n):
            kk = rev(pos)
            acc = (0, 0)
            for j in range(n):
                t = gmul(x[j], gpow(omega, j * kk, p), p)
                acc = ((acc[0] + t[0]) % p, (acc[1] + t[1]) % p)
            assert (int(out[2 * pos]), int(out[2 * pos + 1])) == acc, f"n {n} position {pos}"


def test_exactness_closed_form():
    """All-ones limbs at k = L/2, c = 4 (common/mod.rs:23-53)."""
    common.exactness(CpuTransform(), 1 << 14)


def test_transform_batched_products():
    """Three vectors per call with scales [1, 3, 4] and none (common/mod.rs:55-91)."""
    common.transform_batched(CpuTransform(), 1 << 11)


@pytest.mark.parametrize("h", [2, 4, 8, 256])
def test_half_partner(h):
    out = (C.c_uint64 * h)()
    olib.oc_half_partner(h, out)
    bits = h.bit_length() - 1
    rev = lambda p: int(format(p, f"0{bits}b")[::-1], 2)
    for p in range(h):
        assert out[p] == rev((h - rev(p)) % h)
        # Closed form used by the device kernels.
        want = 0 if p == 0 else (3 << (p.bit_length() - 1)) - 1 - p
        assert out[p] == want


def _golden_steps():
    with open(os.path.join(GOLD, "steps.json")) as f:
        return json.load(f)


def test_golden_steps_gmp_closed_form():
    for v in _golden_steps():
        n, x, c, k = int(v["n"], 16), int(v["x"], 16), v["c"], v["k"]
        r = 1 << (32 * k)
        nprime = (-pow(n, -1, r)) % r
        assert gmp_step(x, n, nprime, k, c) == int(v["out"], 16)


@pytest.mark.parametrize("wrap", [True, False])
def test_golden_steps_cpu(wrap):
    """The Cpu restatement, full-k with and without the wraparound path, on every vector."""
    groups = {}
    for v in _golden_steps():
        groups.setdefault((v["L"], v["c"]), []).append(v)
    for (L, c), vs in groups.items():
        make = (lambda: Cpu(wraparound=True)) if wrap else (lambda: Cpu(full_k=True))
        moduli = [int(v["n"], 16) for v in vs]
        b = ModBatch(make(), moduli, max(c, 1))
        assert b.params().k == vs[0]["k"]
        for i, v in enumerate(vs):
            b.load_raw(i, int(v["x"], 16))
            b.set_exponent(i, 1)
        b.run(1)
        for i, v in enumerate(vs):
            assert b.read_raw(i) == int(v["out"], 16), f"L {L} c {c} bits {v['bits']}"


def test_golden_chains():
    with open(os.path.join(GOLD, "chains.json")) as f:
        chains = json.load(f)
    for ch in chains:
        n, e = int(ch["n"], 16), int(ch["e"], 16)
        # The raw chain uses k = L/2 (the wraparound path's k); Gmp keeps the minimal k
        # (gmp.rs has no full_k), so only its canonical value is compared.
        b = ModBatch(Cpu(wraparound=True), [n], 3)
        b.load(0, 1)
        b.set_exponent(0, e)
        for raw in ch["raw"]:
            b.run(1)
            assert b.read_raw(0) == int(raw, 16)
        assert b.read(0) == int(ch["pow"], 16)
        g = ModBatch(Gmp(), [n], 3)
        g.load(0, 1)
        g.set_exponent(0, e)
        g.run(len(ch["raw"]))
        assert g.read(0) == int(ch["pow"], 16)


def test_golden_tribonacci_prp():
    with open(os.path.join(GOLD, "tribonacci.json")) as f:
        tri = json.load(f)
    ts = [int(t["t"], 16) for t in tri]
    got = [fermat(Cpu(), [t], 3, 1000)[0] for t in ts]
    assert got == [t["prp"] for t in tri]
    assert any(t["prp"] for t in tri) and not all(t["prp"] for t in tri)


def test_wraparound_matches_full_product():
    """Same k = L/2 on both sides: equal raw limbs (tests/cpu.rs:111-121)."""
    for bits in [64, 1000, 20_000]:
        common.assert_same_raw(lambda: Cpu(full_k=True), lambda: Cpu(wraparound=True), bits, 4, 40)
    # Both with the minimal k (tests/gmp.rs:64-66).
    common.assert_same_raw(Gmp, Cpu, 3000, 4, 30)


def test_oracle_flags_carry_out():
    """x = R - 1 lies outside [0, 2N): the squaring carries out and the error names the slot
    (tests/gpu.rs:269-305 on the GPU; the oracle carries the same error word)."""
    from montmul_amd import CarryError
    rng = random.Random(9)
    moduli = [common.rand_odd(rng, 1000), common.rand_odd(rng, 1000)]
    b = ModBatch(Cpu(wraparound=True), moduli, 4)
    k = b.params().k
    for i in range(2):
        b.load(i, 1)
        b.set_exponent(i, 1)
    import ctypes
    st = b.state
    buf = (ctypes.c_uint32 * k)(*([0xFFFFFFFF] * k))
    assert olib.oc_load(st, 1, buf) == 0
    with pytest.raises(CarryError) as ei:
        b.run(1)
    assert ei.value.slot == 1 and "slot 1" in str(ei.value)


def test_cpu_case_4096_bits_oracle():
    """The reference's CPU-runnable case on the oracle's Cpu restatement (wraparound, k = L/2):
    64 moduli of 4,096 bits from the score workload seeded 2026, e = (N - 1) >> 3,584; every slot's
    read equals 3^e mod N and its raw limbs the GMP closed-form chain."""
    import montmul_amd as mm

    moduli, exps = mm.score_workload(4096, 64, 512, 2026)
    a, b = ModBatch(Cpu(wraparound=True), moduli, 3), ModBatch(Gmp(full_k=True), moduli, 3)
    for bb in (a, b):
        for i, e in enumerate(exps):
            bb.load(i, 1)
            bb.set_exponent(i, e)
        bb.run(512)
    for i, (n, e) in enumerate(zip(moduli, exps)):
        assert a.read(i) == pow(3, e, n), f"slot {i}"
        assert a.read_raw(i) == b.read_raw(i), f"slot {i} raw"


@pytest.mark.parametrize("wrap", [True, False])
def test_set_modulus_oracle(wrap):
    """Admission of a new modulus into a running batch, on the Cpu restatement."""
    for bits in (60, 1000, 5000):
        common.set_modulus_admission(lambda: Cpu(wraparound=wrap), bits)


def test_set_modulus_gmp_oracle():
    common.set_modulus_admission(lambda: Gmp(), 1000)


def test_mixbatch_oracle():
    """Mixed lengths on the Cpu restatement (no joint run: the groups run one after another)."""
    lengths = common.mixed_lengths(lambda: Cpu(wraparound=True), [64, 200, 300, 700, 1500, 90, 640], 40, 9)
    assert lengths == [8, 16, 32, 64, 128]
