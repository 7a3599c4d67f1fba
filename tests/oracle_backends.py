"""Test-side access to the CPU oracle (oracle/_build/libmmoracle.so): the restated `Cpu` and
`Gmp` backends as `mm_backend` tables, so the host API drives them exactly as it drives the
HIP backend, plus the oracle's transform-layer and pinning entry points.

Only tests, the smoke check and the bench's CPU-baseline leg use this module.
"""

import ctypes as C
import os
import subprocess

import montmul_amd as mm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "libmmoracle.so")


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def _load():
    if not os.path.exists(ORACLE_LIB):
        build_oracle()
    L = C.CDLL(ORACLE_LIB)
    u32p, u64p, vp = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_void_p
    sig = {
        "oc_limits": (None, [vp, u64p, u64p, u32p, u32p, C.POINTER(C.c_int32)]),
        "oc_t_upload": (None, [u32p, C.c_size_t, C.c_size_t, u32p, u64p]),
        "oc_t_download": (None, [u32p, u64p, C.c_size_t, C.c_size_t, u64p]),
        "oc_t_forward": (None, [u32p, u64p, C.c_size_t, C.c_size_t]),
        "oc_t_inverse": (None, [u32p, u64p, C.c_size_t, C.c_size_t]),
        "oc_t_pointwise": (None, [u32p, u64p, u32p, u64p, u32p, u64p, u32p, C.c_size_t, C.c_size_t]),
        "oc_sylow2": (None, [u32p, u64p]),
        "oc_tables": (None, [C.c_size_t, u32p, u32p, u32p, u32p, u64p, u64p, u64p, u64p]),
        "oc_crt_inv": (C.c_uint32, []),
        "oc_crt": (None, [C.c_uint32, C.c_uint64, u64p]),
        "oc_mul1": (None, [C.c_uint32, C.c_uint32, u32p]),
        "oc_mul2": (None, [C.c_uint64, C.c_uint64, u64p]),
        "oc_cmul1": (None, [u32p, u32p, u32p]),
        "oc_cmul2": (None, [u64p, u64p, u64p]),
        "oc_half_partner": (None, [C.c_size_t, u64p]),
        "oc_half_factors": (None, [u32p, C.c_size_t, u32p, u64p]),
        "oc_dif": (None, [u32p, u64p, C.c_size_t]),
        "og_step": (None, [u32p, u32p, u32p, C.c_size_t, C.c_uint32, u32p]),
        "oc_load": (C.c_int, [vp, C.c_size_t, u32p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


olib = _load()


class _CpuConfig(C.Structure):
    _fields_ = [("wraparound", C.c_int32), ("full_k", C.c_int32), ("threads", C.c_int32)]


class _GmpConfig(C.Structure):
    _fields_ = [("threads", C.c_int32), ("full_k", C.c_int32)]


def _table(prefix, cfg, bound, min_len, max_len, full_k):
    t = mm.BackendTable()
    t.self = C.cast(C.pointer(cfg), C.c_void_p)
    t.create = mm._CREATE((prefix + "create", olib))
    t.destroy = mm._DESTROY((prefix + "destroy", olib))
    t.load = mm._LOAD((prefix + "load", olib))
    t.set_exponent = mm._SETEXP((prefix + "set_exponent", olib))
    t.run = mm._RUN((prefix + "run", olib))
    t.mul = mm._MUL((prefix + "mul", olib))
    t.read = mm._READ((prefix + "read", olib))
    t.remaining = mm._REMAINING((prefix + "remaining", olib))
    t.set_modulus = mm._SETMOD((prefix + "set_modulus", olib))
    if hasattr(olib, prefix + "failed_slots"):
        t.failed_slots = mm._FAILED((prefix + "failed_slots", olib))
    t.last_error = mm._LASTERR(("oc_last_error", olib))
    t.bound_lo, t.bound_hi = bound & (2**64 - 1), bound >> 64
    t.min_len, t.max_len, t.full_k = min_len, max_len, int(full_k)
    return t


class Cpu(mm.Backend):
    """The oracle's restatement of the reference `Cpu` backend (cpu.rs:9-47, 499-578)."""

    def __init__(self, wraparound=False, full_k=False, threads=0):
        self._cfg = _CpuConfig(int(wraparound), int(full_k), threads)
        lo, hi, mn, mx, fk = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32(), C.c_int32()
        olib.oc_limits(C.byref(self._cfg), C.byref(lo), C.byref(hi), C.byref(mn), C.byref(mx), C.byref(fk))
        self.table = _table("oc_", self._cfg, lo.value | (hi.value << 64), mn.value, mx.value, fk.value)

    @staticmethod
    def with_wraparound(on=True):
        return Cpu(wraparound=on)


class Gmp(mm.Backend):
    """The oracle's restatement of the reference `Gmp` backend (gmp.rs:1-122): exact, so the
    coefficient bound is u128::MAX (gmp.rs:20-23)."""

    def __init__(self, threads=0, full_k=False):
        # full_k widens k to L/2 like the wraparound path, so that raw limbs are the closed
        # form with the same R as the GPU's: a test-side option, not in the reference's Gmp.
        self._cfg = _GmpConfig(threads, int(full_k))
        self.table = _table("og_", self._cfg, 2**128 - 1, 8, 1 << 24, full_k)


# ---- Transform layer (cpu.rs:349-407) on numpy arrays ---------------------------------------
import numpy as np


def _p32(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


class CpuBuf:
    def __init__(self, length, count):
        n = length // 2
        self.len, self.count = length, count
        self.f1 = np.zeros(2 * n * count, dtype=np.uint32)
        self.f2 = np.zeros(2 * n * count, dtype=np.uint64)


def t_upload(coeffs, length, count):
    b = CpuBuf(length, count)
    c = np.ascontiguousarray(coeffs, dtype=np.uint32)
    olib.oc_t_upload(_p32(c), length, count, _p32(b.f1), _p64(b.f2))
    return b


def t_download(b):
    out = np.zeros(2 * b.len * b.count, dtype=np.uint64)
    olib.oc_t_download(_p32(b.f1), _p64(b.f2), b.len, b.count, _p64(out))
    return [int(out[2 * i]) | (int(out[2 * i + 1]) << 64) for i in range(b.len * b.count)]


def t_forward(b):
    olib.oc_t_forward(_p32(b.f1), _p64(b.f2), b.len, b.count)


def t_inverse(b):
    olib.oc_t_inverse(_p32(b.f1), _p64(b.f2), b.len, b.count)


def t_pointwise(out, a, b, scales):
    s = None if scales is None else np.ascontiguousarray(scales, dtype=np.uint32)
    olib.oc_t_pointwise(_p32(out.f1), _p64(out.f2), _p32(a.f1), _p64(a.f2), _p32(b.f1), _p64(b.f2),
                        None if s is None else _p32(s), out.len, out.count)


def gmp_step(x, n, nprime, k, c):
    """One closed-form step x' = (x^2 c + m N) / R through the oracle's GMP code."""
    xs, ns, ps = (np.frombuffer(v.to_bytes(4 * k, "little"), dtype=np.uint32).copy() for v in (x, n, nprime))
    out = np.zeros(k, dtype=np.uint32)
    olib.og_step(_p32(xs), _p32(ns), _p32(ps), k, c, _p32(out))
    return int.from_bytes(out.tobytes(), "little")


def _limbs(v, k=None):
    k = k if k is not None else max(1, (v.bit_length() + 31) // 32)
    return np.frombuffer(v.to_bytes(4 * k, "little"), dtype=np.uint32).copy()


def gmp_powm(b, e, m):
    """b^e mod m by GMP's mpz_powm (b, e >= 0)."""
    bs, es, ms = _limbs(b), _limbs(e), _limbs(m)
    out = np.zeros(len(ms), dtype=np.uint32)
    olib.og_powm(_p32(bs), len(bs), _p32(es), len(es), _p32(ms), len(ms), _p32(out))
    return int.from_bytes(out.tobytes(), "little")


def gmp_mulmod(a, b, m):
    """a b mod m by GMP (a, b >= 0)."""
    as_, bs, ms = _limbs(a), _limbs(b), _limbs(m)
    out = np.zeros(len(ms), dtype=np.uint32)
    olib.og_mulmod(_p32(as_), len(as_), _p32(bs), len(bs), _p32(ms), len(ms), _p32(out))
    return int.from_bytes(out.tobytes(), "little")


def gmp_invert(a, m):
    """a^-1 mod m by GMP's mpz_invert; ValueError when there is none."""
    as_, ms = _limbs(a), _limbs(m)
    out = np.zeros(len(ms), dtype=np.uint32)
    if not olib.og_invert(_p32(as_), len(as_), _p32(ms), len(ms), _p32(out)):
        raise ValueError("not invertible")
    return int.from_bytes(out.tobytes(), "little")


class CpuTransform:
    """The oracle's transform layer (Cpu's Transform impl, cpu.rs:349-407) behind the adapter
    interface of common.exactness / common.transform_batched."""

    def upload(self, coeffs, length, count):
        return t_upload(coeffs, length, count)

    def forward(self, b, length, count):
        assert (b.len, b.count) == (length, count)
        t_forward(b)

    def inverse(self, b, length, count):
        assert (b.len, b.count) == (length, count)
        t_inverse(b)

    def pointwise(self, out, a, b, scales, length, count):
        t_pointwise(out, a, b, scales)

    def download_words(self, b):
        out = np.zeros(2 * b.len * b.count, dtype=np.uint64)
        olib.oc_t_download(_p32(b.f1), _p64(b.f2), b.len, b.count, _p64(out))
        return out

    def download_raw(self, b):
        return b.f1, b.f2
