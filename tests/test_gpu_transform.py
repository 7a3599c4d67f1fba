"""The transform layer on the GPU against the oracle's Cpu restatement (tests/gpu.rs:26-90).

Spectra are compared element for element (raw residues, the reference's bit-reversed order),
products against integer multiplication and the exactness closed form.
"""

import numpy as np
import pytest

import common
from oracle_backends import CpuTransform, t_forward, t_pointwise, t_upload

pytestmark = pytest.mark.gpu


class GpuTransform:
    """Adapter over montmul_amd.Hip's transform layer for the common checks."""

    def __init__(self, hip):
        self.h = hip

    def upload(self, coeffs, length, count):
        return self.h.upload(np.ascontiguousarray(coeffs, dtype=np.uint32), length, count)

    def forward(self, b, length, count):
        self.h.forward(b, length, count)

    def inverse(self, b, length, count):
        self.h.inverse(b, length, count)

    def pointwise(self, out, a, b, scales, length, count):
        self.h.pointwise(out, a, b, None if scales is None else self.h.upload_scales(scales), length, count)

    def download_words(self, b):
        return np.frombuffer(self.h.download_words(b), dtype=np.uint64)


def raw(hip, b):
    f1, f2 = hip.download_raw(b)
    return np.frombuffer(f1, dtype=np.uint32), np.frombuffer(f2, dtype=np.uint64)


def test_transform_matches_cpu_bit_for_bit(hip):
    """tests/gpu.rs:26-44: forward spectra raw-equal, round trip equal, lengths 2^3 .. 2^20."""
    rng = np.random.default_rng(26)
    ct = CpuTransform()
    for e in range(3, 21):
        length, count = 1 << e, 3
        coeffs = np.zeros(length * count, dtype=np.uint32)
        for v in range(count):
            coeffs[v * length:v * length + length // 2] = rng.integers(0, 2**32, size=length // 2, dtype=np.uint64)
        gx = hip.upload(coeffs, length, count)
        cx = t_upload(coeffs, length, count)
        hip.forward(gx, length, count)
        t_forward(cx)
        g1, g2 = raw(hip, gx)
        assert np.array_equal(g1, cx.f1) and np.array_equal(g2, cx.f2), f"forward, len {length}"
        hip.inverse(gx, length, count)
        ct.inverse(cx, length, count)
        gw = np.frombuffer(hip.download_words(gx), dtype=np.uint64)
        assert np.array_equal(gw, ct.download_words(cx)), f"round trip, len {length}"
        # the round trip returns the input coefficients
        assert np.array_equal(gw.reshape(-1, 2)[:, 0], coeffs.astype(np.uint64)) and not gw.reshape(-1, 2)[:, 1].any()


def test_exactness_bound(hip):
    """tests/gpu.rs:46-50: exactness(gpu, 2^18)."""
    common.exactness(GpuTransform(hip), 1 << 18)


def test_transform_count_and_scales(hip):
    """tests/gpu.rs:52-56: transform_batched(gpu, 2^16)."""
    common.transform_batched(GpuTransform(hip), 1 << 16)


def test_pointwise_scales_above_m31_match_cpu(hip):
    """tests/gpu.rs:58-73: scales at and above M31 reduce as the Cpu's do."""
    rng = np.random.default_rng(58)
    length, count = 64, 4
    scales = [0x7FFFFFFF, 0x80000004, 0xFFFFFFFE, 0xFFFFFFFF]
    coeffs = rng.integers(0, 2**32, size=length * count, dtype=np.uint64).astype(np.uint32)
    gx, cx = hip.upload(coeffs, length, count), t_upload(coeffs, length, count)
    hip.forward(gx, length, count)
    t_forward(cx)
    zeros = np.zeros(length * count, dtype=np.uint32)
    go, co = hip.upload(zeros, length, count), t_upload(zeros, length, count)
    hip.pointwise(go, gx, gx, hip.upload_scales(scales), length, count)
    t_pointwise(co, cx, cx, scales)
    g1, g2 = raw(hip, go)
    assert np.array_equal(g1, co.f1) and np.array_equal(g2, co.f2)


def test_load_forward_matches_upload_forward(hip):
    """tests/gpu.rs:75-90: load_forward of k = L/2 - 1 limbs equals upload + forward of the padded
    vectors."""
    rng = np.random.default_rng(75)
    for e in (3, 8, 13):
        length, count = 1 << e, 5
        k = length // 2 - 1
        limbs = rng.integers(0, 2**32, size=k * count, dtype=np.uint64).astype(np.uint32)
        padded = np.zeros(length * count, dtype=np.uint32)
        for v in range(count):
            padded[v * length:v * length + k] = limbs[v * k:(v + 1) * k]
        cx = t_upload(padded, length, count)
        t_forward(cx)
        g1, g2 = raw(hip, hip.load_forward(limbs, k, length, count))
        assert np.array_equal(g1, cx.f1) and np.array_equal(g2, cx.f2), f"len {length}"


def test_mismatched_length_is_error(hip):
    """Gpu asserts that a buffer's length matches the call's (gpu/mod.rs:300-308)."""
    import montmul_amd as mm

    b = hip.upload(np.zeros(16, dtype=np.uint32), 16, 1)
    with pytest.raises(mm.MontmulError):
        hip.forward(b, 32, 1)
