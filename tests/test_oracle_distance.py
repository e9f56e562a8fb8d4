"""Pins the oracle's squared-L2 (the restatement of src/metric.cpp:45-78).

1. Bit-equal to avx_pin.cpp: the reference's AVX construction (tail blocks
   first, 8 lanes, lanes summed left to right) compiled the reference's way
   (g++ -O3 -mavx2 -mfma, contraction on), for dimensions hitting every tail
   case.  The same kernel built with contraction off must differ somewhere,
   so the comparison can tell the fused from the unfused order.
2. Exact on integer data in [0, 255] at d = 128 (all partial sums < 2^24).
3. Within 1e-6 relative of FP64 on float data.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from conftest import ROOT


def _pin(name):
    path = os.path.join(ROOT, "oracle", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built")
    lib = ctypes.CDLL(path)
    lib.pin_l2sqr_avx.restype = ctypes.c_float
    lib.pin_l2sqr_avx.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    return lib


@pytest.mark.parametrize("d", [8, 16, 24, 40, 96, 100, 128, 200, 256, 960])
def test_matches_reference_avx_kernel_bitwise(d, rng):
    pin = _pin("libfkpin.so")
    for trial in range(200):
        scale = [1e-3, 1.0, 37.0, 1e4][trial % 4]
        a = oracle.pad8(rng.normal(size=d).astype(np.float32) * scale)
        b = oracle.pad8(rng.normal(size=d).astype(np.float32) * scale)
        ref = pin.pin_l2sqr_avx(a.ctypes.data, b.ctypes.data, d)
        got = oracle.l2(a, b)
        assert np.float32(ref).view(np.uint32) == np.float32(got).view(np.uint32), (d, trial, ref, got)


def test_unfused_order_is_distinguishable(rng):
    fused = _pin("libfkpin.so")
    plain = _pin("libfkpin_nofma.so")
    diffs = 0
    for _ in range(500):
        a = rng.normal(size=128).astype(np.float32)
        b = rng.normal(size=128).astype(np.float32)
        diffs += fused.pin_l2sqr_avx(a.ctypes.data, b.ctypes.data, 128) != plain.pin_l2sqr_avx(
            a.ctypes.data, b.ctypes.data, 128)
    assert diffs > 0


def test_integer_data_is_exact(rng):
    x = rng.integers(0, 256, size=(200, 128)).astype(np.float32)
    a = rng.integers(0, 200, 5000)
    b = rng.integers(0, 200, 5000)
    d32, d64 = oracle.pair_distances(x, a, b, with_f64=True)
    exact = ((x[a].astype(np.int64) - x[b].astype(np.int64)) ** 2).sum(1)
    assert np.array_equal(d64, exact.astype(np.float64))
    assert np.array_equal(d32.astype(np.float64), exact.astype(np.float64))


@pytest.mark.parametrize("d", [20, 100, 128])
def test_shifted_byte_gram_identity(d, rng):
    """The byte-row Gram join: rows of values in [0, 255], zero-padded to
    dpad (a multiple of 8) and shifted to signed bytes b ^ 0x80 = b - 128,
    give |x - y|^2 = n'(x) + n'(y) - 2 (x - 128).(y - 128) with n'(x) the sum
    over all dpad shifted bytes squared -- equal to the canonical FP32
    distance of the unpadded rows, and unchanged by one permutation of the
    dimensions applied to both rows (the chain-major layout)."""
    dpad = (d + 7) // 8 * 8
    x = rng.integers(0, 256, size=(300, d)).astype(np.uint8)
    xp = np.zeros((300, dpad), np.uint8)
    xp[:, :d] = x
    s = (xp ^ 0x80).view(np.int8).astype(np.int64)
    assert np.array_equal(s, xp.astype(np.int64) - 128)
    n8 = (s * s).sum(1)
    a = rng.integers(0, 300, 4000)
    b = rng.integers(0, 300, 4000)
    perm = rng.permutation(dpad)
    got = n8[a] + n8[b] - 2 * (s[a] * s[b]).sum(1)
    got_perm = n8[a] + n8[b] - 2 * (s[a][:, perm] * s[b][:, perm]).sum(1)
    exact = ((x[a].astype(np.int64) - x[b].astype(np.int64)) ** 2).sum(1)
    assert np.array_equal(got, exact) and np.array_equal(got_perm, exact)
    assert exact.max() < 2 ** 24  # converts to FP32 without rounding
    d32 = oracle.pair_distances(x.astype(np.float32), a, b)
    assert np.array_equal(d32, got.astype(np.float32))


@pytest.mark.parametrize("d", [96, 128, 960])
def test_float_data_within_tolerance_of_fp64(d, rng):
    x = rng.uniform(0, 1, size=(300, d)).astype(np.float32)
    a = rng.integers(0, 300, 3000)
    b = rng.integers(0, 300, 3000)
    d32, d64 = oracle.pair_distances(x, a, b, with_f64=True)
    nz = d64 > 0
    rel = np.abs(d32[nz] - d64[nz]) / d64[nz]
    assert rel.max() < 1e-6


def test_symmetric_and_zero_on_self(rng):
    x = rng.normal(size=(50, 96)).astype(np.float32)
    a = np.arange(50)
    b = rng.permutation(50)
    assert np.array_equal(oracle.pair_distances(x, a, b), oracle.pair_distances(x, b, a))
    assert np.all(oracle.pair_distances(x, a, a) == 0)
