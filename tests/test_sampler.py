"""Sampler pins: Sobol generator matrices, Sobol net properties, the integer hash and
the FP32 transcendentals shared by the oracle and the kernels."""
import hashlib
import json
import os

import numpy as np
import pytest

from zillumgl_amd import _lib
from oracle import oracle as orc

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sobol_fixture.json")))


def test_sobol_matrices_match_reference_table():
    """sha256 of the 256x32 table and raw columns from src/math/SobolMatrices256x32.h."""
    for m in (_lib.sobol_matrices(), orc.sobol_matrices()):
        assert hashlib.sha256(m.astype("<u4").tobytes()).hexdigest() == FIX["sha256"]
        for d, cols in FIX["dims"].items():
            assert m[int(d)].tolist() == cols


def test_sobol_dim0_is_bit_reversal():
    for i in [0, 1, 2, 3, 5, 131071, 65536, 12345]:
        rev = int(f"{i:032b}"[::-1], 2)
        assert orc.lib().orc_sobol_sample(i, 0) == rev
        assert _lib.lib().zgl_sobol_sample(i, 0, 0) == rev


def test_sobol_scramble_is_xor():
    L = _lib.lib()
    for i, d, s in [(7, 3, 0xdeadbeef), (100, 200, 12345)]:
        assert L.zgl_sobol_sample(i, d, s) == (L.zgl_sobol_sample(i, d, 0) ^ s)


def _points(dim, n):
    return np.array([orc.lib().orc_sobol_sample(i, dim) for i in range(n)], np.uint64)


@pytest.mark.parametrize("dim", [0, 1, 2, 7, 63, 255])
def test_sobol_each_dim_stratifies_dyadic_intervals(dim):
    """The first 2^k points of every dimension put one point in each [j/2^k, (j+1)/2^k)."""
    p = _points(dim, 1024)
    for k in range(1, 11):
        cells = (p[: 1 << k] >> np.uint64(32 - k)).astype(np.int64)
        assert np.array_equal(np.sort(cells), np.arange(1 << k))


def test_sobol_dims01_form_0m2_net():
    """Dims (0, 1): every elementary interval of volume 2^-m holds exactly one of the first 2^m points."""
    m = 8
    x, y = _points(0, 1 << m), _points(1, 1 << m)
    for a in range(m + 1):
        b = m - a
        cx = (x >> np.uint64(32 - a)).astype(np.int64) if a else np.zeros(len(x), np.int64)
        cy = (y >> np.uint64(32 - b)).astype(np.int64) if b else np.zeros(len(y), np.int64)
        cell = cx * (1 << b) + cy
        assert len(np.unique(cell)) == (1 << m)


def test_hash_is_a_bijection_with_avalanche():
    """random.glsl:5-13 is built from invertible steps (xor-shift, odd multiply), so it
    permutes uint32; flipping one input bit flips about half the output bits."""
    xs = np.arange(0, 1 << 18, dtype=np.uint64) * np.uint64(16411)
    hs = np.array([orc.hash_u32(int(x)) for x in xs])
    assert len(np.unique(hs)) == len(hs)
    flips = []
    rng = np.random.default_rng(0)
    for x in rng.integers(0, 2 ** 32, 2000, dtype=np.uint64):
        b = int(rng.integers(0, 32))
        d = orc.hash_u32(int(x)) ^ orc.hash_u32(int(x) ^ (1 << b))
        flips.append(bin(d).count("1"))
    assert 14 < np.mean(flips) < 18


def test_hash_known_steps():
    """Pin the constants: 0 -> ((61*9) ^ ((61*9) >> 4)) * 0x27d4eb2d, then ^ >> 15."""
    s = (61 * 9) & 0xFFFFFFFF
    s ^= s >> 4
    s = (s * 0x27D4EB2D) & 0xFFFFFFFF
    s ^= s >> 15
    assert orc.hash_u32(0) == s


def _ulp_err(got, ref):
    got = got.astype(np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got - ref) / np.maximum(ulp, np.finfo(np.float32).tiny)


def test_sin_cos_within_2ulp():
    x = np.random.default_rng(1).uniform(-20, 20, 200_000).astype(np.float32)
    s, c = orc.fmath("sin", x), orc.fmath("cos", x)
    xd = x.astype(np.float64)
    # away from zeros of the function the error is relative; near zeros it is absolute
    es = np.abs(s - np.sin(xd)) / np.maximum(np.spacing(np.float32(1.0)) * 0.25, np.spacing(np.abs(np.sin(xd)).astype(np.float32)))
    ec = np.abs(c - np.cos(xd)) / np.maximum(np.spacing(np.float32(1.0)) * 0.25, np.spacing(np.abs(np.cos(xd)).astype(np.float32)))
    assert es.max() <= 2.0 and ec.max() <= 2.0


def test_atan2_within_3ulp():
    rng = np.random.default_rng(2)
    y = rng.normal(size=100_000).astype(np.float32)
    x = rng.normal(size=100_000).astype(np.float32)
    got = orc.fmath("atan2", y, x)
    assert _ulp_err(got, np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() <= 3.0
    assert orc.fmath("atan2", np.float32([0.0]), np.float32([0.0]))[0] == 0.0
    assert orc.fmath("atan2", np.float32([1.0]), np.float32([0.0]))[0] == np.float32(np.pi / 2)


def test_log_exp_pow():
    rng = np.random.default_rng(3)
    x = rng.uniform(1e-6, 1e6, 100_000).astype(np.float32)
    assert _ulp_err(orc.fmath("log", x), np.log(x.astype(np.float64))).max() <= 3.0
    e = rng.uniform(-80, 80, 100_000).astype(np.float32)
    assert _ulp_err(orc.fmath("exp", e), np.exp(e.astype(np.float64))).max() <= 3.0
    a = rng.uniform(1e-4, 1.0, 10_000).astype(np.float32)
    b = rng.uniform(0.0, 1.0, 10_000).astype(np.float32)
    rel = np.abs(orc.fmath("pow", a, b) / np.power(a.astype(np.float64), b.astype(np.float64)) - 1)
    assert rel.max() < 1e-5


def test_sample_offset_wraps_as_glsl_int_arithmetic():
    """sampleOffset = uSpp * uSampleDim, less uSampleNum * uSampleDim past it
    (path_integ_naive.glsl:163-164), in GLSL int arithmetic, which wraps modulo 2^32: pass
    2^24 has offset 2^32 mod 2^32 = 0, the offset of pass 0, and pass 131073 has
    131073 * 256 - 131072 * 256 = 256, the offset of pass 1.  With the same free counter
    the seeds are the same too, so those passes render the same frames bit for bit; an
    int product that overflowed, or no subtraction, would not."""
    from zillumgl_amd import scene
    ps = scene.pack(scene.config1(16, 12))
    base, _ = orc.render_pixels(ps, orc.params(0, 0, 2, 5, 4))
    top, _ = orc.render_pixels(ps, orc.params(0, 1 << 24, 2, 5, 4))
    past, _ = orc.render_pixels(ps, orc.params(0, 1, 1, 5, 4))
    wrapped, _ = orc.render_pixels(ps, orc.params(0, 131073, 1, 5, 4))
    assert np.array_equal(base.view(np.uint32), top.view(np.uint32))
    assert np.array_equal(past.view(np.uint32), wrapped.view(np.uint32))
    other, _ = orc.render_pixels(ps, orc.params(0, 131072, 1, 5, 4))  # offset 2^25: past the table
    assert not np.array_equal(other.view(np.uint32), past.view(np.uint32))
