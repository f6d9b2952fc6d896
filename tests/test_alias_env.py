"""Alias tables (src/math/AliasTable.h:11-57) and the environment two-level table
(src/core/EnvironmentMap.cpp:8-114): library == oracle bit for bit, and the tables
reproduce the distribution they were built from."""
import numpy as np
import pytest

from zillumgl_amd import _lib, scene
from oracle import oracle as orc


def alias_distribution(alias, prob):
    n = len(prob)
    p = np.clip(prob.astype(np.float64), 0, 1)
    P = p.copy()
    np.add.at(P, alias, 1 - p)
    return P / n


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (1000, 3), (4096, 4)])
def test_alias_matches_oracle_and_reconstructs(n, seed):
    rng = np.random.default_rng(seed)
    pdf = (rng.lognormal(size=n) * (rng.uniform(size=n) > 0.2)).astype(np.float32)
    pdf[0] = max(pdf[0], 0.1)
    a1, p1 = _lib.alias_build(pdf)
    a2, p2 = orc.alias_build(pdf)
    assert np.array_equal(a1, a2) and np.array_equal(p1.view(np.uint32), p2.view(np.uint32))
    P = alias_distribution(a1, p1)
    assert np.allclose(P, pdf / pdf.sum(), atol=2e-6 * max(1, n / 1000), rtol=1e-4)


def test_alias_sampling_chi2():
    rng = np.random.default_rng(5)
    pdf = rng.uniform(0.1, 2.0, 16).astype(np.float32)
    alias, prob = _lib.alias_build(pdf)
    u = rng.uniform(size=(2, 400_000)).astype(np.float32)
    cx = (16 * u[0]).astype(int)
    s = np.where(u[1] < prob[cx], cx, alias[cx])
    hist = np.bincount(s, minlength=16)
    exp = pdf / pdf.sum() * len(s)
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 45  # 15 dof, p ~ 1e-4


def test_env_alias_matches_oracle_and_reconstructs():
    env = scene.sky_env(64, 32, seed=3, peak=50.0, sun_radius_deg=5.0)
    a1, p1, s1 = _lib.env_alias_build(env)
    a2, p2, s2 = orc.env_alias_build(env)
    assert np.array_equal(a1, a2) and np.array_equal(p1.view(np.uint32), p2.view(np.uint32)) and s1 == s2
    h, w = env.shape[:2]
    lum = 0.2126 * env[..., 0] + 0.7152 * env[..., 1] + 0.0722 * env[..., 2]
    wgt = lum * np.sin((np.arange(h) + 0.5) / h * np.pi)[:, None]
    assert abs(s1 - wgt.sum()) <= 1e-4 * wgt.sum()
    rows = alias_distribution(a1[:, w], p1[:, w])
    joint = np.stack([alias_distribution(a1[r, :w], p1[r, :w]) for r in range(h)]) * rows[:, None]
    assert np.allclose(joint, wgt / wgt.sum(), rtol=1e-3, atol=1e-7)


def test_env_pdf_is_uniform_for_constant_map():
    """Constant env: the sin(theta)-weighted table samples solid angle uniformly, so the
    pdf used by envPdfLi (portion * w * h / (2 pi^2)) equals 1 / (4 pi) up to the int
    truncation of the sum."""
    w, h = 128, 64
    env = np.ones((h, w, 3), np.float32)
    _, _, s = _lib.env_alias_build(env)
    env_sum = float(int(s))
    pdf = 1.0 / env_sum * w * h * 0.5 / np.pi ** 2
    assert abs(pdf * 4 * np.pi - 1) < 2e-3


def test_zero_env_gives_zero_sum():
    """The reference requires an env map; a black 1x1 map yields uEnvSum = 0."""
    _, _, s = _lib.env_alias_build(np.zeros((1, 1, 3), np.float32))
    assert int(s) == 0


def test_float_to_half_round_to_nearest_even():
    x = np.random.default_rng(6).normal(scale=1000, size=100_000).astype(np.float32)
    x = np.concatenate([x, np.float32([0, -0.0, 65504, 65520, 65519.99, 1e-8, 6e-8, 3e-8, np.inf, -np.inf])])
    h = _lib.float_to_half(x)
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(h, ref)
