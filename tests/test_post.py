"""Display transform (post_proc.glsl:1-59) and image files.

Oracle: the tone maps against their closed forms in FP64 (filmic: John Hable's curve
with the reference's constants and white point 11.2; ACES: Narkowicz's fit), the gamma
against numpy's power, clamping and the byte conversion (GL: clamp, round(255 x), NaN
-> 0, rows written top-down as captureImage flips them).  Image files: PNG / PFM / RGBE
round trips."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from zillumgl_amd import image


def _frame(vals):
    v = np.asarray(vals, np.float32)
    f = np.zeros((1, len(v), 4), np.float32)
    f[0, :, 0] = f[0, :, 1] = f[0, :, 2] = v
    f[..., 3] = 1
    return f


def _filmic64(x):
    A, B, C, D, E, F = 0.22, 0.3, 0.1, 0.2, 0.01, 0.3
    calc = lambda t: (t * (t * A + B * C) + D * E) / (t * (t * A + B) + D * F) - E / F
    return calc(1.6 * x) / calc(11.2)


def _aces64(x):
    return (x * (x * 2.51 + 0.03)) / (x * (x * 2.43 + 0.59) + 0.14)


@pytest.mark.parametrize("tone,fn", [(0, lambda x: x), (1, _filmic64), (2, _aces64)])
def test_tone_maps_match_closed_forms(tone, fn):
    x = np.concatenate([np.linspace(0, 4, 2001), np.geomspace(1e-4, 1e4, 500)]).astype(np.float32)
    out, _ = orc.post_process(_frame(x), 1.0, tone)
    got = out[0, :, 0].astype(np.float64)
    # in the linear domain: the FP32 filmic curve cancels to ~1e-9 at 0, which the
    # 1/2.2 gamma would magnify
    lin = fn(x.astype(np.float64))
    assert np.all(np.abs(got ** 2.2 - lin) <= 1e-5 * np.abs(lin) + 1e-8), np.abs(got ** 2.2 - lin).max()
    # the gamma itself against numpy where the value is well away from 0
    m = lin > 1e-3
    ref = np.power(lin[m], 1 / 2.2)
    assert np.all(np.abs(got[m] - ref) <= 3e-6 * ref)
    assert np.all(out[0, :, 3] == 1)


def test_clamp_scale_and_bytes():
    x = [-5.0, 0.0, 0.5, 1.0, 2.0, 1e30, 3e38, np.nan]
    out, rgb = orc.post_process(_frame(x), 0.5, 0)
    m = out[0, :, 0]
    assert m[0] == 0 and m[1] == 0                      # clamped below
    assert abs(m[2] - 0.25 ** (1 / 2.2)) < 1e-6          # scaled then gamma
    assert np.isfinite(m[5]) and m[5] > 1e12             # 0.5e30 ** (1/2.2)
    assert rgb[0, 0, 0] == 0 and rgb[0, 1, 0] == 0 and rgb[0, 4, 0] == 255 and rgb[0, 5, 0] == 255
    assert rgb[0, 7, 0] == 0                             # NaN -> 0
    # round(255 x) with x = 0.25^(1/2.2)
    assert rgb[0, 2, 0] == int(np.floor(255 * 0.25 ** (1 / 2.2) + 0.5))


def test_rows_are_written_top_down():
    f = np.zeros((3, 2, 4), np.float32)
    f[0, :, :3] = 1.0   # frame row 0 is the bottom of the picture
    _, rgb = orc.post_process(f, 1.0, 0)
    assert (rgb[2] == 255).all() and (rgb[0] == 0).all()


def test_png_pfm_hdr_round_trips(tmp_path):
    rng = np.random.default_rng(1)
    rgb8 = rng.integers(0, 256, (17, 23, 3), dtype=np.uint8)
    image.write_png(str(tmp_path / "a.png"), rgb8)
    assert np.array_equal(image.read_png_rgb8(str(tmp_path / "a.png")), rgb8)
    f = rng.uniform(0, 100, (9, 11, 3)).astype(np.float32)
    image.write_pfm(str(tmp_path / "a.pfm"), f)
    assert np.array_equal(image.read_pfm(str(tmp_path / "a.pfm")), f)
    image.write_hdr(str(tmp_path / "a.hdr"), f)
    g = image.read_hdr(str(tmp_path / "a.hdr"))
    # RGBE keeps 8 mantissa bits of the largest channel
    assert np.all(np.abs(g - f) <= f.max(axis=2, keepdims=True) / 128.0)
