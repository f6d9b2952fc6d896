"""The oracle's restatement of the global-queue path tracer (GlobalQueuePathIntegrator,
src/integrator/GlobalQueuePath.cpp:100-155; pt_global_queue_primary.glsl and
pt_global_queue_streaming.glsl), pinned two ways:

  * primary rays: with max_depth 0 only primary contributions (env or light seen
    directly) remain, and the frame equals the naive tracer's bit for bit (same Sobol
    dims 0-3, same camera, same per-pass order);
  * bounces: the estimator differs from the naive tracer's only in which random numbers
    each vertex uses (dims 4 + 9 depth, seeds restarted per depth), so both are unbiased
    estimators of the same image: over many passes their means agree within the
    statistical error (z-test on 8x8-pixel block means over 256 passes); a 3% bias
    planted into the oracle's output fails it, and a dropped term (NEE, an MIS weight,
    the cosine in the throughput) biases the image by far more.
"""
import numpy as np
import pytest

from zillumgl_amd import scene
from oracle import oracle as orc


@pytest.fixture(scope="module")
def c1_small():
    return scene.pack(scene.config1(32, 32))


def _sum(ps, integrator, n, depth, first=0, **kw):
    img, _ = orc.render_pixels(ps, orc.params(integrator, first, n, 1 + first, depth, **kw))
    return img.astype(np.float64)


def test_depth0_equals_naive(c1_small):
    for kw in (dict(), dict(sample_light=False)):
        a = _sum(c1_small, 3, 4, 0, **kw)
        b = _sum(c1_small, 0, 4, 0, **kw)
        assert np.array_equal(a.astype(np.float32).view(np.uint32), b.astype(np.float32).view(np.uint32))
        assert (a > 0).any()


def _block_means(ps, integrator, passes, chunk, depth, **kw):
    """Per-chunk mean images (chunks of `chunk` passes, disjoint pass ranges)."""
    out = []
    for first in range(0, passes, chunk):
        out.append(_sum(ps, integrator, chunk, depth, first=first, **kw) / chunk)
    return np.stack(out)


def _z(ps, depth, **kw):
    """z-score of the difference of the two estimators' image means (8x8-pixel blocks)."""
    gq = _block_means(ps, 3, 256, 32, depth, **kw)
    nv = _block_means(ps, 0, 256, 32, depth, **kw)

    def blocks(a):
        n, h, w, c = a.shape
        return a.reshape(n, h // 8, 8, w // 8, 8, c).mean(axis=(2, 4))

    g, v = blocks(gq), blocks(nv)
    diff = g.mean(0) - v.mean(0)
    se = np.sqrt(g.var(0, ddof=1) / g.shape[0] + v.var(0, ddof=1) / v.shape[0])
    m = se > 0
    return diff[m] / se[m], g.mean(0)[m], v.mean(0)[m]


@pytest.mark.parametrize("kw", [dict(), dict(russian_roulette=True), dict(sample_light=False)])
def test_bounces_unbiased_against_naive(c1_small, kw):
    z, g, v = _z(c1_small, 4, **kw)
    # 48 block-channels; |z| above 4.5 happens by chance with probability ~3e-4
    assert np.abs(z).max() < 4.5, (kw, np.abs(z).max())
    assert abs(g.sum() / v.sum() - 1) < 0.02


def test_planted_bias_is_detected(c1_small):
    """The check has power: the global-queue means scaled by 0.97 (a 3% bias, planted on
    the CPU into the oracle's output) fail the same z-test.  Dropping a term -- NEE, an
    MIS weight, the cosine of the throughput -- biases the image by far more than 3%."""
    gq = _block_means(c1_small, 3, 256, 32, 4) * 0.97
    nv = _block_means(c1_small, 0, 256, 32, 4)

    def blocks(a):
        n, h, w, c = a.shape
        return a.reshape(n, h // 8, 8, w // 8, 8, c).mean(axis=(2, 4))

    g, v = blocks(gq), blocks(nv)
    se = np.sqrt(g.var(0, ddof=1) / g.shape[0] + v.var(0, ddof=1) / v.shape[0])
    z = (g.mean(0) - v.mean(0))[se > 0] / se[se > 0]
    assert np.abs(z).max() >= 4.5
