"""The naive tracer on the wavefront schedule (ZGL_INTEGRATOR_NAIVE_WAVEFRONT: camera /
trace / shade kernels over compacted queues, many passes per batch, samples added to the
frame in pass order) renders the persistent kernel's frame bit for bit: every option,
both samplers, and the 1080p headline frame, whose 64 passes take two batches."""
import hashlib

import numpy as np
import pytest

from zillumgl_amd import scene, render, _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    print("[gpu] wavefront: creating context on device 0", flush=True)
    c = render.Context(0)
    yield c
    c.close()


def _frames(ctx, ps, first, n, fc, depth, **kw):
    out = []
    for integ in (_lib.INTEGRATOR_NAIVE, _lib.INTEGRATOR_NAIVE_WAVEFRONT):
        ctx.upload(ps)
        ctx.clear()
        ctx.render(integ, first, n, fc, depth, **kw)
        out.append(ctx.read_frame())
    return out


@pytest.mark.parametrize("kw", [dict(max_depth=5), dict(max_depth=5, russian_roulette=True), dict(max_depth=0),
                                dict(max_depth=3, sample_light=False),
                                dict(max_depth=4, light_env_uniform=True, light_portion=0.3)])
def test_wavefront_equals_persistent_c1(ctx, kw):
    print(f"[gpu] wavefront: C1 64x64 {kw}", flush=True)
    depth = kw.pop("max_depth")
    a, b = _frames(ctx, scene.pack(scene.config1(64, 64)), 0, 4, 1, depth, **kw)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_wavefront_independent_sampler_offset_passes(ctx):
    print("[gpu] wavefront: independent sampler, passes 5..10", flush=True)
    ps = scene.pack(scene.config1(48, 40))
    ps.sampler = 0
    a, b = _frames(ctx, ps, 5, 6, 9, 4, russian_roulette=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_wavefront_disney_env(ctx):
    print("[gpu] wavefront: Disney + env 96x54 depth 8", flush=True)
    ps = scene.pack(scene.config2(width=96, height=54, n_spheres=6, subdiv=3, env_size=(256, 128), seed=11))
    a, b = _frames(ctx, ps, 0, 4, 1, 8)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_wavefront_headline_frame(ctx):
    print("[gpu] wavefront: 1080p headline frame, 64 passes", flush=True)
    a, b = _frames(ctx, scene.pack(scene.config2()), 0, 64, 1, 8)
    assert hashlib.sha1(a.tobytes()).hexdigest() == hashlib.sha1(b.tobytes()).hexdigest()
