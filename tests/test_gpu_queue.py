"""GPU parity of the global-queue (wavefront) path tracer: the HIP schedule (compacted
per-depth queues, device-side counts, wave-aggregated pushes) against the oracle's
restatement of GlobalQueuePathIntegrator (src/integrator/GlobalQueuePath.cpp:100-155)
on the same inputs: frame sums bit for bit, for both samplers and every option."""
import numpy as np
import pytest

from zillumgl_amd import scene, render, _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    print("[gpu] queue: creating context on device 0", flush=True)
    c = render.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c1():
    return scene.pack(scene.config1(64, 64))


def _gpu(ctx, ps, first, n, fc, depth, **kw):
    ctx.upload(ps)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_GLOBAL_QUEUE, first, n, fc, depth, **kw)
    return ctx.read_frame()


@pytest.mark.parametrize("kw", [dict(max_depth=5), dict(max_depth=5, russian_roulette=True), dict(max_depth=0),
                                dict(max_depth=3, sample_light=False),
                                dict(max_depth=4, light_env_uniform=True, light_portion=0.3)])
def test_global_queue_c1_matches_oracle(ctx, c1, kw):
    print(f"[gpu] queue: C1 64x64 4 passes {kw}", flush=True)
    depth = kw.pop("max_depth")
    gpu = _gpu(ctx, c1, 0, 4, 1, depth, **kw)
    ref, _ = orc.render_pixels(c1, orc.params(3, 0, 4, 1, depth, **kw))
    assert np.array_equal(gpu[..., :3].view(np.uint32), ref.view(np.uint32)), kw
    assert (gpu[..., 3][(gpu[..., :3] != 0).any(-1)] == 1.0).all()


def test_global_queue_independent_sampler_and_offset_passes(ctx, c1):
    print("[gpu] queue: independent sampler, passes 7..9", flush=True)
    ps = scene.pack(scene.config1(64, 64))
    ps.sampler = 0
    gpu = _gpu(ctx, ps, 7, 3, 12, 4)
    ref, _ = orc.render_pixels(ps, orc.params(3, 7, 3, 12, 4))
    assert np.array_equal(gpu[..., :3].view(np.uint32), ref.view(np.uint32))


def test_global_queue_disney_env_matches_oracle(ctx):
    print("[gpu] queue: Disney + env 96x54 4 passes depth 8", flush=True)
    ps = scene.pack(scene.config2(width=96, height=54, n_spheres=6, subdiv=3, env_size=(256, 128), seed=11))
    gpu = _gpu(ctx, ps, 0, 4, 1, 8)
    ref, _ = orc.render_pixels(ps, orc.params(3, 0, 4, 1, 8))
    assert np.array_equal(gpu[..., :3].view(np.uint32), ref.view(np.uint32))
    assert (ref > 0).mean() > 0.5


def test_global_queue_integrator_passes_accumulate(ctx, c1):
    """render_one_pass per pass == one call of all passes; result scale 1/(N+1)."""
    ctx.upload(c1)
    a = render.GlobalQueuePathIntegrator(max_depth=4)
    a.init(ctx)
    for _ in range(6):
        a.render_one_pass(1)
    fa = a.get_frame()
    b = render.GlobalQueuePathIntegrator(max_depth=4)
    b.init(ctx)
    b.render_one_pass(6)
    fb = b.get_frame()
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
    assert a.result_scale() == b.result_scale() == 1.0 / 7.0


def test_global_queue_rejects_count_bytes(ctx, c1):
    ctx.upload(c1)
    with pytest.raises(_lib.ZglError):
        ctx.render(_lib.INTEGRATOR_GLOBAL_QUEUE, 0, 1, 1, 4, count_bytes=True)
