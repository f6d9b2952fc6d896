"""GPU parity: the HIP kernels through the C ABI against the CPU oracle on the same inputs.

Tolerances:
  * closest hit (prim, t bits, barycentric bin): bit-exact;
  * naive path tracer frame sums at matched Sobol indices: bit-exact (each sample's
    operations are the oracle's, and each pixel's sum is formed in pass order);
  * light tracer / triple path films: 1e-4 relative per pixel above 1e-3 of the
    image mean, since FP32 atomic splat order differs from the oracle's FP64 sums.
"""
import numpy as np
import pytest

from zillumgl_amd import scene, render, _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _rel_err(a, b, floor):
    return np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b.astype(np.float64)), floor)


@pytest.fixture(scope="module")
def ctx():
    print("[gpu] creating context on device 0", flush=True)
    c = render.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c1():
    return scene.pack(scene.config1())


def test_device_present():
    assert _lib.device_count() >= 1


def test_closest_hit_c1_1m_rays_bit_exact(ctx, c1):
    print("[gpu] closest hit, C1, 1M rays", flush=True)
    ctx.upload(c1)
    rays = scene.ray_batch(1_000_000, seed=0)
    prim, t, b = ctx.trace_closest(rays)
    rp, rt, rb, _ = orc.trace_closest(c1, rays)
    assert np.array_equal(prim, rp), f"{np.count_nonzero(prim != rp)} prim mismatches"
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32)), "t bits differ"
    assert np.array_equal(b, rb), "barycentric bins differ"
    # the box is open on the camera side (y = 0), so roughly 1/7 of the rays escape
    assert 0.8 < (prim >= 0).mean() < 0.95


def _special_rays(n, seed):
    """Rays inside the Cornell box with directions on or near the axes, with components
    at or below 1e-6, signed zeros, and a few NaN/inf components: every case of the box
    test, mixed with general rays in the same waves."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.05, 0.95, (n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(-1.0, 0.95, n).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    kind = rng.integers(0, 6, n)
    axis = rng.integers(0, 3, n)
    sgn = np.where(rng.random(n) < 0.5, -1.0, 1.0).astype(np.float32)
    idx = np.arange(n)
    m = kind == 1                             # exact axes, other components +0 / -0
    d[m] = 0.0
    d[idx[m], axis[m]] = sgn[m]
    d[idx[m], (axis[m] + 1) % 3] = np.where(rng.random(m.sum()) < 0.5, -0.0, 0.0).astype(np.float32)
    m = kind == 2                             # |d_a| around 1 - 1e-6
    d[idx[m], axis[m]] = sgn[m] * rng.choice(np.float32([1 - 5e-7, 1 - 1e-6, 1 - 2e-6]), m.sum())
    m = kind == 3                             # one component at or below 1e-6
    d[idx[m], axis[m]] = rng.choice(np.float32([0.0, -0.0, 4e-7, -8e-7, 1e-6]), m.sum())
    m = (kind == 4) & (rng.random(n) < 0.02)  # a few NaN / inf components
    d[idx[m], axis[m]] = rng.choice(np.float32([np.nan, np.inf]), m.sum())
    return np.ascontiguousarray(np.concatenate([o.T, d.T]).astype(np.float32))


def test_closest_hit_special_rays_bit_exact(ctx, c1):
    print("[gpu] closest hit, C1, special directions", flush=True)
    ctx.upload(c1)
    rays = _special_rays(400_000, seed=3)
    prim, t, b = ctx.trace_closest(rays)
    rp, rt, rb, _ = orc.trace_closest(c1, rays)
    assert np.array_equal(prim, rp), f"{np.count_nonzero(prim != rp)} prim mismatches"
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32)), "t bits differ"
    assert np.array_equal(b, rb)
    assert 0.3 < (prim >= 0).mean() < 0.99


def test_upload_rejects_nonfinite_bounds(ctx, c1):
    import copy
    bad = copy.copy(c1)
    bad.bounds = c1.bounds.copy()
    bad.bounds[3, 4] = np.inf
    with pytest.raises(_lib.ZglError):
        ctx.upload(bad)
    ctx.upload(c1)


def test_closest_hit_counters_match_oracle(ctx, c1):
    ctx.upload(c1)
    rays = scene.ray_batch(100_000, seed=5)
    ctx.reset_counters()
    prim, t, b = ctx.trace_closest(rays, counted=True)
    cnt = ctx.counters()
    _, _, _, rc = orc.trace_closest(c1, rays)
    assert list(cnt[:4]) == list(rc[:4])


def test_any_hit_matches_closest(ctx, c1):
    ctx.upload(c1)
    rays = scene.ray_batch(200_000, seed=9)
    prim, t, _ = ctx.trace_closest(rays)
    hit = prim >= 0
    # with tmax just above the closest t every hit ray is occluded; just below, none of
    # the rays can report a triangle closer than the closest one
    occ_hi = ctx.trace_any(rays, np.where(hit, t * 1.0001 + 1e-6, 1e8).astype(np.float32))
    assert np.array_equal(occ_hi, hit)
    occ_lo = ctx.trace_any(rays, np.where(hit, t * 0.9999 - 1e-6, 1e8).astype(np.float32))
    assert not occ_lo[hit].any()


def test_naive_c1_matches_oracle(ctx, c1):
    print("[gpu] naive C1 256x256 16spp depth 5", flush=True)
    ctx.upload(c1)
    integ = render.NaivePathIntegrator(max_depth=5)
    integ.init(ctx)
    integ.render_one_pass(16)
    gpu = integ.get_frame()[..., :3]
    ref, _ = orc.render_pixels(c1, orc.params(0, 0, 16, 1, 5))
    err = _rel_err(gpu, ref, 1e-3)
    exact = np.mean(gpu.view(np.uint32) == ref.view(np.uint32))
    print(f"[gpu] naive C1: max rel err {err.max():.3g}, bit-exact fraction {exact:.6f}", flush=True)
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


def test_naive_c1_pass_by_pass_equals_batched(ctx, c1):
    """16 launches of one pass give the same sums as one launch of 16 passes."""
    # (the persistent kernel forms each pixel's sum in pass order within a launch)
    ctx.upload(c1)
    a = render.NaivePathIntegrator(max_depth=5)
    a.init(ctx)
    for _ in range(16):
        a.render_one_pass(1)
    fa = a.get_frame()
    b = render.NaivePathIntegrator(max_depth=5)
    b.init(ctx)
    b.free_counter = 0
    b.render_one_pass(16)
    fb = b.get_frame()
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))


def test_naive_c1_counters_match_oracle(ctx, c1):
    ctx.upload(c1)
    ctx.reset_counters()
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_NAIVE, 0, 2, 1, 5, count_bytes=True)
    cnt = ctx.counters()
    _, rc = orc.render_pixels(c1, orc.params(0, 0, 2, 1, 5))
    assert list(cnt[:4]) == list(rc[:4])


@pytest.fixture(scope="module")
def c2_small():
    d = scene.config2(width=96, height=54, n_spheres=6, subdiv=3, env_size=(256, 128), seed=11)
    return scene.pack(d)


def test_naive_disney_env_matches_oracle(ctx, c2_small):
    print("[gpu] naive Disney+env 96x54 4spp depth 8", flush=True)
    ctx.upload(c2_small)
    integ = render.NaivePathIntegrator(max_depth=8)
    integ.init(ctx)
    integ.render_one_pass(4)
    gpu = integ.get_frame()[..., :3]
    ref, _ = orc.render_pixels(c2_small, orc.params(0, 0, 4, 1, 8))
    err = _rel_err(gpu, ref, 1e-3)
    print(f"[gpu] naive C2-small: max rel err {err.max():.3g}", flush=True)
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


@pytest.fixture(scope="module")
def c3_small():
    return scene.pack(scene.config3(width=128, height=128))


def test_light_tracer_matches_oracle(ctx, c3_small):
    print("[gpu] light tracer C3 128x128, 1 pass, depth 8", flush=True)
    ctx.upload(c3_small)
    integ = render.LightPathIntegrator(max_depth=8, thread_blocks_one_pass=32)
    integ.init(ctx)
    integ.render_one_pass(1)
    gpu = ctx.read_film()
    ref, _ = orc.render_light(c3_small, orc.params(1, 0, 1, 1, 8, blocks_one_pass=32))
    floor = 1e-3 * max(ref.mean(), 1e-12)
    sig = ref > floor
    err = _rel_err(gpu, ref, floor)
    print(f"[gpu] light: max rel err {err[sig].max():.3g} over {sig.sum()} texels", flush=True)
    assert err[sig].max() <= 1e-4
    assert abs(gpu.sum() - ref.sum()) <= 1e-5 * ref.sum()


def test_triple_matches_oracle(ctx, c3_small):
    print("[gpu] triple C3 128x128, 1 pass, depth 8", flush=True)
    ctx.upload(c3_small)
    integ = render.TriplePathIntegrator(max_depth=8, lpt_blocks_one_pass=8)
    integ.init(ctx)
    integ.render_one_pass(1)
    gpu = ctx.read_film()
    pt, _ = orc.render_pixels(c3_small, orc.params(2, 0, 1, 1, 8, blocks_one_pass=8))
    lpt, _ = orc.render_light(c3_small, orc.params(2, 0, 1, 1, 8, blocks_one_pass=8))
    ref = pt.astype(np.float64) + lpt
    floor = 1e-3 * max(ref.mean(), 1e-12)
    sig = ref > floor
    err = _rel_err(gpu, ref, floor)
    print(f"[gpu] triple: max rel err {err[sig].max():.3g}", flush=True)
    assert err[sig].max() <= 1e-4


@pytest.mark.parametrize("threshold", [1, 20, 63])
def test_triple_persistent_thresholds_match_oracle(ctx, c3_small, threshold):
    """The triple tracer's camera pass on the persistent schedule, passes 2..4 (free
    counter 3) in one call, at several regeneration thresholds; and the one-thread-per-
    pixel camera pass (the byte-counting variant) gives the same film."""
    print(f"[gpu] triple persistent threshold {threshold}", flush=True)
    ctx.upload(c3_small)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_TRIPLE, 2, 3, 3, 8, blocks_one_pass=8, regen_threshold=threshold)
    gpu = ctx.read_film()
    prm = orc.params(2, 2, 3, 3, 8, blocks_one_pass=8)
    pt, _ = orc.render_pixels(c3_small, prm)
    lpt, _ = orc.render_light(c3_small, prm)
    ref = pt.astype(np.float64) + lpt
    floor = 1e-3 * max(ref.mean(), 1e-12)
    sig = ref > floor
    err = _rel_err(gpu, ref, floor)
    print(f"[gpu] triple persistent: max rel err {err[sig].max():.3g}", flush=True)
    assert err[sig].max() <= 1e-4
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_TRIPLE, 2, 3, 3, 8, blocks_one_pass=8, count_bytes=True)
    plain = ctx.read_film()
    assert _rel_err(plain, gpu.astype(np.float64), floor)[sig].max() <= 1e-4


@pytest.mark.parametrize("threshold", [1, 16, 40, 63])
def test_persistent_thresholds_match_oracle(ctx, c2_small, threshold):
    """Whatever the regeneration threshold, each walk visits the same rows and each
    pixel's sum is formed in pass order: the frame is the oracle's bit for bit (passes
    3..7, free counter 4)."""
    print(f"[gpu] persistent threshold {threshold}", flush=True)
    ctx.upload(c2_small)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_NAIVE, 3, 5, 4, 8, regen_threshold=threshold)
    gpu = ctx.read_frame()[..., :3]
    ref, _ = orc.render_pixels(c2_small, orc.params(0, 3, 5, 4, 8))
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("threshold", [1, 20, 63])
def test_persistent_roulette_is_exact_and_repeatable(ctx, c2_small, threshold):
    """Russian roulette on the Disney + environment scene at depth 8 (the roulette ends
    paths at every depth): the frame is the oracle's bit for bit at every regeneration
    threshold, and a second launch gives the same bits."""
    print(f"[gpu] persistent roulette threshold {threshold}", flush=True)
    ctx.upload(c2_small)
    frames = []
    for _ in range(2):
        ctx.clear()
        ctx.render(_lib.INTEGRATOR_NAIVE, 1, 4, 2, 8, russian_roulette=True, regen_threshold=threshold)
        frames.append(ctx.read_frame()[..., :3].copy())
    ref, _ = orc.render_pixels(c2_small, orc.params(0, 1, 4, 2, 8, russian_roulette=True))
    assert np.array_equal(frames[0].view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(frames[0].view(np.uint32), frames[1].view(np.uint32))


def test_triple_persistent_roulette_matches_plain_camera_pass(ctx, c3_small):
    """With roulette, the persistent camera pass and the one-thread-per-pixel one (the
    byte-counting variant) give the same film up to the light pass's atomic order."""
    ctx.upload(c3_small)
    films = []
    for count in (False, True):
        ctx.clear()
        ctx.render(_lib.INTEGRATOR_TRIPLE, 0, 4, 1, 8, russian_roulette=True, blocks_one_pass=8, count_bytes=count)
        films.append(ctx.read_film().astype(np.float64))
    floor = 1e-3 * max(films[1].mean(), 1e-12)
    sig = films[1] > floor
    assert _rel_err(films[0], films[1], floor)[sig].max() <= 1e-4


@pytest.mark.parametrize("kw", [dict(max_depth=5, russian_roulette=True), dict(max_depth=0),
                                dict(max_depth=3, sample_light=False), dict(max_depth=4, light_env_uniform=True,
                                                                            light_portion=0.3)])
def test_naive_options_match_oracle(ctx, c1, kw):
    ctx.upload(c1)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_NAIVE, 0, 4, 1, **kw)
    gpu = ctx.read_frame()[..., :3]
    ref, _ = orc.render_pixels(c1, orc.params(0, 0, 4, 1, **kw))
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), kw


def test_light_and_triple_with_rr_match_oracle(ctx):
    ps = scene.pack(scene.config3(64, 64))
    ctx.upload(ps)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_LIGHT, 0, 1, 1, 6, russian_roulette=True, blocks_one_pass=4)
    g = ctx.read_film()
    r, _ = orc.render_light(ps, orc.params(1, 0, 1, 1, 6, russian_roulette=True, blocks_one_pass=4))
    fl = 1e-3 * r.mean()
    m = r > fl
    assert (np.abs(g - r) / np.maximum(r, fl))[m].max() <= 1e-4
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_TRIPLE, 0, 1, 1, 6, russian_roulette=True, blocks_one_pass=2)
    g = ctx.read_film()
    pt, _ = orc.render_pixels(ps, orc.params(2, 0, 1, 1, 6, russian_roulette=True, blocks_one_pass=2))
    lp, _ = orc.render_light(ps, orc.params(2, 0, 1, 1, 6, russian_roulette=True, blocks_one_pass=2))
    r = pt + lp
    fl = 1e-3 * r.mean()
    m = r > fl
    assert (np.abs(g - r) / np.maximum(r, fl))[m].max() <= 1e-4


def test_render_rejects_bad_threshold(ctx, c1):
    ctx.upload(c1)
    with pytest.raises(_lib.ZglError):
        ctx.render(_lib.INTEGRATOR_NAIVE, 0, 1, 1, 4, regen_threshold=64)
    with pytest.raises(_lib.ZglError):
        ctx.render(_lib.INTEGRATOR_NAIVE, 0, 1, 1, 4, regen_threshold=-1)


@pytest.mark.parametrize("tone", [0, 1, 2])
def test_post_process_matches_oracle(ctx, c1, tone):
    """post_proc.glsl on the GPU == the oracle bit for bit (floats and RGB8 bytes), on a
    rendered frame with extreme values written into a few pixels."""
    print(f"[gpu] post-process tone {tone}", flush=True)
    ctx.upload(c1)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_NAIVE, 0, 4, 1, 5)
    f = ctx.read_frame().copy()
    f[0, 0, :3] = [-1.0, 0.0, 1e30]
    f[0, 1, :3] = [3e38, 1e-30, 0.5]
    f[1, 0, :3] = [np.inf, 2.0, 1e-7]
    ctx.write_frame(f)
    rgba, rgb = ctx.post_process(np.float32(1.0 / 5.0), tone)
    ref_rgba, ref_rgb = orc.post_process(f, np.float32(1.0 / 5.0), tone)
    assert np.array_equal(rgba.view(np.uint32), ref_rgba.view(np.uint32))
    assert np.array_equal(rgb, ref_rgb)
    assert 0 < rgb.mean() < 255
