"""GPU parity on the material-branch scene: the metal-workflow BSDF, the rough dielectric
(GGX normal sampling without visible normals), the thin dielectric, the principled BRDF's
clearcoat lobe (GTR1 sampling) with sheen and subsurface, and the thin-lens camera
(lensRadius > 0) -- every integrator against the CPU oracle.

Tolerances as in test_gpu_parity.py: camera-path integrators bit-exact (naive, global
queue), light-path films within 1e-4 relative above 1e-3 of the image mean (FP32 atomic
splat order)."""
import numpy as np
import pytest

from zillumgl_amd import scene, _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _rel_err(a, b, floor):
    return np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b.astype(np.float64)), floor)


@pytest.fixture(scope="module")
def ctx():
    from zillumgl_amd import render
    print("[gpu] materials: creating context on device 0", flush=True)
    c = render.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module", params=[0.0, 0.05], ids=["pinhole", "thinlens"])
def mats(request):
    return scene.pack(scene.config_materials(96, 96, lens_radius=request.param))


@pytest.mark.parametrize("kw", [dict(), dict(russian_roulette=True), dict(regen_threshold=1),
                                dict(regen_threshold=40, russian_roulette=True)])
def test_naive_bit_exact(ctx, mats, kw):
    print(f"[gpu] materials naive lens {mats.camera.lens_radius} {kw}", flush=True)
    ctx.upload(mats)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_NAIVE, 2, 4, 3, 6, **kw)
    gpu = ctx.read_frame()[..., :3]
    okw = {k: v for k, v in kw.items() if k != "regen_threshold"}
    ref, _ = orc.render_pixels(mats, orc.params(0, 2, 4, 3, 6, **okw))
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), \
        f"{np.count_nonzero(gpu.view(np.uint32) != ref.view(np.uint32))} texels differ"


def test_global_queue_bit_exact(ctx, mats):
    print(f"[gpu] materials global queue lens {mats.camera.lens_radius}", flush=True)
    ctx.upload(mats)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_GLOBAL_QUEUE, 0, 3, 1, 6, russian_roulette=True)
    gpu = ctx.read_frame()[..., :3]
    ref, _ = orc.render_pixels(mats, orc.params(3, 0, 3, 1, 6, russian_roulette=True))
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


def test_wavefront_bit_exact(ctx, mats):
    print(f"[gpu] materials wavefront lens {mats.camera.lens_radius}", flush=True)
    ctx.upload(mats)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_NAIVE_WAVEFRONT, 0, 3, 1, 6, russian_roulette=True)
    gpu = ctx.read_frame()[..., :3]
    ref, _ = orc.render_pixels(mats, orc.params(0, 0, 3, 1, 6, russian_roulette=True))
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("rr", [False, True])
def test_light_tracer(ctx, mats, rr):
    print(f"[gpu] materials light tracer lens {mats.camera.lens_radius} rr {rr}", flush=True)
    ctx.upload(mats)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_LIGHT, 0, 1, 1, 6, russian_roulette=rr, blocks_one_pass=16)
    gpu = ctx.read_film()
    ref, _ = orc.render_light(mats, orc.params(1, 0, 1, 1, 6, russian_roulette=rr, blocks_one_pass=16))
    floor = 1e-3 * max(ref.mean(), 1e-12)
    sig = ref > floor
    err = _rel_err(gpu, ref, floor)
    print(f"[gpu] light: max rel err {err[sig].max():.3g} over {sig.sum()} texels", flush=True)
    assert sig.sum() > 1000 and err[sig].max() <= 1e-4


@pytest.mark.parametrize("rr", [False, True])
def test_triple(ctx, mats, rr):
    print(f"[gpu] materials triple lens {mats.camera.lens_radius} rr {rr}", flush=True)
    ctx.upload(mats)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_TRIPLE, 1, 2, 2, 6, russian_roulette=rr, blocks_one_pass=4)
    gpu = ctx.read_film()
    prm = orc.params(2, 1, 2, 2, 6, russian_roulette=rr, blocks_one_pass=4)
    pt, _ = orc.render_pixels(mats, prm)
    lpt, _ = orc.render_light(mats, prm)
    ref = pt.astype(np.float64) + lpt
    floor = 1e-3 * max(ref.mean(), 1e-12)
    sig = ref > floor
    err = _rel_err(gpu, ref, floor)
    print(f"[gpu] triple: max rel err {err[sig].max():.3g}", flush=True)
    assert err[sig].max() <= 1e-4


@pytest.fixture(scope="module")
def textured():
    return scene.pack(scene.config_textured(96, 96))


@pytest.mark.parametrize("integrator", [_lib.INTEGRATOR_NAIVE, _lib.INTEGRATOR_GLOBAL_QUEUE,
                                        _lib.INTEGRATOR_NAIVE_WAVEFRONT])
def test_textured_camera_paths_bit_exact(ctx, textured, integrator):
    """Albedo textures (texIndex != -1, material_loader.glsl:8-97): repeat past uv 1, a
    layer smaller than the array, principled and metal-workflow materials."""
    print(f"[gpu] textured scene, integrator {integrator}", flush=True)
    ctx.upload(textured)
    ctx.clear()
    ctx.render(integrator, 1, 3, 2, 5, russian_roulette=True)
    gpu = ctx.read_frame()[..., :3]
    orc_int = {_lib.INTEGRATOR_GLOBAL_QUEUE: 3}.get(integrator, 0)
    ref, _ = orc.render_pixels(textured, orc.params(orc_int, 1, 3, 2, 5, russian_roulette=True))
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))


def test_textured_light_paths(ctx, textured):
    print("[gpu] textured scene, light and triple", flush=True)
    ctx.upload(textured)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_LIGHT, 0, 1, 1, 5, blocks_one_pass=16)
    gpu = ctx.read_film()
    ref, _ = orc.render_light(textured, orc.params(1, 0, 1, 1, 5, blocks_one_pass=16))
    floor = 1e-3 * max(ref.mean(), 1e-12)
    sig = ref > floor
    assert _rel_err(gpu, ref, floor)[sig].max() <= 1e-4
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_TRIPLE, 0, 2, 1, 5, blocks_one_pass=4)
    gpu = ctx.read_film()
    prm = orc.params(2, 0, 2, 1, 5, blocks_one_pass=4)
    ref = orc.render_pixels(textured, prm)[0].astype(np.float64) + orc.render_light(textured, prm)[0]
    floor = 1e-3 * max(ref.mean(), 1e-12)
    sig = ref > floor
    assert _rel_err(gpu, ref, floor)[sig].max() <= 1e-4


def test_texture_state_errors(ctx, textured):
    """A scene naming texture layers that are not set cannot render; the differentiable
    mode refuses textured scenes."""
    ctx.upload(textured)
    ctx.set_textures(None, None)
    with pytest.raises(_lib.ZglError) as e:
        ctx.render(_lib.INTEGRATOR_NAIVE, 0, 1, 1, 3)
    assert e.value.status == _lib.ZGL_E_STATE
    ctx.upload(textured)
    with pytest.raises(_lib.ZglError) as e:
        ctx.render_replay(0, 1, 1, 3)
    assert e.value.status == _lib.ZGL_E_INVALID


@pytest.mark.parametrize("threshold", [1, 20, 63])
def test_light_persistent_equals_one_thread_per_path(ctx, mats, threshold):
    """The light tracer's persistent schedule (regenerating lanes, walks advanced a few
    rows per round, connections splatted after their visibility walk) and its
    one-thread-per-path form (the byte-counting variant) splat the same contributions:
    films equal within 1e-4 (FP32 atomic order), with roulette, at several regeneration
    thresholds."""
    print(f"[gpu] materials light persistent vs one thread per path, threshold {threshold}", flush=True)
    ctx.upload(mats)
    films = []
    for count in (False, True):
        ctx.clear()
        ctx.render(_lib.INTEGRATOR_LIGHT, 2, 3, 4, 6, russian_roulette=True, blocks_one_pass=16,
                   count_bytes=count, regen_threshold=threshold)
        films.append(ctx.read_film().astype(np.float64))
    floor = 1e-3 * max(films[1].mean(), 1e-12)
    sig = films[1] > floor
    assert sig.sum() > 1000 and _rel_err(films[0], films[1], floor)[sig].max() <= 1e-4
