"""Integrator pins on the CPU oracle: closed-form furnace scenes and estimator identities."""
import numpy as np

from zillumgl_amd import scene
from oracle import oracle as orc


def furnace_scene(albedo=0.5, env=1.0, w=24, h=24):
    """An infinite-looking Lambertian floor under a constant environment: every
    camera ray hits the floor, whose outgoing radiance is albedo * env exactly."""
    d = scene.SceneDesc()
    fl = scene.quad((-1000, -1000, 0), (1000, -1000, 0), (1000, 1000, 0), (-1000, 1000, 0), (0, 0, 1))
    d.objects = [(fl, scene.Material(base_color=(albedo,) * 3))]
    d.env = np.full((32, 64, 3), env, np.float32)
    d.camera = scene.CameraDesc(pos=(0, 0, 1), angle=(0, -60, 0), fov=20)
    d.width, d.height = w, h
    return scene.pack(d)


def test_env_furnace_closed_form():
    ps = furnace_scene(0.5, 1.0)
    spp = 64
    img, _ = orc.render_pixels(ps, orc.params(0, 0, spp, 1, 2))
    mean = img.mean() / spp
    assert abs(mean - 0.5) < 0.01, mean


def test_env_furnace_nee_off_same_mean():
    ps = furnace_scene(0.7, 2.0)
    spp = 64
    on, _ = orc.render_pixels(ps, orc.params(0, 0, spp, 1, 2, sample_light=True))
    off, _ = orc.render_pixels(ps, orc.params(0, 0, spp, 1, 2, sample_light=False))
    assert abs(on.mean() / spp - 1.4) < 0.02
    assert abs(off.mean() / spp - 1.4) < 0.05


def test_depth_zero_sees_only_emitters():
    ps = furnace_scene(0.5, 1.0)
    img, _ = orc.render_pixels(ps, orc.params(0, 0, 4, 1, 0))
    assert np.all(img == 0)


def test_fp64_oracle_agrees_with_fp32_in_mean():
    ps = scene.pack(scene.config1(32, 32))
    a, _ = orc.render_pixels(ps, orc.params(0, 0, 32, 1, 5))
    b, _ = orc.render_pixels(ps, orc.params(0, 0, 32, 1, 5), fp64=True)
    # different sample streams (the FP64 seeds round differently) but the same estimator
    assert abs(a.mean() - b.mean()) < 0.03 * b.mean()


def primary_light_mask(ps):
    """Pixels whose pinhole centre ray (or a neighbour's) sees an emitter directly."""
    W, H = ps.width, ps.height
    c = ps.camera
    F, R, U, P = (np.array(v[:], np.float64) for v in (c.front, c.right, c.up, c.pos))
    ys, xs = np.mgrid[0:H, 0:W]
    nx = (xs + 0.5) / W * 2 - 1
    ny = (ys + 0.5) / H * 2 - 1
    d = (R[None, None] * (nx * c.aspect * c.tan_fov)[..., None] + U[None, None] * (ny * c.tan_fov)[..., None]
         + F[None, None])
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = np.broadcast_to(P, d.shape)
    rays = np.ascontiguousarray(np.concatenate([o.reshape(-1, 3).T, d.reshape(-1, 3).T]).astype(np.float32))
    prim, _, _, _ = orc.trace_closest(ps, rays)
    m = (prim >= ps.obj_prim_count).reshape(H, W)
    mm = m.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            mm |= np.roll(np.roll(m, dy, 0), dx, 1)
    return mm


def test_light_tracer_matches_path_tracer_off_emitters():
    """Light tracing and path tracing estimate the same image, except on pixels that see
    an emitter directly: the reference's light->camera connection omits the cosine at
    the light (light_path_integ.glsl:240), which the restatement keeps."""
    ps = scene.pack(scene.config1(32, 32))
    pt, _ = orc.render_pixels(ps, orc.params(0, 0, 256, 1, 5))
    pt = pt / 256
    passes, blocks = 8, 4
    film, _ = orc.render_light(ps, orc.params(1, 0, passes, 1, 5, blocks_one_pass=blocks))
    lt = film / (passes * blocks * 1536 / (32 * 32))
    mask = ~primary_light_mask(ps)
    a, b = lt[mask].mean(), pt[mask].mean()
    assert abs(a - b) < 0.03 * b, (a, b)
    # where the light is seen directly the light tracer is brighter (no cos at the light)
    assert lt[~mask].mean() > pt[~mask].mean()
