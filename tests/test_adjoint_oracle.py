"""The oracle's differentiable naive estimator (replay with detached sampling) and its
FP64 adjoint, pinned on the CPU:
  * at the scene's own parameters the replay estimator is the forward renderer's,
    bit for bit (oracle render_pixels, path_integ_naive.glsl:35-174);
  * the FP64 adjoint equals FP64 central differences of the FP64 value of the same
    recorded paths, for material parameters and environment texels;
  * the FP32 replay frame at perturbed parameters agrees with that FP64 value.
"""
import numpy as np
import pytest

from zillumgl_amd import scene
from oracle import oracle as orc

REGION = dict(x0=30, y0=14, w=24, h=16)


@pytest.fixture(scope="module")
def c2s():
    return scene.pack(scene.config2(width=96, height=54, n_spheres=6, subdiv=3, env_size=(256, 128), seed=11))


def eval_point(ps):
    return ps.materials.copy(), ps.env_half.view(np.float16).astype(np.float32)


def _prm():
    return orc.params(0, 0, 3, 1, 4)


def test_replay_at_scene_parameters_is_the_forward_renderer(c2s):
    em, ee = eval_point(c2s)
    rep = orc.render_replay(c2s, _prm(), em, ee, **REGION)
    fwd, _ = orc.render_pixels(c2s, _prm(), REGION["x0"], REGION["y0"], REGION["w"], REGION["h"])
    assert np.array_equal(rep.view(np.uint32), fwd.view(np.uint32))
    assert (rep > 0).mean() > 0.5


def _perturbed(ps, seed):
    rng = np.random.default_rng(seed)
    em, ee = eval_point(ps)
    params = np.stack([em[:, 0], em[:, 1], em[:, 2], em[:, 3], em[:, 5]], 1).astype(np.float64)
    return rng, em, ee, params


def _set(em, params):
    em = em.copy()
    em[:, 0], em[:, 1], em[:, 2], em[:, 3], em[:, 5] = (params[:, k].astype(np.float32) for k in range(5))
    return em


def test_oracle_adjoint_matches_fp64_central_differences(c2s):
    rng, em, ee, params = _perturbed(c2s, 3)
    d_frame = rng.uniform(-1, 1, (REGION["h"], REGION["w"], 3)).astype(np.float32)
    gm, ge = orc.adjoint(c2s, _prm(), em, ee, d_frame, **REGION)
    assert np.abs(gm).max() > 0 and np.abs(ge).max() > 0
    # material parameters: every parameter of the materials seen in the region
    seen = np.nonzero(np.abs(gm).max(1) > 0)[0]
    assert len(seen) >= 2
    for m in seen[:4]:
        for k in range(5):
            h = 1e-4
            pp, pm = params.copy(), params.copy()
            pp[m, k] += h
            pm[m, k] -= h
            # the FP64 value reads the records as FP32: step by the exact FP32 difference
            step = float(np.float32(pp[m, k])) - float(np.float32(pm[m, k]))
            fd = (orc.replay_loss64(c2s, _prm(), _set(em, pp), ee, d_frame, **REGION) -
                  orc.replay_loss64(c2s, _prm(), _set(em, pm), ee, d_frame, **REGION)) / step
            assert abs(gm[m, k] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gm).max()), (m, k, gm[m, k], fd)
    # environment texels with the largest gradients
    flat = np.argsort(-np.abs(ge).ravel())[:6]
    for f in flat:
        h = 1e-2
        ep, emn = ee.copy().ravel(), ee.copy().ravel()
        ep[f] += h
        emn[f] -= h
        step = float(ep[f]) - float(emn[f])
        fd = (orc.replay_loss64(c2s, _prm(), em, ep.reshape(ee.shape), d_frame, **REGION) -
              orc.replay_loss64(c2s, _prm(), em, emn.reshape(ee.shape), d_frame, **REGION)) / step
        assert abs(ge.ravel()[f] - fd) <= 1e-6 * max(abs(fd), 1e-9), (f, ge.ravel()[f], fd)


def test_fp32_replay_frame_agrees_with_fp64_value(c2s):
    rng, em, ee, params = _perturbed(c2s, 4)
    params = np.clip(params + rng.normal(0, 0.1, params.shape), 0.0, 1.0)
    em2 = _set(em, params)
    ee2 = (ee * rng.lognormal(0, 0.1, ee.shape)).astype(np.float32)
    d_frame = rng.uniform(0, 1, (REGION["h"], REGION["w"], 3)).astype(np.float32)
    frame = orc.render_replay(c2s, _prm(), em2, ee2, **REGION)
    v32 = float((frame.astype(np.float64) * d_frame).sum())
    v64 = orc.replay_loss64(c2s, _prm(), em2, ee2, d_frame, **REGION)
    assert abs(v32 - v64) <= 1e-5 * abs(v64)
