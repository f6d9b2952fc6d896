"""GPU parity of the differentiable naive estimator and its path-replay adjoint.

  * at the uploaded parameters the replay kernel's frame equals zgl_render's bit for bit;
  * at perturbed parameters it equals the oracle's FP32 replay (same operations), per
    pixel within 1e-4 relative;
  * the adjoint equals the oracle's FP64 adjoint of the FP32 estimator per gradient
    component (components above 1e-3 of the largest) within 1e-4 relative, environment
    texels and material parameters alike.  The oracle's BSDF derivatives are FP64
    central differences; the kernel's are dual numbers, with FP64 tangents in the
    specular lobe, where (alpha^2 - 1) cos^2 + 1 at the scene generator's roughness
    floor (alpha^2 ~ 1e-5) would lose 1e-3 of a tangent to FP32 rounding.  Measured:
    2.3e-5 at most;
  * through the torch.autograd.Function, the gradient of an L2 image loss against a
    target rendered at parameters perturbed by N(0, 0.1) (clamped to [0, 1]) agrees
    with fourth-order central finite differences on 32 random material parameters and
    32 random environment texels (material steps h = 2^-10 ~ 1e-3; env steps larger,
    the loss being quadratic in the texels) within 1e-3 relative (the roughness of
    glossy lobes, where the FP32 estimator cannot resolve such a step, is left to the
    FP64 oracle comparison).
"""
import numpy as np
import pytest
import torch

from zillumgl_amd import scene, render, diff
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

PASSES, DEPTH = 3, 4


@pytest.fixture(scope="module")
def ctx():
    print("[gpu] adjoint: creating context on device 0", flush=True)
    c = render.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c2s():
    return scene.pack(scene.config2(width=96, height=54, n_spheres=6, subdiv=3, env_size=(256, 128), seed=11))


def _records(ps, params):
    em = ps.materials.copy()
    em[:, 0], em[:, 1], em[:, 2], em[:, 3], em[:, 5] = (params[:, k] for k in range(5))
    return em


def _perturb(ps, seed):
    """Parameters perturbed by N(0, 0.1) and clamped to the generator's ranges (raw
    roughness [0.05, 1], others [0, 1]); env texels scaled by lognormal(0, 0.1)."""
    rng = np.random.default_rng(seed)
    p0, e0 = diff.scene_parameters(ps)
    p1 = np.clip(p0 + rng.normal(0, 0.1, p0.shape), 0, 1).astype(np.float32)
    p1[:, 3] = np.maximum(p1[:, 3], np.float32(0.05))
    e1 = np.clip(e0 * rng.lognormal(0, 0.1, e0.shape), 0, None).astype(np.float32)
    return p0, e0, p1, e1


def test_replay_at_uploaded_parameters_equals_forward(ctx, c2s):
    print("[gpu] adjoint: replay at uploaded parameters vs zgl_render", flush=True)
    ctx.upload(c2s)
    ctx.clear()
    ctx.render(0, 0, PASSES, 1, DEPTH)
    fwd = ctx.read_frame()[..., :3]
    ctx.set_eval_params(None, None)
    ctx.clear()
    ctx.render_replay(0, PASSES, 1, DEPTH)
    rep = ctx.read_frame()[..., :3]
    assert np.array_equal(fwd.view(np.uint32), rep.view(np.uint32))
    # setting the uploaded values explicitly gives the same frame
    p0, e0 = diff.scene_parameters(c2s)
    ctx.set_eval_params(p0, e0)
    ctx.clear()
    ctx.render_replay(0, PASSES, 1, DEPTH)
    assert np.array_equal(fwd.view(np.uint32), ctx.read_frame()[..., :3].view(np.uint32))


def test_replay_at_perturbed_parameters_matches_oracle(ctx, c2s):
    print("[gpu] adjoint: replay at perturbed parameters vs oracle", flush=True)
    ctx.upload(c2s)
    _, _, p1, e1 = _perturb(c2s, 1)
    ctx.set_eval_params(p1, e1)
    ctx.clear()
    ctx.render_replay(0, PASSES, 1, DEPTH)
    gpu = ctx.read_frame()[..., :3]
    ref = orc.render_replay(c2s, orc.params(0, 0, PASSES, 1, DEPTH), _records(c2s, p1), e1)
    err = np.abs(gpu.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-3)
    print(f"[gpu] adjoint: replay max rel err {err.max():.3g}", flush=True)
    assert err.max() <= 1e-4
    assert not np.array_equal(gpu, orc.render_pixels(c2s, orc.params(0, 0, PASSES, 1, DEPTH))[0])


def _close(g, ref, rel):
    floor = 1e-3 * np.abs(ref).max()
    m = np.abs(ref) > floor
    err = np.abs(g.astype(np.float64) - ref) / np.maximum(np.abs(ref), floor)
    return err[m].max(), m.sum()


def test_adjoint_matches_oracle_fp64(ctx, c2s):
    print("[gpu] adjoint: GPU adjoint vs oracle FP64 adjoint", flush=True)
    ctx.upload(c2s)
    rng = np.random.default_rng(2)
    _, _, p1, e1 = _perturb(c2s, 2)
    d_frame = rng.uniform(-1, 1, (c2s.height, c2s.width, 3)).astype(np.float32)
    ctx.set_eval_params(p1, e1)
    gm, ge = ctx.adjoint(d_frame, p1.shape[0], e1.shape, 0, PASSES, 1, DEPTH)
    rm, re = orc.adjoint(c2s, orc.params(0, 0, PASSES, 1, DEPTH), _records(c2s, p1), e1, d_frame, fp32_values=True)
    em, nm = _close(gm, rm, 1e-3)
    ee, ne = _close(ge, re, 1e-3)
    print(f"[gpu] adjoint: material max rel err {em:.3g} over {nm}, env {ee:.3g} over {ne}", flush=True)
    assert nm >= 20 and ne >= 1000
    assert em <= 1e-4 and ee <= 1e-4


def test_adjoint_offset_passes_and_split_launches(ctx, c2s):
    """Passes that start past 0 (first_pass 5, free_counter 9) against the oracle, with
    Russian roulette on; and the gradients of passes [5, 11) equal those of [5, 8) plus
    [8, 11) (each launch logs and differentiates its own passes), up to the order of the
    FP64 atomic sums."""
    print("[gpu] adjoint: offset passes vs oracle, split launches", flush=True)
    ctx.upload(c2s)
    rng = np.random.default_rng(6)
    _, _, p1, e1 = _perturb(c2s, 6)
    d_frame = rng.uniform(-1, 1, (c2s.height, c2s.width, 3)).astype(np.float32)
    ctx.set_eval_params(p1, e1)
    gm, ge = ctx.adjoint(d_frame, p1.shape[0], e1.shape, 5, 6, 9, DEPTH, russian_roulette=True)
    rm, re = orc.adjoint(c2s, orc.params(0, 5, 6, 9, DEPTH, True), _records(c2s, p1), e1, d_frame, fp32_values=True)
    em, nm = _close(gm, rm, 1e-3)
    ee, ne = _close(ge, re, 1e-3)
    print(f"[gpu] adjoint offset: material max rel err {em:.3g} over {nm}, env {ee:.3g} over {ne}", flush=True)
    assert nm >= 20 and ne >= 1000
    assert em <= 1e-4 and ee <= 1e-4
    am, ae = ctx.adjoint(d_frame, p1.shape[0], e1.shape, 5, 3, 9, DEPTH, russian_roulette=True)
    bm, be = ctx.adjoint(d_frame, p1.shape[0], e1.shape, 8, 3, 12, DEPTH, russian_roulette=True)
    for whole, parts in ((gm, am.astype(np.float64) + bm), (ge, ae.astype(np.float64) + be)):
        scale = np.abs(whole).max()
        assert np.abs(whole - parts).max() <= 1e-5 * scale


def test_adjoint_deep_paths_match_oracle_fp64(ctx, c2s):
    """Depth 10 issues up to 21 traversal queries per sample, more than the replay's hit
    log holds (16): the queries past it are walked again in the replay pass."""
    print("[gpu] adjoint: depth 10 vs oracle FP64 adjoint", flush=True)
    ctx.upload(c2s)
    rng = np.random.default_rng(4)
    _, _, p1, e1 = _perturb(c2s, 3)
    d_frame = rng.uniform(-1, 1, (c2s.height, c2s.width, 3)).astype(np.float32)
    ctx.set_eval_params(p1, e1)
    gm, ge = ctx.adjoint(d_frame, p1.shape[0], e1.shape, 0, 2, 1, 10)
    rm, re = orc.adjoint(c2s, orc.params(0, 0, 2, 1, 10), _records(c2s, p1), e1, d_frame, fp32_values=True)
    em, nm = _close(gm, rm, 1e-3)
    ee, ne = _close(ge, re, 1e-3)
    print(f"[gpu] adjoint depth 10: material max rel err {em:.3g} over {nm}, env {ee:.3g} over {ne}", flush=True)
    assert nm >= 20 and ne >= 1000
    assert em <= 1e-4 and ee <= 1e-4


def test_adjoint_reuses_the_image_log_only_for_its_scene_and_passes(ctx, c2s):
    """zgl_render_replay logs every sample's traversal results; a zgl_adjoint of the same
    passes right after differentiates from that log, and after another scene is uploaded
    (or other passes are asked for) it logs its own.  Each against the oracle's FP64
    adjoint: a log of other rays would give other gradients."""
    print("[gpu] adjoint: traversal log of the image reused, and not across scenes", flush=True)
    rng = np.random.default_rng(8)
    d_frame = rng.uniform(-1, 1, (c2s.height, c2s.width, 3)).astype(np.float32)
    _, _, p1, e1 = _perturb(c2s, 8)
    prm = orc.params(0, 0, PASSES, 1, DEPTH)
    ctx.upload(c2s)
    ctx.set_eval_params(p1, e1)
    ctx.clear()
    ctx.render_replay(0, PASSES, 1, DEPTH)
    gm, ge = ctx.adjoint(d_frame, p1.shape[0], e1.shape, 0, PASSES, 1, DEPTH)
    rm, re = orc.adjoint(c2s, prm, _records(c2s, p1), e1, d_frame, fp32_values=True)
    em, _ = _close(gm, rm, 1e-3)
    ee, _ = _close(ge, re, 1e-3)
    print(f"[gpu] adjoint after the image: material {em:.3g}, env {ee:.3g}", flush=True)
    assert em <= 1e-4 and ee <= 1e-4
    # the image's log of this scene, then another scene of the same size
    other = scene.pack(scene.config2(width=96, height=54, n_spheres=6, subdiv=3, env_size=(256, 128), seed=12))
    ctx.clear()
    ctx.render_replay(0, PASSES, 1, DEPTH)
    ctx.upload(other)
    p0o, e0o = diff.scene_parameters(other)
    gm2, ge2 = ctx.adjoint(d_frame, p0o.shape[0], e0o.shape, 0, PASSES, 1, DEPTH)
    rm2, re2 = orc.adjoint(other, prm, _records(other, p0o), e0o, d_frame, fp32_values=True)
    em2, _ = _close(gm2, rm2, 1e-3)
    ee2, _ = _close(ge2, re2, 1e-3)
    print(f"[gpu] adjoint of another scene: material {em2:.3g}, env {ee2:.3g}", flush=True)
    assert em2 <= 1e-4 and ee2 <= 1e-4
    # other passes of the same scene: the log of passes [0, 3) is not theirs
    gm3, ge3 = ctx.adjoint(d_frame, p0o.shape[0], e0o.shape, 4, PASSES, 5, DEPTH)
    rm3, re3 = orc.adjoint(other, orc.params(0, 4, PASSES, 5, DEPTH), _records(other, p0o), e0o, d_frame,
                           fp32_values=True)
    em3, _ = _close(gm3, rm3, 1e-3)
    ee3, _ = _close(ge3, re3, 1e-3)
    assert em3 <= 1e-4 and ee3 <= 1e-4


def test_autograd_l2_gradient_matches_central_differences(ctx, c2s):
    print("[gpu] adjoint: torch autograd L2 loss vs central differences", flush=True)
    ctx.upload(c2s)
    rd = diff.Renderer(ctx, c2s, 16, DEPTH)
    p0, e0, pt, et = _perturb(c2s, 5)
    target = torch.from_numpy(rd.image(pt, et))
    theta_m = torch.tensor(p0, requires_grad=True)
    theta_e = torch.tensor(e0, requires_grad=True)
    loss = ((diff.render_image(theta_m, theta_e, rd) - target) ** 2).sum()
    loss.backward()
    gm, ge = theta_m.grad.numpy(), theta_e.grad.numpy()

    def loss_at(pm, pe):
        img = rd.image(pm, pe).astype(np.float64)
        return ((img - target.numpy().astype(np.float64)) ** 2).sum()

    def fd(kind, i):
        """Fourth-order central difference, steps h and 2h.  Materials: h = 2^-10 of the
        scale.  Env texels: the image is linear in them and the loss quadratic, so the
        difference is exact for any step; h = 1/4 of the scale keeps the change of every
        pixel far above FP32 resolution."""
        base = (p0 if kind == "m" else e0).ravel()
        h = np.float32((2.0 ** -10 if kind == "m" else 0.25) * max(1.0, abs(float(base[i]))))
        vals = {}
        for m in (-2, -1, 1, 2):
            q = base.copy()
            q[i] = np.float32(base[i] + np.float32(m) * h)
            q = q.reshape(p0.shape if kind == "m" else e0.shape)
            vals[m] = (loss_at(q, e0) if kind == "m" else loss_at(p0, q), float(q.ravel()[i]) - float(base[i]))
        c = {m: (vals[m][0] - vals[-m][0]) / (vals[m][1] - vals[-m][1]) for m in (1, 2)}
        return (4 * c[1] - c[2]) / 3

    # 64 random parameters with a gradient.  Excluded: the roughness of materials whose
    # GGX term (alpha^2 - 1) cos^2 + 1 cannot resolve a 1e-3 step in FP32 to 1e-4, i.e.
    # 2 h dalpha^2/dr < 2^-23 / 1e-4 (alpha^2 = r'^4, r' = mix(0.0134, 1, r)): there the
    # FP32 estimator is a staircase at the step's scale, and the adjoint is checked
    # against the oracle's FP64 adjoint instead (test_adjoint_matches_oracle_fp64).
    rough = 0.0134 * (1 - p0[:, 3].astype(np.float64)) + p0[:, 3]
    resolvable = 2 * 2.0 ** -10 * 4 * rough ** 3 * 0.9866 >= 2.0 ** -23 / 1e-4
    # gradients below 1e-3 of the largest of their kind are numerically zero for a
    # difference of FP32 images and are not sampled
    cand_m = [i for i in range(gm.size)
              if abs(gm.ravel()[i]) >= 1e-3 * np.abs(gm).max() and (i % 5 != 3 or resolvable[i // 5])]
    cand_e = list(np.nonzero(np.abs(ge.ravel()) >= 1e-3 * np.abs(ge).max())[0])
    # half of the draws from the material parameters: drawn with the env texels, whose
    # candidates outnumber them a thousandfold, they would almost never be checked
    rng = np.random.default_rng(9)
    n_m = min(32, len(cand_m))
    assert n_m >= 16
    params = [("m", cand_m[k]) for k in rng.choice(len(cand_m), n_m, replace=False)]
    params += [("e", cand_e[k]) for k in rng.choice(len(cand_e), 64 - n_m, replace=False)]
    worst = 0.0
    for kind, i in params:
        f4 = fd(kind, i)
        g = float((gm if kind == "m" else ge).ravel()[i])
        rel = abs(g - f4) / max(abs(f4), 1e-30)
        worst = max(worst, rel)
        assert rel <= 1e-3, (kind, i, g, f4, rel)
    print(f"[gpu] adjoint: 64 central differences ({n_m} material, {64 - n_m} env), worst rel err {worst:.3g}",
          flush=True)
