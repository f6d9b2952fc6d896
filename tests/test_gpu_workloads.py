"""GPU tests on the full-size batched and differentiable workloads.

Batched workload: scenes drawn from the 64-scene distribution as the batch bench draws
them (1k-2M triangles, 128^2-1024^2, 4-64 spp, depth 3-10), among them scene 34 (1.29M
triangles at 1024^2, 45 spp, depth 8): one batch launch writes every frame bit for bit
as per-scene zgl_render does, and scene 34's frame equals the oracle's on bands of rows.

Differentiable workload: the 201,588-triangle Cornell box with 32 Disney materials and a
256 x 128 environment at 512^2, depth 6: the adjoint against the oracle's FP64 adjoint of
the FP32 estimator over the whole image, and the gradient of the L2 loss at 32 spp against
fourth-order central differences on 32 random material parameters and 32 random
environment texels.

Tolerances: frames bit-exact; adjoint components above 1e-3 of the largest of their kind
within 1e-4 relative; central differences within 1e-3 relative."""
import numpy as np
import pytest
import torch

from zillumgl_amd import scene, render, diff, _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

C4_PICK = [0, 2, 12, 25, 34, 54]
C4_BIG = 34


@pytest.fixture(scope="module")
def c4():
    descs = scene.config4(only=C4_PICK)
    out = []
    for k, (d, spp, depth) in zip(C4_PICK, descs):
        print(f"[gpu] batched workload: packing scene {k} ({d.width}^2, {spp} spp, depth {depth})", flush=True)
        out.append((k, scene.pack(d), spp, depth))
    return out


def test_batched_workload_equals_per_scene_and_oracle(c4):
    ctxs = []
    try:
        for k, ps, spp, depth in c4:
            c = render.Context(0)
            c.upload(ps)
            ctxs.append(c)
        refs = []
        for c, (k, ps, spp, depth) in zip(ctxs, c4):
            print(f"[gpu] batched workload: zgl_render scene {k}", flush=True)
            c.clear()
            c.render(_lib.INTEGRATOR_NAIVE, 0, spp, 1, depth)
            refs.append(c.read_frame())
        b = render.Batch(0)
        try:
            b.set(ctxs, [(0, spp, 1, depth) for (_, _, spp, depth) in c4])
            for c in ctxs:
                c.clear()
            print("[gpu] batched workload: one batch launch", flush=True)
            b.launch()
            b.synchronize()
            for c, r, (k, ps, _, _) in zip(ctxs, refs, c4):
                f = c.read_frame()
                assert np.array_equal(f.view(np.uint32), r.view(np.uint32)), f"scene {k}"
        finally:
            b.close()
        # the largest 1024^2 scene against the oracle on two bands of 8 rows
        i = C4_PICK.index(C4_BIG)
        k, ps, spp, depth = c4[i]
        assert ps.indices.shape[0] > 1_000_000 and ps.width == 1024
        for y0 in (256, 768):
            print(f"[gpu] batched workload: scene {k} rows {y0}..{y0 + 7} vs oracle", flush=True)
            ref, _ = orc.render_pixels(ps, orc.params(0, 0, spp, 1, depth), 0, y0, ps.width, 8)
            got = refs[i][y0:y0 + 8, :, :3]
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), y0
    finally:
        for c in ctxs:
            c.close()


@pytest.fixture(scope="module")
def c5():
    print("[gpu] differentiable workload: packing", flush=True)
    return scene.pack(scene.config5())


@pytest.fixture(scope="module")
def ctx5(c5):
    c = render.Context(0)
    c.upload(c5)
    yield c
    c.close()


def _records(ps, params):
    em = ps.materials.copy()
    em[:, 0], em[:, 1], em[:, 2], em[:, 3], em[:, 5] = (params[:, k] for k in range(5))
    return em


def _perturb(ps, seed):
    rng = np.random.default_rng(seed)
    p0, e0 = diff.scene_parameters(ps)
    p1 = np.clip(p0 + rng.normal(0, 0.1, p0.shape), 0, 1).astype(np.float32)
    p1[:, 3] = np.maximum(p1[:, 3], np.float32(0.05))
    e1 = np.clip(e0 * rng.lognormal(0, 0.1, e0.shape), 0, None).astype(np.float32)
    return p0, e0, p1, e1


def _resolvable_roughness(params):
    """Raw roughness whose GGX term (alpha^2 - 1) cos^2 + 1 resolves a 1e-3 step to 1e-4
    in FP32 (alpha^2 = r'^4, r' = mix(0.0134, 1, r))."""
    r = 0.0134 * (1 - params[:, 3].astype(np.float64)) + params[:, 3]
    return 2 * 2.0 ** -10 * 4 * r ** 3 * 0.9866 >= 2.0 ** -23 / 1e-4


def test_differentiable_workload_adjoint_matches_oracle_fp64(ctx5, c5):
    _, _, p1, e1 = _perturb(c5, 21)
    d_frame = np.random.default_rng(22).uniform(-1, 1, (c5.height, c5.width, 3)).astype(np.float32)
    ctx5.set_eval_params(p1, e1)
    print("[gpu] differentiable workload: adjoint, 2 passes, depth 6", flush=True)
    gm, ge = ctx5.adjoint(d_frame, p1.shape[0], e1.shape, 0, 2, 1, 6)
    print("[gpu] differentiable workload: oracle FP64 adjoint", flush=True)
    rm, re = orc.adjoint(c5, orc.params(0, 0, 2, 1, 6), _records(c5, p1), e1, d_frame, fp32_values=True)
    ctx5.set_eval_params(None, None)
    floor_m = 1e-3 * np.abs(rm).max()
    floor_e = 1e-3 * np.abs(re).max()
    err_m = np.abs(gm - rm) / np.maximum(np.abs(rm), floor_m)
    err_e = np.abs(ge - re) / np.maximum(np.abs(re), floor_e)
    sig_m = np.abs(rm) > floor_m
    sig_e = np.abs(re) > floor_e
    loose = np.zeros_like(sig_m)
    loose[:, 3] = ~_resolvable_roughness(p1)
    print(f"[gpu] differentiable workload: materials max rel err {err_m[sig_m & ~loose].max():.3g} "
          f"(unresolvable roughness {err_m[sig_m & loose].max() if (sig_m & loose).any() else 0:.3g}), "
          f"env {err_e[sig_e].max():.3g} over {sig_e.sum()}", flush=True)
    assert sig_m.sum() >= 100 and sig_e.sum() >= 5000
    assert err_m[sig_m].max() <= 1e-4
    assert err_e[sig_e].max() <= 1e-4


def test_differentiable_workload_gradient_matches_central_differences(ctx5, c5):
    rd = diff.Renderer(ctx5, c5, 32, 6)
    p0, e0, pt, et = _perturb(c5, 23)
    print("[gpu] differentiable workload: target and autograd gradient, 32 spp", flush=True)
    target = torch.from_numpy(rd.image(pt, et))
    theta_m = torch.tensor(p0, requires_grad=True)
    theta_e = torch.tensor(e0, requires_grad=True)
    loss = ((diff.render_image(theta_m, theta_e, rd) - target) ** 2).sum()
    loss.backward()
    gm, ge = theta_m.grad.numpy(), theta_e.grad.numpy()
    tgt = target.numpy().astype(np.float64)

    def loss_at(pm, pe):
        return ((rd.image(pm, pe).astype(np.float64) - tgt) ** 2).sum()

    def fd(kind, i):
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

    resolvable = _resolvable_roughness(p0)
    cand_m = [i for i in range(gm.size)
              if abs(gm.ravel()[i]) >= 1e-3 * np.abs(gm).max() and (i % 5 != 3 or resolvable[i // 5])]
    cand_e = list(np.nonzero(np.abs(ge.ravel()) >= 1e-3 * np.abs(ge).max())[0])
    # half of the draws from the material parameters (they would almost never be drawn
    # with the far more numerous env texels)
    rng = np.random.default_rng(24)
    n_m = min(32, len(cand_m))
    assert n_m >= 16
    params = [("m", cand_m[k]) for k in rng.choice(len(cand_m), n_m, replace=False)]
    params += [("e", cand_e[k]) for k in rng.choice(len(cand_e), 64 - n_m, replace=False)]
    worst = 0.0
    for n, (kind, i) in enumerate(params):
        f4 = fd(kind, i)
        g = float((gm if kind == "m" else ge).ravel()[i])
        rel = abs(g - f4) / max(abs(f4), 1e-30)
        worst = max(worst, rel)
        if n % 16 == 15:
            print(f"[gpu] differentiable workload: {n + 1} differences ({kind}), worst rel err {worst:.3g}",
                  flush=True)
        assert rel <= 1e-3, (kind, i, g, f4, rel)
