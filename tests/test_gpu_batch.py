"""GPU parity of the batched persistent launch: one graph launch over several contexts
writes into every frame exactly (bit for bit) what zgl_render of the same passes writes,
whatever the scene order, regeneration threshold and blocks per CU, and after tuning;
so do the per-pass form it is timed against and the differentiable mode's replay at the
uploaded parameters."""
import numpy as np
import pytest

from zillumgl_amd import scene, render, _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_scenes():
    descs = scene.config4(n_scenes=6, seed=9, min_tris=200, max_tris=40_000)
    rng = np.random.default_rng(3)
    out = []
    for k, (d, spp, depth) in enumerate(descs):
        d.width, d.height = int(rng.choice([24, 40, 64])), int(rng.choice([24, 40, 64]))
        out.append((scene.pack(d), int(rng.integers(1, 5)), depth, int(rng.integers(0, 3))))
    return out


@pytest.fixture(scope="module")
def contexts(batch_scenes):
    print("[gpu] batch: creating 6 contexts", flush=True)
    cs = []
    for ps, _, _, _ in batch_scenes:
        c = render.Context(0)
        c.upload(ps)
        cs.append(c)
    yield cs
    for c in cs:
        c.close()


def _reference(contexts, batch_scenes):
    refs = []
    for c, (ps, spp, depth, first) in zip(contexts, batch_scenes):
        c.clear()
        c.render(_lib.INTEGRATOR_NAIVE, first, spp, 1 + first, depth)
        refs.append(c.read_frame())
    return refs


def _batch_frames(b, contexts):
    b.launch()
    b.synchronize()
    return [c.read_frame() for c in contexts]


def test_batch_equals_per_scene_launches(contexts, batch_scenes):
    print("[gpu] batch: 6 scenes, one graph launch vs 6 zgl_render", flush=True)
    refs = _reference(contexts, batch_scenes)
    b = render.Batch(0)
    b.set(contexts, [(first, spp, 1 + first, depth) for (_, spp, depth, first) in batch_scenes])
    for c in contexts:
        c.clear()
    frames = _batch_frames(b, contexts)
    for k, (f, r) in enumerate(zip(frames, refs)):
        assert np.array_equal(f.view(np.uint32), r.view(np.uint32)), k
    assert all((r[..., :3] > 0).any() for r in refs)
    # other orders / thresholds / occupancy: the same frames
    cfg = b.config()
    for th, bpc, order in [(1, 0, [5, 4, 3, 2, 1, 0]), (63, 1, [2, 0, 4, 1, 5, 3]),
                           (16, cfg["max_blocks_per_cu"], None)]:
        b.configure(th, bpc, order)
        for c in contexts:
            c.clear()
        frames = _batch_frames(b, contexts)
        for k, (f, r) in enumerate(zip(frames, refs)):
            assert np.array_equal(f.view(np.uint32), r.view(np.uint32)), (th, bpc, order, k)
    b.close()


def test_replay_and_per_pass_forms_equal_per_scene_launches(contexts, batch_scenes):
    """zgl_render_replay at the uploaded parameters (the logged forward pass and the
    replay from its log) and one zgl_render per pass write the frames zgl_render of all the
    passes writes."""
    print("[gpu] batch: replay and per-pass forms vs zgl_render", flush=True)
    refs = _reference(contexts, batch_scenes)
    for k, (c, (ps, spp, depth, first)) in enumerate(zip(contexts, batch_scenes)):
        c.clear()
        c.render_replay(first, spp, 1 + first, depth)
        assert np.array_equal(c.read_frame().view(np.uint32), refs[k].view(np.uint32)), k
        c.clear()
        for j in range(spp):
            c.render(_lib.INTEGRATOR_NAIVE, first + j, 1, 1 + first + j, depth)
        assert np.array_equal(c.read_frame().view(np.uint32), refs[k].view(np.uint32)), k


def test_batch_with_roulette_and_options_equals_per_scene_launches(contexts, batch_scenes):
    """Per-scene options in one batch (roulette on some scenes, NEE off or uniform light
    selection on others): every frame equals zgl_render of the same options bit for bit."""
    print("[gpu] batch: mixed options incl. roulette", flush=True)
    opts = [dict(russian_roulette=True), dict(sample_light=False), dict(light_env_uniform=True, light_portion=0.3),
            dict(russian_roulette=True, sample_light=False), dict(), dict(russian_roulette=True)]
    params = []
    refs = []
    for c, (ps, spp, depth, first), o in zip(contexts, batch_scenes, opts):
        kw = dict(first_pass=first, n_passes=spp, free_counter=1 + first, max_depth=depth, **o)
        params.append(kw)
        c.clear()
        c.render(_lib.INTEGRATOR_NAIVE, first, spp, 1 + first, depth, **o)
        refs.append(c.read_frame())
    b = render.Batch(0)
    b.set(contexts, params)
    for c in contexts:
        c.clear()
    frames = _batch_frames(b, contexts)
    for k, (f, r) in enumerate(zip(frames, refs)):
        assert np.array_equal(f.view(np.uint32), r.view(np.uint32)), k
    b.close()


def test_batch_tune_keeps_results(contexts, batch_scenes):
    print("[gpu] batch: autotuner", flush=True)
    passes = [(0, spp, 1, depth) for (_, spp, depth, _) in batch_scenes]
    for c, (ps, spp, depth, _) in zip(contexts, batch_scenes):
        c.clear()
        c.render(_lib.INTEGRATOR_NAIVE, 0, spp, 1, depth)
    refs = [c.read_frame() for c in contexts]
    b = render.Batch(0)
    b.set(contexts, passes)
    best = b.tune(reps=1)
    cfg = b.config()
    assert best > 0 and 0 <= cfg["regen_threshold"] <= 63 and sorted(cfg["order"]) == list(range(6))
    frames = _batch_frames(b, contexts)
    for k, (f, r) in enumerate(zip(frames, refs)):
        assert np.array_equal(f.view(np.uint32), r.view(np.uint32)), k
    b.close()


def test_batch_rejects_bad_input(contexts):
    b = render.Batch(0)
    with pytest.raises(_lib.ZglError):
        b.set([contexts[0], contexts[0]], [(0, 1, 1, 3), (0, 1, 1, 3)])
    with pytest.raises(_lib.ZglError):
        b.set([contexts[0]], [(0, 0, 1, 3)])
    b.set([contexts[0], contexts[1]], [(0, 1, 1, 3), (0, 1, 1, 3)])
    with pytest.raises(_lib.ZglError):
        b.configure(16, 0, [0, 0])
    with pytest.raises(_lib.ZglError):
        b.configure(64, 0, None)
    b.close()
