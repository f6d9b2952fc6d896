"""GPU parity at the edges of the input space, against the oracle on the same inputs.

  * Films smaller than one 8x8 tile or with a ragged last tile row and column: 1x1, 7x1,
    3x5 and 13x9 pixels of the Cornell box.  The persistent schedules hand out whole
    tiles, and a lane whose slot falls outside the film fetches again, so here most
    slots of every tile are empty.  Naive (persistent and wavefront schedules) and the
    global-queue tracer bit for bit; the triple tracer within 1e-4 (FP32 splat atomics);
    the FP64 reference to the FP32 rounding of the oracle's FP64 mode (1e-6).
  * Pass indices at the top of the accepted range: first_pass 2^24 (the largest that
    zgl_render accepts) with the free counter past it.  There uSpp * uSampleDim wraps
    modulo 2^32 as GLSL int arithmetic does (path_integ_naive.glsl:163; 2^24 * 256 = 2^32
    gives offset 0), and both sides compute it in unsigned arithmetic: a C++ int product
    would overflow, which is undefined.  Naive and global-queue bit for bit.
"""
import numpy as np
import pytest

from zillumgl_amd import scene, render, _lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (7, 1), (3, 5), (13, 9)]


@pytest.fixture(scope="module")
def ctx():
    print("[gpu] edges: creating context on device 0", flush=True)
    c = render.Context(0)
    yield c
    c.close()


def _frame(ctx, ps, integ, first, n, fc, depth, **kw):
    ctx.upload(ps)
    ctx.clear()
    ctx.render(integ, first, n, fc, depth, **kw)
    return ctx.read_frame()[..., :3]


@pytest.mark.parametrize("size", SIZES)
def test_small_and_ragged_films_match_oracle(ctx, size):
    w, h = size
    ps = scene.pack(scene.config1(w, h))
    for name, integ in (("naive", _lib.INTEGRATOR_NAIVE), ("wavefront", 4),
                        ("global queue", _lib.INTEGRATOR_GLOBAL_QUEUE)):
        print(f"[gpu] edges: {name} {w}x{h}, 4 passes, depth 5, roulette", flush=True)
        gpu = _frame(ctx, ps, integ, 0, 4, 1, 5, russian_roulette=True)
        ref, _ = orc.render_pixels(ps, orc.params(3 if name == "global queue" else 0, 0, 4, 1, 5,
                                                  russian_roulette=True))
        assert gpu.shape == (h, w, 3)
        assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), (name, size)
    print(f"[gpu] edges: naive {w}x{h} at regeneration thresholds 1 and 63", flush=True)
    for th in (1, 63):
        gpu = _frame(ctx, ps, _lib.INTEGRATOR_NAIVE, 2, 3, 3, 6, regen_threshold=th)
        ref, _ = orc.render_pixels(ps, orc.params(0, 2, 3, 3, 6))
        assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32)), (th, size)


@pytest.mark.parametrize("size", SIZES)
def test_small_and_ragged_films_triple_and_fp64_reference(ctx, size):
    w, h = size
    ps = scene.pack(scene.config1(w, h))
    ctx.upload(ps)
    print(f"[gpu] edges: triple {w}x{h}, 1 pass, depth 6", flush=True)
    ctx.clear()
    ctx.render(_lib.INTEGRATOR_TRIPLE, 0, 1, 1, 6, blocks_one_pass=2)
    gpu = ctx.read_film()
    pt, _ = orc.render_pixels(ps, orc.params(2, 0, 1, 1, 6, blocks_one_pass=2))
    lpt, _ = orc.render_light(ps, orc.params(2, 0, 1, 1, 6, blocks_one_pass=2))
    ref = pt.astype(np.float64) + lpt
    err = np.abs(gpu - ref) / np.maximum(np.abs(ref), 1e-3 * max(ref.mean(), 1e-12))
    assert gpu.shape == (h, w, 3)
    assert err.max() <= 1e-4, (size, err.max())
    print(f"[gpu] edges: FP64 reference {w}x{h}, 4 passes, depth 5", flush=True)
    ref64 = ctx.render_reference_f64(0, 4, 1, 5, russian_roulette=True)
    o64, _ = orc.render_pixels(ps, orc.params(0, 0, 4, 1, 5, russian_roulette=True), nthreads=4, fp64=True)
    rel = np.abs(ref64 - o64.astype(np.float64)) / np.maximum(np.abs(o64.astype(np.float64)), 1e-3)
    assert rel.max() <= 1e-6, (size, rel.max())


@pytest.mark.parametrize("integ", [_lib.INTEGRATOR_NAIVE, _lib.INTEGRATOR_GLOBAL_QUEUE])
def test_largest_pass_indices_match_oracle(ctx, integ):
    ps = scene.pack(scene.config1(24, 16))
    first = 1 << 24
    print(f"[gpu] edges: integrator {integ}, passes 2^24 .. 2^24 + 2", flush=True)
    gpu = _frame(ctx, ps, integ, first, 3, first + 1, 5)
    ref, _ = orc.render_pixels(ps, orc.params(3 if integ == _lib.INTEGRATOR_GLOBAL_QUEUE else 0, first, 3,
                                              first + 1, 5))
    assert np.array_equal(gpu.view(np.uint32), ref.view(np.uint32))
    assert (ref > 0).any()
    # passes 2^24 + k have the sample offsets of passes k: with the same free counter
    # they render the same frame
    low = _frame(ctx, ps, integ, 0, 3, first + 1, 5)
    assert np.array_equal(gpu.view(np.uint32), low.view(np.uint32))
    with pytest.raises(RuntimeError):
        ctx.render(integ, first + 1, 1, first + 2, 5)
