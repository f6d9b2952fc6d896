"""zgl_render_static -- the naive integrator with its unit counters, Sobol points and
per-sample values in a static device array, allocating nothing -- against zgl_render,
through the C ABI.

Two calls in a row, the second on a smaller film with fewer passes, each checked against
zgl_render's frame bit for bit (the same estimator at the same Sobol indices, the passes
added in the same order). The first call's film takes two chunks of the static array.
Sizes beyond the array and other integrators are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from zillumgl_amd import scene, render, _lib

pytestmark = pytest.mark.gpu


def _params(first, n, depth, rr, threshold=0):
    return _lib.PassParams(_lib.INTEGRATOR_NAIVE, first, n, first, depth, int(rr), 1, 0, 0.5, 32, 1, 0, threshold)


def _frames(ctx, p):
    L = _lib.lib()
    ctx.clear()
    assert L.zgl_render(ctx.h, C.byref(p)) == _lib.ZGL_OK, L.zgl_last_error(ctx.h)
    a = ctx.read_frame()
    ctx.clear()
    st = L.zgl_render_static(ctx.h, C.byref(p))
    assert st == _lib.ZGL_OK, L.zgl_last_error(ctx.h)
    b = ctx.read_frame()
    return a, b


def test_render_static_equals_render_on_two_calls():
    ctx = render.Context(0)
    try:
        # 1024 x 576: 9.4 MB of sample values per pass, 14 passes per chunk, 20 passes
        big = scene.pack(scene.config2(width=1024, height=576, n_spheres=6, subdiv=3, env_size=(256, 128)))
        print("[gpu] no-alloc: 1024x576 Disney + env, 20 passes, depth 6, roulette", flush=True)
        ctx.upload(big)
        a, b = _frames(ctx, _params(3, 20, 6, True))
        assert np.isfinite(a).all() and a[..., :3].mean() > 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
            f"{np.count_nonzero(a.view(np.uint32) != b.view(np.uint32))} texels differ"
        # a smaller input next: the array still holds the first call's values
        small = scene.pack(scene.config1(48, 40))
        print("[gpu] no-alloc: 48x40 Cornell box, 5 passes, depth 5", flush=True)
        ctx.upload(small)
        a, b = _frames(ctx, _params(0, 5, 5, False, threshold=40))
        assert a[..., :3].mean() > 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
            f"{np.count_nonzero(a.view(np.uint32) != b.view(np.uint32))} texels differ"
    finally:
        ctx.close()


def test_render_static_refuses_what_does_not_fit():
    L = _lib.lib()
    ctx = render.Context(0)
    try:
        ctx.upload(scene.pack(scene.config1(32, 32)))
        p = _params(0, 1, 3, False)
        p.integrator = _lib.INTEGRATOR_LIGHT
        assert L.zgl_render_static(ctx.h, C.byref(p)) == _lib.ZGL_E_INVALID
        assert b"naive integrator only" in L.zgl_last_error(ctx.h)
        # 4096 x 2048: 134,217,728 bytes of sample values per pass, over the array's share
        print("[gpu] no-alloc: 4096x2048 film refused", flush=True)
        ctx.upload(scene.pack(scene.config1(4096, 2048)))
        assert L.zgl_render_static(ctx.h, C.byref(_params(0, 1, 3, False))) == _lib.ZGL_E_INVALID
        assert b"static scratch" in L.zgl_last_error(ctx.h)
    finally:
        ctx.close()
