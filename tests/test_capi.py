"""The C ABI library: symbols, error behaviour without a GPU, host packing semantics."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from zillumgl_amd import _lib, scene
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "zgl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zgl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def test_create_without_device_reports_nodevice():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    import ctypes as C
    h = C.c_void_p()
    assert _lib.lib().zgl_create(0, C.byref(h)) == _lib.ZGL_E_NODEVICE
    assert not h.value


def test_null_context_calls_fail_cleanly():
    L = _lib.lib()
    assert L.zgl_render(None, None) == _lib.ZGL_E_INVALID
    assert L.zgl_synchronize(None) == _lib.ZGL_E_INVALID
    assert L.zgl_last_error(None) == b"null context"


def test_host_producers_reject_bad_input():
    import ctypes as C
    L = _lib.lib()
    v = np.zeros((3, 3), np.float32)
    bad = np.array([[0, 1, 5]], np.uint32)
    b = np.zeros((1, 6), np.float32)
    t = np.zeros((6, 1, 3), np.int32)
    assert L.zgl_build_mtbvh(_lib.ptr(v, C.c_float), 3, _lib.ptr(bad, C.c_uint32), 1, _lib.ptr(b, C.c_float),
                             _lib.ptr(t, C.c_int32)) == _lib.ZGL_E_INVALID
    assert L.zgl_build_mtbvh(_lib.ptr(v, C.c_float), 3, _lib.ptr(bad, C.c_uint32), 0, _lib.ptr(b, C.c_float),
                             _lib.ptr(t, C.c_int32)) == _lib.ZGL_E_INVALID
    assert L.zgl_alias_build(None, -1, None, None) == _lib.ZGL_E_INVALID
    assert L.zgl_env_alias_build(None, 0, 0, None, None, None) == _lib.ZGL_E_INVALID


def test_packing_semantics():
    """Scene::createGLContext (src/core/Scene.cpp:133-270)."""
    ps = scene.pack(scene.config3(32, 32))
    # objects first: matTexIndex = offMaterial + (-1 << 16 | 0): texId -1, matId = object index
    mti = ps.mat_tex_indices
    assert np.all((mti >> 16) == -1)
    assert len(mti) == ps.obj_prim_count
    mids = mti & 0xFFFF
    assert mids.min() == 0 and mids.max() == ps.materials.shape[0] - 1
    # light triangles at the end; their powers add up to each mesh's XML radiance
    assert ps.n_light_triangles == 2 + 4
    assert np.allclose(ps.light_power[:2].sum(0), [17, 12, 4], rtol=1e-6)
    assert np.allclose(ps.light_power[2:], 50.0, rtol=1e-6)
    lum = ps.light_power @ np.float32([0.299, 0.587, 0.114])
    assert abs(ps.light_sum - lum.sum()) < 1e-4 * lum.sum()
    # the material type int lives in slot 13 of the 64-byte record
    types = ps.materials.view(np.int32)[:, 13]
    assert scene.MAT_DIELECTRIC in types and scene.MAT_LAMBERTIAN in types
    # no env map: a black 1x1 one with uEnvSum = 0
    assert ps.env_width == 1 and ps.env_sum == 0.0


def test_camera_uniforms_orthonormal_and_roundtrip():
    d = scene.config2(width=64, height=36, n_spheres=1, subdiv=0, env_size=(16, 8))
    ps = scene.pack(d)
    c = ps.camera
    F, R, U = (np.array(x[:], np.float64) for x in (c.front, c.right, c.up))
    for a in (F, R, U):
        assert abs(np.linalg.norm(a) - 1) < 1e-6
    assert abs(F @ R) < 1e-6 and abs(F @ U) < 1e-6 and abs(R @ U) < 1e-6
    Minv = np.array(c.mat_inv[:]).reshape(3, 3).T
    assert np.allclose(Minv @ np.stack([R, U, F], 1), np.eye(3), atol=1e-6)
    assert abs(c.tan_fov - np.tan(np.radians(22.5))) < 1e-6
    for u, v in [(0.5, 0.5), (0.1, 0.9), (0.73, 0.21)]:
        uv = orc.camera_roundtrip(ps, u, v)
        assert np.allclose(uv, [u, v], atol=2e-6)


def _plan(wh, passes, first=None):
    n = len(wh)
    films = np.ascontiguousarray(np.array(wh, np.int32).ravel())
    params = (_lib.PassParams * n)()
    for i in range(n):
        params[i] = _lib.PassParams(_lib.INTEGRATOR_NAIVE, 0 if first is None else first[i], passes[i], 1, 8, 0, 1, 0,
                                    0.5, 32, 1, 0, 0)
    nb, sf, sc = C.c_int64(0), C.c_int32(0), C.c_int32(0)
    st = _lib.lib().zgl_batch_plan(_lib.ptr(films, C.c_int32), params, n, C.byref(nb), C.byref(sf), C.byref(sc))
    return st, nb.value, sf.value, sc.value


def test_batch_plan_limits():
    """The batch checks, without a device: each scene's (8x8 tile, pass) work units stay
    below 2^31 - 2^24 (the persistent kernel counts them in int32), the sample values
    below 64 GiB, and the Sobol table below 65,537 passes however far apart the scenes'
    pass ranges are."""
    st, nb, sf, sc = _plan([(1920, 1080), (128, 128)], [64, 4])
    assert st == _lib.ZGL_OK and nb == (1920 * 1080 * 64 + 128 * 128 * 4) * 16 and (sf, sc) == (0, 65)
    # 1920 x 1080 is 240 x 135 tiles of 64 units: 2,073,600 units per pass; 1000 passes fit
    # the unit limit, 1036 do not (2^31 - 2^24 = 2,130,706,432)
    assert _plan([(1920, 1080)], [1000])[0] == _lib.ZGL_OK
    assert _plan([(1920, 1080)], [1036])[0] == _lib.ZGL_E_INVALID
    # 64 GiB of sample values: 64 scenes x 1024^2 pixels x 64 passes x 16 B = 2^36 B exactly
    assert _plan([(1024, 1024)] * 64, [64] * 64)[0] == _lib.ZGL_OK
    assert _plan([(1024, 1024)] * 65, [1024] * 65)[0] == _lib.ZGL_E_INVALID
    # far-apart pass ranges: the table is capped, the rest comes from the generator matrices
    st, _, sf, sc = _plan([(64, 64), (64, 64)], [4, 4], first=[0, 1 << 24])
    assert st == _lib.ZGL_OK and sf == 0 and sc == 65537
    assert _plan([(0, 64)], [4])[0] == _lib.ZGL_E_INVALID
    assert _plan([(64, 64)], [0])[0] == _lib.ZGL_E_INVALID
