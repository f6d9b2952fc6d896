"""The C ABI library: symbols, error behaviour without a GPU, host packing semantics."""
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
