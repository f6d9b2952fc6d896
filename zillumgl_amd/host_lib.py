"""ctypes binding of libzgl_host.so (include/zgl_host.h): the C++ host layer's scene
loader, packer and Integrator classes.  Built in-tree with libzgl.so; missing -> raises."""
import ctypes as C
import os

import numpy as np

from . import _lib

HOST_PATH = os.path.join(_lib.HERE, "libzgl_host.so")

_vp = C.c_void_p
SIGNATURES = [
    ("zglh_last_error", C.c_char_p, []),
    ("zglh_load_scene", C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(_vp)]),
    ("zglh_scene_destroy", None, [_vp]),
    ("zglh_scene_buffers", C.c_int32, [_vp, C.POINTER(_lib.SceneBuffers), C.POINTER(_lib.Camera),
                                       C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32)]),
    ("zglh_upload", C.c_int32, [_vp, _vp]),
    ("zglh_integrator_create", C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    ("zglh_integrator_destroy", None, [_vp]),
    ("zglh_integrator_init", C.c_int32, [_vp, _vp, C.c_int32, C.c_int32]),
    ("zglh_integrator_render_one_pass", C.c_int32, [_vp, C.c_int32]),
    ("zglh_integrator_reset", C.c_int32, [_vp]),
    ("zglh_integrator_get_frame", C.c_int32, [_vp, C.POINTER(C.c_float), C.c_int64]),
    ("zglh_integrator_result_scale", C.c_int32, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
]

_host = None


def lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_PATH):
            raise RuntimeError(f"{HOST_PATH} is missing: run `python -m zillumgl_amd.build` first")
        _lib.lib()  # libzgl.so first (the host library links against it)
        L = C.CDLL(HOST_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _host = L
    return _host


def check(st):
    if st != 0:
        raise _lib.ZglError(st, lib().zglh_last_error().decode())


class HostScene:
    """A scene loaded and packed by the C++ host (zglh_load_scene)."""

    def __init__(self, xml_path, device_bvh=False):
        self.h = _vp()
        check(lib().zglh_load_scene(os.fsencode(xml_path), int(device_bvh), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().zglh_scene_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        """The packed buffers as numpy copies, keyed like scene.PackedScene's fields,
        plus the camera struct."""
        sb, cam = _lib.SceneBuffers(), _lib.Camera()
        noise = C.POINTER(C.c_float)()
        w, h = C.c_int32(), C.c_int32()
        check(lib().zglh_scene_buffers(self.h, C.byref(sb), C.byref(cam), C.byref(noise), C.byref(w), C.byref(h)))

        def arr(p, n, shape, dtype):
            if n == 0 or not p:
                return np.zeros(shape, dtype)
            return np.ctypeslib.as_array(p, shape=(n,)).copy().astype(dtype, copy=False).reshape(shape)
        nv, nt, bs = sb.n_vertices, sb.n_triangles, sb.bvh_size
        nl, nm = sb.n_light_triangles, sb.n_materials
        ew, eh = sb.env_width, sb.env_height
        out = dict(
            vertices=arr(sb.vertices, nv * 3, (nv, 3), np.float32),
            normals=arr(sb.normals, nv * 3, (nv, 3), np.float32),
            texcoords=arr(sb.texcoords, sb.n_texcoords * 2, (sb.n_texcoords, 2), np.float32),
            indices=arr(sb.indices, nt * 3, (nt, 3), np.int32),
            bounds=arr(sb.bounds, bs * 6, (bs, 6), np.float32) if sb.bounds else None,
            hit_table=arr(sb.hit_table, 6 * bs * 3, (6, bs, 3), np.int32) if sb.hit_table else None,
            mat_tex_indices=arr(sb.mat_tex_indices, sb.obj_prim_count, (sb.obj_prim_count,), np.int32),
            materials=arr(sb.materials, nm * 16, (nm, 16), np.float32),
            light_power=arr(sb.light_power, nl * 3, (nl, 3), np.float32),
            light_alias=arr(sb.light_alias, nl, (nl,), np.int32),
            light_prob=arr(sb.light_prob, nl, (nl,), np.float32),
            env_half=arr(sb.env_rgb16f, ew * eh * 3, (eh, ew, 3), np.uint16),
            env_alias=arr(sb.env_alias, (ew + 1) * eh, ((ew + 1) * eh,), np.int32),
            env_prob=arr(sb.env_prob, (ew + 1) * eh, ((ew + 1) * eh,), np.float32),
            noise=arr(noise, w.value * h.value * 2, (h.value, w.value, 2), np.float32),
            obj_prim_count=sb.obj_prim_count, n_light_triangles=nl, light_sum=sb.light_sum, env_sum=sb.env_sum,
            env_rotation=sb.env_rotation, sampler=sb.sampler, width=w.value, height=h.value)
        return out, cam

    def upload(self, ctx):
        check(lib().zglh_upload(ctx.h, self.h))
        w, h = C.c_int32(), C.c_int32()
        check(lib().zglh_scene_buffers(self.h, None, None, None, C.byref(w), C.byref(h)))
        ctx.width, ctx.height = w.value, h.value


class HostIntegrator:
    """One of the C++ host's Integrator classes (zglh_integrator_*)."""

    def __init__(self, integrator, max_depth=4, russian_roulette=False):
        self.h = _vp()
        check(lib().zglh_integrator_create(integrator, max_depth, int(russian_roulette), C.byref(self.h)))

    def init(self, ctx):
        self.ctx = ctx
        check(lib().zglh_integrator_init(self.h, ctx.h, ctx.width, ctx.height))

    def render_one_pass(self, n_passes=1):
        check(lib().zglh_integrator_render_one_pass(self.h, n_passes))

    def reset(self):
        check(lib().zglh_integrator_reset(self.h))

    def get_frame(self):
        out = np.empty((self.ctx.height, self.ctx.width, 4), np.float32)
        check(lib().zglh_integrator_get_frame(self.h, _lib.ptr(out, C.c_float), out.size))
        return out

    def result_scale(self):
        s, n = C.c_float(), C.c_int32()
        check(lib().zglh_integrator_result_scale(self.h, C.byref(s), C.byref(n)))
        return s.value, n.value

    def close(self):
        if self.h:
            lib().zglh_integrator_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
