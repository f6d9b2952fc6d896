"""ctypes binding of libzgl.so (the C ABI declared in include/zgl.h).

The library is built in-tree (python -m zillumgl_amd.build); importing this module
when it is missing raises, there is no fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzgl.so")

ZGL_OK, ZGL_E_INVALID, ZGL_E_HIP, ZGL_E_OOM, ZGL_E_STATE, ZGL_E_NODEVICE = range(6)
INTEGRATOR_NAIVE, INTEGRATOR_LIGHT, INTEGRATOR_TRIPLE, INTEGRATOR_GLOBAL_QUEUE = 0, 1, 2, 3
INTEGRATOR_NAIVE_WAVEFRONT = 4
STATUS_NAMES = {0: "ZGL_OK", 1: "ZGL_E_INVALID", 2: "ZGL_E_HIP", 3: "ZGL_E_OOM", 4: "ZGL_E_STATE",
                5: "ZGL_E_NODEVICE"}

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)
_u16p = C.POINTER(C.c_uint16)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)


class SceneBuffers(C.Structure):
    _fields_ = [
        ("vertices", _fp), ("normals", _fp), ("n_vertices", C.c_int64),
        ("texcoords", _fp), ("n_texcoords", C.c_int64),
        ("indices", _ip), ("n_triangles", C.c_int64),
        ("bounds", _fp), ("hit_table", _ip), ("bvh_size", C.c_int32),
        ("mat_tex_indices", _ip), ("obj_prim_count", C.c_int32),
        ("materials", _fp), ("n_materials", C.c_int32),
        ("light_power", _fp), ("light_alias", _ip), ("light_prob", _fp),
        ("n_light_triangles", C.c_int32), ("light_sum", C.c_float),
        ("env_rgb16f", _u16p), ("env_width", C.c_int32), ("env_height", C.c_int32),
        ("env_alias", _ip), ("env_prob", _fp), ("env_sum", C.c_float), ("env_rotation", C.c_float),
        ("sampler", C.c_int32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("front", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3), ("pos", C.c_float * 3),
        ("mat_inv", C.c_float * 9), ("tan_fov", C.c_float), ("aspect", C.c_float),
        ("lens_radius", C.c_float), ("focal_dist", C.c_float),
        ("film_width", C.c_int32), ("film_height", C.c_int32),
    ]


class PassParams(C.Structure):
    _fields_ = [
        ("integrator", C.c_int32), ("first_pass", C.c_int32), ("n_passes", C.c_int32),
        ("free_counter", C.c_int32), ("max_depth", C.c_int32), ("russian_roulette", C.c_int32),
        ("sample_light", C.c_int32), ("light_env_uniform", C.c_int32), ("light_portion", C.c_float),
        ("blocks_one_pass", C.c_int32), ("loops_per_pass", C.c_int32), ("count_bytes", C.c_int32),
        ("regen_threshold", C.c_int32),
    ]


# (name, restype, argtypes) for every symbol in include/zgl.h
SIGNATURES = [
    ("zgl_build_mtbvh", C.c_int32, [_fp, C.c_int64, _up, C.c_int64, _fp, _ip]),
    ("zgl_alias_build", C.c_int32, [_fp, C.c_int32, _ip, _fp]),
    ("zgl_env_alias_build", C.c_int32, [_fp, C.c_int32, C.c_int32, _ip, _fp, _fp]),
    ("zgl_sobol_sample", C.c_uint32, [C.c_uint32, C.c_int32, C.c_uint32]),
    ("zgl_sobol_matrices", None, [_up]),
    ("zgl_float_to_half", None, [_fp, _u16p, C.c_int64]),
    ("zgl_device_count", C.c_int32, []),
    ("zgl_create", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("zgl_destroy", None, [C.c_void_p]),
    ("zgl_last_error", C.c_char_p, [C.c_void_p]),
    ("zgl_set_stream", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("zgl_get_stream", C.c_void_p, [C.c_void_p]),
    ("zgl_synchronize", C.c_int32, [C.c_void_p]),
    ("zgl_upload_scene", C.c_int32, [C.c_void_p, C.POINTER(SceneBuffers)]),
    ("zgl_set_camera", C.c_int32, [C.c_void_p, C.POINTER(Camera)]),
    ("zgl_set_noise", C.c_int32, [C.c_void_p, _fp, C.c_int32, C.c_int32]),
    ("zgl_set_textures", C.c_int32, [C.c_void_p, _u8p, C.c_int32, C.c_int32, C.c_int32, _fp]),
    ("zgl_trace_closest", C.c_int32, [C.c_void_p, _fp, C.c_int64, _ip, _fp, _up]),
    ("zgl_trace_closest_device", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("zgl_trace_closest_counted", C.c_int32, [C.c_void_p, _fp, C.c_int64, _ip, _fp, _up]),
    ("zgl_trace_any", C.c_int32, [C.c_void_p, _fp, _fp, C.c_int64, _u8p]),
    ("zgl_render", C.c_int32, [C.c_void_p, C.POINTER(PassParams)]),
    ("zgl_render_static", C.c_int32, [C.c_void_p, C.POINTER(PassParams)]),
    ("zgl_render_reference_f64", C.c_int32, [C.c_void_p, C.POINTER(PassParams), C.POINTER(C.c_double)]),
    ("zgl_clear_frame", C.c_int32, [C.c_void_p]),
    ("zgl_post_process", C.c_int32, [C.c_void_p, C.c_float, C.c_int32, _fp, C.c_int64, _u8p, C.c_int64]),
    ("zgl_read_frame", C.c_int32, [C.c_void_p, _fp, C.c_int64]),
    ("zgl_read_film", C.c_int32, [C.c_void_p, _fp, C.c_int64]),
    ("zgl_write_frame", C.c_int32, [C.c_void_p, _fp, C.c_int64]),
    ("zgl_frame_device_ptr", C.c_void_p, [C.c_void_p]),
    ("zgl_read_counters", C.c_int32, [C.c_void_p, _u64p, C.c_int32]),
    ("zgl_reset_counters", C.c_int32, [C.c_void_p]),
    ("zgl_set_eval_params", C.c_int32, [C.c_void_p, _fp, _fp]),
    ("zgl_render_replay", C.c_int32, [C.c_void_p, C.POINTER(PassParams)]),
    ("zgl_adjoint", C.c_int32, [C.c_void_p, C.POINTER(PassParams), _fp, _fp, _fp]),
    ("zgl_batch_create", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("zgl_batch_destroy", None, [C.c_void_p]),
    ("zgl_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("zgl_batch_set", C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(PassParams), C.c_int32]),
    ("zgl_batch_configure", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _ip]),
    ("zgl_batch_plan", C.c_int32, [_ip, C.POINTER(PassParams), C.c_int32, C.POINTER(C.c_int64), _ip, _ip]),
    ("zgl_batch_launch", C.c_int32, [C.c_void_p]),
    ("zgl_batch_synchronize", C.c_int32, [C.c_void_p]),
    ("zgl_batch_tune", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    ("zgl_batch_get_config", C.c_int32, [C.c_void_p, _ip, _ip, _ip, _ip]),
    ("zgl_bvh_builder_create", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("zgl_bvh_builder_destroy", None, [C.c_void_p]),
    ("zgl_bvh_builder_last_error", C.c_char_p, [C.c_void_p]),
    ("zgl_bvh_builder_upload", C.c_int32, [C.c_void_p, _fp, C.c_int64, _up, C.c_int64]),
    ("zgl_bvh_builder_build", C.c_int32, [C.c_void_p, C.c_int32]),
    ("zgl_bvh_builder_download", C.c_int32, [C.c_void_p, _fp, _ip]),
    ("zgl_bvh_builder_stats", C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("zgl_build_mtbvh_device", C.c_int32, [C.c_int32, _fp, C.c_int64, _up, C.c_int64, _fp, _ip]),
]

_lib = None


def lib():
    """Load libzgl.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m zillumgl_amd.build` first")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


class ZglError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status, ctx=None, what=""):
    if status != ZGL_OK:
        msg = lib().zgl_last_error(ctx).decode() if ctx else what
        raise ZglError(status, msg or what)


# ----------------------------------------------------------------- host producers
def build_mtbvh(vertices, indices):
    """MTBVH bounds ((2n-1),6) and hit tables (6,(2n-1),3) for a triangle mesh."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
    n = i.shape[0]
    bounds = np.empty(((2 * n - 1), 6), np.float32)
    table = np.empty((6, 2 * n - 1, 3), np.int32)
    check(lib().zgl_build_mtbvh(ptr(v, C.c_float), v.shape[0], ptr(i, C.c_uint32), n,
                                ptr(bounds, C.c_float), ptr(table, C.c_int32)), what="zgl_build_mtbvh")
    return bounds, table


def build_mtbvh_device(vertices, indices, device=0):
    """The same MTBVH as build_mtbvh, built by the GPU kernels (zgl_build_mtbvh_device)."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
    n = i.shape[0]
    bounds = np.empty(((2 * n - 1), 6), np.float32)
    table = np.empty((6, 2 * n - 1, 3), np.int32)
    check(lib().zgl_build_mtbvh_device(device, ptr(v, C.c_float), v.shape[0], ptr(i, C.c_uint32), n,
                                       ptr(bounds, C.c_float), ptr(table, C.c_int32)),
          what="zgl_build_mtbvh_device")
    return bounds, table


def alias_build(pdf):
    p = np.ascontiguousarray(pdf, dtype=np.float32)
    alias = np.zeros(len(p), np.int32)
    prob = np.zeros(len(p), np.float32)
    check(lib().zgl_alias_build(ptr(p, C.c_float), len(p), ptr(alias, C.c_int32), ptr(prob, C.c_float)),
          what="zgl_alias_build")
    return alias, prob


def env_alias_build(rgb):
    img = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w, _ = img.shape
    alias = np.zeros((h, w + 1), np.int32)
    prob = np.zeros((h, w + 1), np.float32)
    s = C.c_float(0)
    check(lib().zgl_env_alias_build(ptr(img, C.c_float), w, h, ptr(alias, C.c_int32), ptr(prob, C.c_float),
                                    C.byref(s)), what="zgl_env_alias_build")
    return alias, prob, s.value


def sobol_matrices():
    out = np.zeros(256 * 32, np.uint32)
    lib().zgl_sobol_matrices(ptr(out, C.c_uint32))
    return out.reshape(256, 32)


def float_to_half(x):
    a = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(a.shape, np.uint16)
    lib().zgl_float_to_half(ptr(a, C.c_float), ptr(out, C.c_uint16), a.size)
    return out


def device_count():
    return int(lib().zgl_device_count())
