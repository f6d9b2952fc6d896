"""Device context and the progressive integrators over the C ABI.

``NaivePathIntegrator``, ``LightPathIntegrator`` and ``TriplePathIntegrator`` keep
the reference's Integrator interface (src/core/Integrator.h:25-52): init,
renderOnePass, reset, getFrame, resultScale, with the same pass counters
(mCurSample, mFreeCounter, samplePerPixel) and parameter defaults
(Integrator.h:56-143).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, ptr


class Context:
    """One zgl_context (device buffers + one HIP stream)."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = C.c_void_p()
        st = L.zgl_create(device, C.byref(h))
        if st != _lib.ZGL_OK:
            raise _lib.ZglError(st, f"zgl_create(device={device}) failed")
        self.h = h
        self.width = self.height = 0
        self._keep = None

    def close(self):
        if self.h:
            _lib.lib().zgl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        check(st, self.h)

    def upload(self, packed):
        sb = packed.as_struct()
        self._check(_lib.lib().zgl_upload_scene(self.h, C.byref(sb)))
        self.set_camera(packed.camera)
        self.set_noise(packed.noise)
        self.set_textures(getattr(packed, "tex_rgb8", None), getattr(packed, "tex_scale", None))
        return self

    def set_textures(self, rgb8, uv_scale):
        """Albedo texture array (layers, h, w, 3) uint8 sRGB and per-layer uv scales
        (layers, 2); None removes the textures."""
        if rgb8 is None:
            self._check(_lib.lib().zgl_set_textures(self.h, None, 0, 0, 0, None))
            return
        t = np.ascontiguousarray(rgb8, np.uint8)
        sc = np.ascontiguousarray(uv_scale, np.float32)
        self._check(_lib.lib().zgl_set_textures(self.h, ptr(t, C.c_uint8), t.shape[0], t.shape[2], t.shape[1],
                                                ptr(sc, C.c_float)))

    def set_camera(self, cam):
        self._check(_lib.lib().zgl_set_camera(self.h, C.byref(cam)))
        self.width, self.height = cam.film_width, cam.film_height

    def set_noise(self, noise):
        n = np.ascontiguousarray(noise, np.float32)
        self._check(_lib.lib().zgl_set_noise(self.h, ptr(n, C.c_float), n.shape[1], n.shape[0]))

    def set_stream(self, stream_handle):
        self._check(_lib.lib().zgl_set_stream(self.h, C.c_void_p(stream_handle)))

    def stream(self):
        return _lib.lib().zgl_get_stream(self.h)

    def synchronize(self):
        self._check(_lib.lib().zgl_synchronize(self.h))

    def trace_closest(self, rays, counted=False):
        r = np.ascontiguousarray(rays, np.float32)
        n = r.shape[-1] if r.ndim == 2 else r.size // 6
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        b = np.empty(n, np.uint32)
        f = _lib.lib().zgl_trace_closest_counted if counted else _lib.lib().zgl_trace_closest
        self._check(f(self.h, ptr(r, C.c_float), n, ptr(prim, C.c_int32), ptr(t, C.c_float), ptr(b, C.c_uint32)))
        return prim, t, b

    def trace_closest_device(self, d_rays, n, d_prim, d_t, d_bin):
        self._check(_lib.lib().zgl_trace_closest_device(self.h, C.c_void_p(d_rays), n, C.c_void_p(d_prim),
                                                        C.c_void_p(d_t), C.c_void_p(d_bin)))

    def trace_any(self, rays, tmax):
        r = np.ascontiguousarray(rays, np.float32)
        n = r.size // 6
        tm = np.ascontiguousarray(tmax, np.float32)
        occ = np.empty(n, np.uint8)
        self._check(_lib.lib().zgl_trace_any(self.h, ptr(r, C.c_float), ptr(tm, C.c_float), n,
                                             ptr(occ, C.c_uint8)))
        return occ.astype(bool)

    def render(self, integrator, first_pass, n_passes, free_counter, max_depth, russian_roulette=False,
               sample_light=True, light_env_uniform=False, light_portion=0.5, blocks_one_pass=32,
               loops_per_pass=1, count_bytes=False, regen_threshold=0):
        p = _lib.PassParams(integrator, first_pass, n_passes, free_counter, max_depth, int(russian_roulette),
                            int(sample_light), int(light_env_uniform), light_portion, blocks_one_pass,
                            loops_per_pass, int(count_bytes), regen_threshold)
        self._check(_lib.lib().zgl_render(self.h, C.byref(p)))

    def render_reference_f64(self, first_pass, n_passes, free_counter, max_depth, russian_roulette=False,
                             sample_light=True, light_env_uniform=False, light_portion=0.5, rgb_sum=None):
        """The naive integrator's passes in FP64 (zgl_render_reference_f64): returns
        rgb_sum (H, W, 3 float64, a new zero array if None) with each pixel's sum of the
        n_passes samples added."""
        if rgb_sum is None:
            rgb_sum = np.zeros((self.height, self.width, 3), np.float64)
        assert rgb_sum.dtype == np.float64 and rgb_sum.shape == (self.height, self.width, 3)
        assert rgb_sum.flags.c_contiguous
        p = _lib.PassParams(_lib.INTEGRATOR_NAIVE, first_pass, n_passes, free_counter, max_depth,
                            int(russian_roulette), int(sample_light), int(light_env_uniform), light_portion, 32, 1,
                            0, 0)
        self._check(_lib.lib().zgl_render_reference_f64(self.h, C.byref(p), ptr(rgb_sum, C.c_double)))
        return rgb_sum

    # ------------------------------------------------------------ differentiable mode
    def _params(self, first_pass, n_passes, free_counter, max_depth, russian_roulette, sample_light,
                light_env_uniform, light_portion):
        return _lib.PassParams(_lib.INTEGRATOR_NAIVE, first_pass, n_passes, free_counter, max_depth,
                               int(russian_roulette), int(sample_light), int(light_env_uniform), light_portion, 32, 1,
                               0, 0)

    def set_eval_params(self, material_params=None, env_rgb=None):
        """Evaluation point of the differentiable estimator: material_params (n, 5) =
        (baseColor rgb, raw roughness, metallic), env_rgb (env_h, env_w, 3); None keeps
        the uploaded values."""
        m = None if material_params is None else np.ascontiguousarray(material_params, np.float32)
        e = None if env_rgb is None else np.ascontiguousarray(env_rgb, np.float32)
        self._check(_lib.lib().zgl_set_eval_params(self.h, ptr(m, C.c_float), ptr(e, C.c_float)))

    def render_replay(self, first_pass, n_passes, free_counter, max_depth, russian_roulette=False,
                      sample_light=True, light_env_uniform=False, light_portion=0.5):
        p = self._params(first_pass, n_passes, free_counter, max_depth, russian_roulette, sample_light,
                         light_env_uniform, light_portion)
        self._check(_lib.lib().zgl_render_replay(self.h, C.byref(p)))

    def adjoint(self, d_frame, n_materials, env_shape, first_pass, n_passes, free_counter, max_depth,
                russian_roulette=False, sample_light=True, light_env_uniform=False, light_portion=0.5):
        """Gradients of sum(d_frame . frame) (d_frame (H, W, 3) or (H, W, 4)) w.r.t. the
        material parameters (n, 5) and the env texels env_shape + (3,)."""
        d = np.zeros((self.height, self.width, 4), np.float32)
        d[..., :3] = np.asarray(d_frame, np.float32)[..., :3]
        gm = np.zeros((n_materials, 5), np.float32)
        ge = np.zeros(tuple(env_shape[:2]) + (3,), np.float32)
        p = self._params(first_pass, n_passes, free_counter, max_depth, russian_roulette, sample_light,
                         light_env_uniform, light_portion)
        self._check(_lib.lib().zgl_adjoint(self.h, C.byref(p), ptr(d, C.c_float), ptr(gm, C.c_float),
                                           ptr(ge, C.c_float)))
        return gm, ge

    def clear(self):
        self._check(_lib.lib().zgl_clear_frame(self.h))

    def read_frame(self):
        out = np.empty((self.height, self.width, 4), np.float32)
        self._check(_lib.lib().zgl_read_frame(self.h, ptr(out, C.c_float), out.size))
        return out

    def post_process(self, result_scale, tone_mapper=1):
        """The display transform of post_proc.glsl: (rgba float32 (H, W, 4), rgb8 (H, W, 3)
        rows top-down, as captureImage writes them).  tone_mapper 0 none, 1 filmic
        (the reference's default), 2 ACES."""
        rgba = np.empty((self.height, self.width, 4), np.float32)
        rgb = np.empty((self.height, self.width, 3), np.uint8)
        self._check(_lib.lib().zgl_post_process(self.h, result_scale, tone_mapper, ptr(rgba, C.c_float), rgba.size,
                                                ptr(rgb, C.c_uint8), rgb.size))
        return rgba, rgb

    def read_film(self):
        out = np.empty((self.height, self.width * 3), np.float32)
        self._check(_lib.lib().zgl_read_film(self.h, ptr(out, C.c_float), out.size))
        return out.reshape(self.height, self.width, 3)

    def write_frame(self, rgba):
        a = np.ascontiguousarray(rgba, np.float32)
        self._check(_lib.lib().zgl_write_frame(self.h, ptr(a, C.c_float), a.size))

    def frame_device_ptr(self):
        return _lib.lib().zgl_frame_device_ptr(self.h)

    def counters(self):
        out = np.zeros(4, np.uint64)
        self._check(_lib.lib().zgl_read_counters(self.h, ptr(out, C.c_uint64), 4))
        return out

    def reset_counters(self):
        self._check(_lib.lib().zgl_reset_counters(self.h))


class Batch:
    """Several contexts rendered by one persistent launch captured in a HIP graph
    (zgl_batch_*).  Each frame receives what Context.render of the same passes writes."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = C.c_void_p()
        st = L.zgl_batch_create(device, C.byref(h))
        if st != _lib.ZGL_OK:
            raise _lib.ZglError(st, f"zgl_batch_create(device={device}) failed")
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            _lib.lib().zgl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != _lib.ZGL_OK:
            raise _lib.ZglError(st, _lib.lib().zgl_batch_last_error(self.h).decode())

    def set(self, contexts, passes):
        """contexts: [Context]; passes: [(first_pass, n_passes, free_counter, max_depth)] or
        dicts of Context.render keyword arguments."""
        n = len(contexts)
        hs = (C.c_void_p * n)(*[c.h.value for c in contexts])
        ps = (_lib.PassParams * n)()
        for i, p in enumerate(passes):
            if isinstance(p, dict):
                kw = dict(russian_roulette=False, sample_light=True, light_env_uniform=False, light_portion=0.5)
                kw.update(p)
                ps[i] = _lib.PassParams(_lib.INTEGRATOR_NAIVE, kw["first_pass"], kw["n_passes"], kw["free_counter"],
                                        kw["max_depth"], int(kw["russian_roulette"]), int(kw["sample_light"]),
                                        int(kw["light_env_uniform"]), kw["light_portion"], 32, 1, 0, 0)
            else:
                first, count, fc, depth = p
                ps[i] = _lib.PassParams(_lib.INTEGRATOR_NAIVE, first, count, fc, depth, 0, 1, 0, 0.5, 32, 1, 0, 0)
        self._keep = (contexts, hs, ps)
        self._check(_lib.lib().zgl_batch_set(self.h, hs, ps, n))
        self.n = n

    def configure(self, regen_threshold=0, blocks_per_cu=0, order=None):
        o = None if order is None else np.ascontiguousarray(order, np.int32)
        self._check(_lib.lib().zgl_batch_configure(self.h, regen_threshold, blocks_per_cu, ptr(o, C.c_int32)))

    def launch(self):
        self._check(_lib.lib().zgl_batch_launch(self.h))

    def synchronize(self):
        self._check(_lib.lib().zgl_batch_synchronize(self.h))

    def tune(self, reps=1):
        best = C.c_double(0)
        self._check(_lib.lib().zgl_batch_tune(self.h, reps, C.byref(best)))
        return best.value

    def config(self):
        th, bpc, mx = C.c_int32(), C.c_int32(), C.c_int32()
        order = np.zeros(self.n, np.int32)
        self._check(_lib.lib().zgl_batch_get_config(self.h, C.byref(th), C.byref(bpc), ptr(order, C.c_int32),
                                                     C.byref(mx)))
        return dict(regen_threshold=th.value, blocks_per_cu=bpc.value, order=order, max_blocks_per_cu=mx.value)


class NaivePathIntegrator:
    """Naive MIS path tracer (src/integrator/NaivePath.cpp:77-115); defaults from
    PathIntegParam (src/core/Integrator.h:56-66)."""

    def __init__(self, max_depth=4, russian_roulette=False, sample_light=True, light_env_uniform=False,
                 light_portion=0.5):
        self.max_depth = max_depth
        self.russian_roulette = russian_roulette
        self.sample_light = sample_light
        self.light_env_uniform = light_env_uniform
        self.light_portion = light_portion
        self.cur_sample = 0
        self.free_counter = 0

    def init(self, ctx):
        self.ctx = ctx
        self.reset()

    def reset(self):
        self.ctx.clear()
        self.cur_sample = 0

    def render_one_pass(self, n_passes=1):
        """renderOnePass; n_passes > 1 renders consecutive passes in one launch."""
        self.free_counter += 1
        self.ctx.render(_lib.INTEGRATOR_NAIVE, self.cur_sample, n_passes, self.free_counter, self.max_depth,
                        self.russian_roulette, self.sample_light, self.light_env_uniform, self.light_portion)
        self.free_counter += n_passes - 1
        self.cur_sample += n_passes

    def get_frame(self):
        return self.ctx.read_frame()

    def result_scale(self):
        # 1 / (mCurSample + 1) as the reference displays it (Integrator.h:79)
        return 1.0 / (self.cur_sample + 1)


class GlobalQueuePathIntegrator:
    """Global-queue (wavefront) path tracer (src/integrator/GlobalQueuePath.cpp:100-176);
    QueuedPTParam defaults (src/core/Integrator.h:211-222).  Its estimator differs from
    NaivePathIntegrator's (Sobol dims 4 + 9 depth per vertex, seeds restarted per depth,
    contributions added to the frame as produced)."""

    def __init__(self, max_depth=4, russian_roulette=False, sample_light=True, light_env_uniform=False,
                 light_portion=0.5):
        self.max_depth = max_depth
        self.russian_roulette = russian_roulette
        self.sample_light = sample_light
        self.light_env_uniform = light_env_uniform
        self.light_portion = light_portion
        self.cur_sample = 0
        self.free_counter = 0
        self.sample_per_pixel = 0.0

    def init(self, ctx):
        self.ctx = ctx
        self.reset()

    def reset(self):
        self.ctx.clear()
        self.cur_sample = 0
        self.sample_per_pixel = 1.0

    def render_one_pass(self, n_passes=1):
        """renderOnePass; n_passes > 1 renders consecutive passes in one call."""
        self.free_counter += 1
        self.ctx.render(_lib.INTEGRATOR_GLOBAL_QUEUE, self.cur_sample, n_passes, self.free_counter, self.max_depth,
                        self.russian_roulette, self.sample_light, self.light_env_uniform, self.light_portion)
        self.free_counter += n_passes - 1
        self.cur_sample += n_passes
        self.sample_per_pixel += n_passes

    def get_frame(self):
        return self.ctx.read_frame()

    def result_scale(self):
        return 1.0 / self.sample_per_pixel


class LightPathIntegrator:
    """Light tracer (src/integrator/LightPath.cpp:83-135); LightPathIntegParam defaults."""
    WORKGROUP = 1536

    def __init__(self, max_depth=4, russian_roulette=False, thread_blocks_one_pass=32):
        self.max_depth = max_depth
        self.russian_roulette = russian_roulette
        self.blocks = thread_blocks_one_pass
        self.cur_sample = 0
        self.free_counter = 0
        self.sample_per_pixel = 0.0

    def init(self, ctx):
        self.ctx = ctx
        self.reset()

    def reset(self):
        self.ctx.clear()
        self.cur_sample = 0
        self.sample_per_pixel = self.blocks * self.WORKGROUP / (self.ctx.width * self.ctx.height)

    def render_one_pass(self, n_passes=1):
        self.free_counter += 1
        self.ctx.render(_lib.INTEGRATOR_LIGHT, self.cur_sample, n_passes, self.free_counter, self.max_depth,
                        self.russian_roulette, blocks_one_pass=self.blocks)
        self.free_counter += n_passes - 1
        self.sample_per_pixel += n_passes * self.blocks * self.WORKGROUP / (self.ctx.width * self.ctx.height)
        self.cur_sample += n_passes

    def get_frame(self):
        return self.ctx.read_frame()

    def result_scale(self):
        return 1.0 / self.sample_per_pixel


class TriplePathIntegrator:
    """s=0 / s=1 / t=1 MIS (src/integrator/TriplePath.cpp:97-156); TriplePathIntegParam defaults."""

    def __init__(self, max_depth=4, russian_roulette=False, lpt_blocks_one_pass=64, lpt_loops_per_pass=1):
        self.max_depth = max_depth
        self.russian_roulette = russian_roulette
        self.blocks = lpt_blocks_one_pass
        self.loops = lpt_loops_per_pass
        self.cur_sample = 0
        self.free_counter = 0
        self.sample_per_pixel = 1.0

    def init(self, ctx):
        self.ctx = ctx
        self.reset()

    def reset(self):
        self.ctx.clear()
        self.cur_sample = 0
        self.sample_per_pixel = 1.0

    def render_one_pass(self, n_passes=1):
        self.free_counter += 1
        self.ctx.render(_lib.INTEGRATOR_TRIPLE, self.cur_sample, n_passes, self.free_counter, self.max_depth,
                        self.russian_roulette, blocks_one_pass=self.blocks, loops_per_pass=self.loops)
        self.free_counter += n_passes - 1
        self.sample_per_pixel += n_passes
        self.cur_sample += n_passes

    def get_frame(self):
        return self.ctx.read_frame()

    def result_scale(self):
        return 1.0 / self.sample_per_pixel


class BvhBuilder:
    """MTBVH construction on the GPU (zgl_bvh_builder_*): the same bounds and hit tables
    as the host build (_lib.build_mtbvh), bit for bit."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = C.c_void_p()
        st = L.zgl_bvh_builder_create(device, C.byref(h))
        if st != _lib.ZGL_OK:
            raise _lib.ZglError(st, f"zgl_bvh_builder_create(device={device}) failed")
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            _lib.lib().zgl_bvh_builder_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != _lib.ZGL_OK:
            raise _lib.ZglError(st, _lib.lib().zgl_bvh_builder_last_error(self.h).decode())

    def upload(self, vertices, indices):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self._check(_lib.lib().zgl_bvh_builder_upload(self.h, _lib.ptr(v, C.c_float), v.shape[0],
                                                      _lib.ptr(i, C.c_uint32), i.shape[0]))
        self.n = i.shape[0]

    def build(self, subtree_size=0):
        """Builds on the device; returns when done."""
        self._check(_lib.lib().zgl_bvh_builder_build(self.h, subtree_size))

    def download(self):
        n = self.n
        bounds = np.empty(((2 * n - 1), 6), np.float32)
        table = np.empty((6, 2 * n - 1, 3), np.int32)
        self._check(_lib.lib().zgl_bvh_builder_download(self.h, _lib.ptr(bounds, C.c_float),
                                                        _lib.ptr(table, C.c_int32)))
        return bounds, table

    def stats(self):
        out = (C.c_int64 * 5)()
        self._check(_lib.lib().zgl_bvh_builder_stats(self.h, out))
        return dict(zip(("levels", "big_nodes", "small_subtrees", "depth", "jumps"), list(out)))
