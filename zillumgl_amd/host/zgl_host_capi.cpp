// extern "C" surface of the host layer (include/zgl_host.h): exceptions become status
// codes and a per-thread message.
#include "../../include/zgl_host.h"

#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "zgl_host.hpp"

struct zglh_scene {
    zgl::host::PackedScene ps;
};
struct zglh_integrator {
    std::unique_ptr<zgl::host::Integrator> it;
    bool ready = false;
};

namespace {
thread_local std::string gError;

template <class F>
zgl_status guarded(F&& f)
{
    try {
        f();
        gError.clear();
        return ZGL_OK;
    } catch (const std::invalid_argument& e) {
        gError = e.what();
        return ZGL_E_INVALID;
    } catch (const std::bad_alloc&) {
        gError = "host allocation failed";
        return ZGL_E_OOM;
    } catch (const std::exception& e) {
        gError = e.what();
        return ZGL_E_STATE;
    } catch (...) {
        gError = "unknown error";
        return ZGL_E_STATE;
    }
}

void need(bool ok, const char* what)
{
    if (!ok) throw std::invalid_argument(what);
}
}  // namespace

extern "C" {

const char* zglh_last_error(void) { return gError.c_str(); }

zgl_status zglh_load_scene(const char* xml_path, int32_t device_bvh, zglh_scene** out)
{
    return guarded([&] {
        need(xml_path && out, "null argument");
        *out = nullptr;
        auto s = std::make_unique<zglh_scene>();
        s->ps = zgl::host::pack(zgl::host::loadScene(xml_path), device_bvh != 0);
        *out = s.release();
    });
}

void zglh_scene_destroy(zglh_scene* scene) { delete scene; }

zgl_status zglh_scene_buffers(const zglh_scene* scene, zgl_scene_buffers* buffers, zgl_camera* camera,
                              const float** noise_rg, int32_t* width, int32_t* height)
{
    return guarded([&] {
        need(scene != nullptr, "null scene");
        if (buffers) *buffers = scene->ps.buffers();
        if (camera) *camera = scene->ps.camera;
        if (noise_rg) *noise_rg = scene->ps.noise.data();
        if (width) *width = scene->ps.width;
        if (height) *height = scene->ps.height;
    });
}

zgl_status zglh_upload(zgl_context* ctx, const zglh_scene* scene)
{
    return guarded([&] {
        need(ctx && scene, "null argument");
        zgl::host::upload(ctx, scene->ps);
    });
}

zgl_status zglh_integrator_create(int32_t type, int32_t max_depth, int32_t russian_roulette, zglh_integrator** out)
{
    return guarded([&] {
        need(out != nullptr, "null argument");
        *out = nullptr;
        need(max_depth >= 0 && max_depth <= 1024, "max_depth out of range");
        auto h = std::make_unique<zglh_integrator>();
        const bool rr = russian_roulette != 0;
        switch (type) {
        case ZGL_INTEGRATOR_NAIVE: h->it = std::make_unique<zgl::host::NaivePathIntegrator>(max_depth, rr); break;
        case ZGL_INTEGRATOR_LIGHT: h->it = std::make_unique<zgl::host::LightPathIntegrator>(max_depth, rr); break;
        case ZGL_INTEGRATOR_TRIPLE: h->it = std::make_unique<zgl::host::TriplePathIntegrator>(max_depth, rr); break;
        case ZGL_INTEGRATOR_GLOBAL_QUEUE:
            h->it = std::make_unique<zgl::host::GlobalQueuePathIntegrator>(max_depth, rr);
            break;
        default: throw std::invalid_argument("unknown integrator type");
        }
        *out = h.release();
    });
}

void zglh_integrator_destroy(zglh_integrator* it) { delete it; }

zgl_status zglh_integrator_init(zglh_integrator* it, zgl_context* ctx, int32_t width, int32_t height)
{
    return guarded([&] {
        need(it && ctx, "null argument");
        it->it->init(ctx, width, height);
        it->ready = true;
    });
}

zgl_status zglh_integrator_render_one_pass(zglh_integrator* it, int32_t n)
{
    return guarded([&] {
        need(it != nullptr, "null integrator");
        if (!it->ready) throw std::runtime_error("integrator: init() first");
        it->it->renderOnePass(n);
    });
}

zgl_status zglh_integrator_reset(zglh_integrator* it)
{
    return guarded([&] {
        need(it != nullptr, "null integrator");
        if (!it->ready) throw std::runtime_error("integrator: init() first");
        it->it->reset();
    });
}

zgl_status zglh_integrator_get_frame(const zglh_integrator* it, float* rgba, int64_t n_floats)
{
    return guarded([&] {
        need(it && rgba, "null argument");
        if (!it->ready) throw std::runtime_error("integrator: init() first");
        const std::vector<float> f = it->it->getFrame();
        need(n_floats >= (int64_t)f.size(), "frame buffer too small");
        std::copy(f.begin(), f.end(), rgba);
    });
}

zgl_status zglh_integrator_result_scale(const zglh_integrator* it, float* scale, int32_t* cur_sample)
{
    return guarded([&] {
        need(it != nullptr, "null integrator");
        if (scale) *scale = it->it->resultScale();
        if (cur_sample) *cur_sample = it->it->curSample();
    });
}

}  // extern "C"
