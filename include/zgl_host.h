/* zgl_host.h -- C ABI of the host layer (libzgl_host.so, above libzgl.so).
 *
 * The C++ classes of zillumgl_amd/host/zgl_host.hpp (scene ingestion, packing and the
 * Integrator classes) for callers that bind C: the Python tests, or a host in another
 * language.  Every function returns a zgl_status (ZGL_OK = 0) and never aborts; the
 * message of the calling thread's last failure is zglh_last_error().
 */
#ifndef ZGL_HOST_H
#define ZGL_HOST_H

#include "zgl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zglh_scene zglh_scene;
typedef struct zglh_integrator zglh_integrator;

/* The calling thread's last failure message (the reference aborts instead,
 * src/util/Error.h:24-42). */
const char* zglh_last_error(void);

/* Scene::load (src/core/Scene.cpp:58-127) of an XML scene with OBJ/MTL models and a .hdr
 * or .pfm environment map, packed as Scene::createGLContext (Scene.cpp:133-270) does.
 * device_bvh != 0 leaves the MTBVH to zgl_upload_scene's device build. */
zgl_status zglh_load_scene(const char* xml_path, int32_t device_bvh, zglh_scene** out);
void zglh_scene_destroy(zglh_scene* scene);
/* The packed buffers (owned by the scene): SceneGLContext (src/core/Scene.h:12-27),
 * the camera uniforms NaivePathIntegrator::updateUniforms sets (NaivePath.cpp:49-60) and
 * the seed texture (Sampler.cpp:66-80). */
zgl_status zglh_scene_buffers(const zglh_scene* scene, zgl_scene_buffers* buffers, zgl_camera* camera,
                              const float** noise_rg, int32_t* width, int32_t* height);
/* zgl_upload_scene + zgl_set_camera + zgl_set_noise: what Scene::createGLContext and
 * updateUniforms hand to GL (Scene.cpp:245-258, NaivePath.cpp:12-67). */
zgl_status zglh_upload(zgl_context* ctx, const zglh_scene* scene);

/* Integrator of type ZGL_INTEGRATOR_NAIVE / _LIGHT / _TRIPLE / _GLOBAL_QUEUE
 * (src/core/Integrator.h:25-52; src/integrator/NaivePath.cpp, LightPath.cpp, TriplePath.cpp,
 * GlobalQueuePath.cpp) with the reference's parameter defaults apart from these two. */
zgl_status zglh_integrator_create(int32_t type, int32_t max_depth, int32_t russian_roulette, zglh_integrator** out);
void zglh_integrator_destroy(zglh_integrator* it);
/* init(scene, w, h): the context holds the uploaded scene; the frame is cleared. */
zgl_status zglh_integrator_init(zglh_integrator* it, zgl_context* ctx, int32_t width, int32_t height);
/* renderOnePass; n > 1 renders n consecutive passes in one call. */
zgl_status zglh_integrator_render_one_pass(zglh_integrator* it, int32_t n);
zgl_status zglh_integrator_reset(zglh_integrator* it);
/* getFrame (the running sum, RGBA32F W x H) and resultScale (Integrator.h:79). */
zgl_status zglh_integrator_get_frame(const zglh_integrator* it, float* rgba, int64_t n_floats);
zgl_status zglh_integrator_result_scale(const zglh_integrator* it, float* scale, int32_t* cur_sample);

#ifdef __cplusplus
}
#endif

#endif /* ZGL_HOST_H */
