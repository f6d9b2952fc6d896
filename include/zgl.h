/* zgl.h -- C ABI of the MI355X rendering library (libzgl.so).
 *
 * Drop-in replacement for the reference's compute-shader dispatch
 * (the src/shader GLSL kernels driven by src/integrator/{NaivePath,LightPath,TriplePath}.cpp).
 * Plain pointers and sizes only; the caller owns every host pointer and the
 * library copies what it needs before returning.  No entry point aborts: each
 * returns a zgl_status and, for context calls, leaves a message readable with
 * zgl_last_error().  A context is not thread-safe (like the reference's single GL
 * context); distinct contexts may be used from distinct threads.
 */
#ifndef ZGL_H
#define ZGL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t zgl_status;
#define ZGL_OK 0
#define ZGL_E_INVALID 1  /* bad argument or inconsistent buffers */
#define ZGL_E_HIP 2      /* HIP runtime error */
#define ZGL_E_OOM 3      /* device or host allocation failed */
#define ZGL_E_STATE 4    /* call order (e.g. render before upload) */
#define ZGL_E_NODEVICE 5 /* no HIP device available */

#define ZGL_INTEGRATOR_NAIVE 0  /* NaivePathIntegrator  (src/integrator/NaivePath.cpp) */
#define ZGL_INTEGRATOR_LIGHT 1  /* LightPathIntegrator  (src/integrator/LightPath.cpp) */
#define ZGL_INTEGRATOR_TRIPLE 2 /* TriplePathIntegrator (src/integrator/TriplePath.cpp) */
#define ZGL_INTEGRATOR_GLOBAL_QUEUE 3 /* GlobalQueuePathIntegrator (src/integrator/GlobalQueuePath.cpp) */
#define ZGL_INTEGRATOR_NAIVE_WAVEFRONT 4 /* NaivePathIntegrator's estimator and frame, wavefront schedule */

/* ------------------------------------------------------------------ */
/* Host producers (no GPU needed)                                       */
/* ------------------------------------------------------------------ */

/* BVH::build (src/accelerator/BVH.cpp:116-144): 16-bucket binned SAH, one triangle
 * per leaf, threaded six ways (buildHitTable, :298-345).
 * bounds_out: (2n-1)*6 floats; hit_table_out: 6*(2n-1)*3 ints. */
zgl_status zgl_build_mtbvh(const float* vertices, int64_t n_vertices, const uint32_t* indices,
                           int64_t n_triangles, float* bounds_out, int32_t* hit_table_out);

/* AliasTable::build<int32_t> (src/math/AliasTable.h:11-57). */
zgl_status zgl_alias_build(const float* pdf, int32_t n, int32_t* alias_out, float* prob_out);

/* EnvironmentMap::EnvironmentMap + setupAliasTable (src/core/EnvironmentMap.cpp:8-114):
 * alias/prob are (width+1)*height; *sum_out is the float sum whose int truncation is
 * the reference's sumPdf(). */
zgl_status zgl_env_alias_build(const float* rgb, int32_t width, int32_t height, int32_t* alias_out,
                               float* prob_out, float* sum_out);

/* Sampler::sobolSample (src/core/Sampler.cpp:19-28). */
uint32_t zgl_sobol_sample(uint32_t index, int32_t dim, uint32_t scramble);
/* SobolMatrices (src/math/SobolMatrices256x32.h): 256*32 generator columns. */
void zgl_sobol_matrices(uint32_t* out);

/* FP32 -> FP16 (round to nearest even) as done by the RGB16F upload of the env map
 * (src/core/EnvironmentMap.cpp:13). */
void zgl_float_to_half(const float* in, uint16_t* out, int64_t n);

/* ------------------------------------------------------------------ */
/* Device context                                                       */
/* ------------------------------------------------------------------ */
typedef struct zgl_context zgl_context;

int32_t zgl_device_count(void);
zgl_status zgl_create(int32_t device, zgl_context** out);
void zgl_destroy(zgl_context* ctx);
const char* zgl_last_error(const zgl_context* ctx);
/* Use a caller-owned hipStream_t (NULL = the context's own stream). */
zgl_status zgl_set_stream(zgl_context* ctx, void* hip_stream);
void* zgl_get_stream(zgl_context* ctx);
zgl_status zgl_synchronize(zgl_context* ctx);

/* Scene buffers in the reference's layout: SceneGLContext (src/core/Scene.h:12-27)
 * plus the scalars NaivePathIntegrator::updateUniforms binds
 * (src/integrator/NaivePath.cpp:39-47). */
typedef struct zgl_scene_buffers {
    const float* vertices;        /* RGB32F, n_vertices */
    const float* normals;         /* RGB32F, n_vertices */
    int64_t n_vertices;
    const float* texcoords;       /* RG32F, object vertices only */
    int64_t n_texcoords;
    const int32_t* indices;       /* 3 per triangle: objects first, then lights */
    int64_t n_triangles;
    const float* bounds;          /* bvh_size nodes x {pMin.xyz, pMax.xyz} */
    const int32_t* hit_table;     /* 6 x bvh_size x {node, prim, miss}; bounds and
                                     hit_table both NULL: the MTBVH is built on the
                                     device (as zgl_build_mtbvh_device) */
    int32_t bvh_size;             /* boxCount = 2n - 1 */
    const int32_t* mat_tex_indices; /* one per object triangle */
    int32_t obj_prim_count;       /* objPrimCount */
    const float* materials;       /* 16 floats per material; [13] holds the int type */
    int32_t n_materials;
    const float* light_power;     /* RGB32F per light triangle */
    const int32_t* light_alias;
    const float* light_prob;
    int32_t n_light_triangles;
    float light_sum;              /* lightSumPdf */
    const uint16_t* env_rgb16f;   /* RGB16F, env_width x env_height, row 0 = v 0 */
    int32_t env_width;
    int32_t env_height;
    const int32_t* env_alias;     /* (env_width+1) x env_height */
    const float* env_prob;        /* (env_width+1) x env_height */
    float env_sum;                /* (float)(int)sum, as sumPdf() returns int */
    float env_rotation;           /* uEnvRotation */
    int32_t sampler;              /* 0 independent, 1 Sobol (scene->sampler) */
} zgl_scene_buffers;

zgl_status zgl_upload_scene(zgl_context* ctx, const zgl_scene_buffers* scene);

/* Camera uniforms (src/integrator/NaivePath.cpp:49-60). mat_inv is column-major. */
typedef struct zgl_camera {
    float front[3], right[3], up[3], pos[3];
    float mat_inv[9];
    float tan_fov;                /* tan(radians(fov * 0.5)) */
    float aspect;
    float lens_radius;
    float focal_dist;
    int32_t film_width, film_height;
} zgl_camera;

/* Sets the camera; (re)allocates and clears the frame when the film size changes. */
zgl_status zgl_set_camera(zgl_context* ctx, const zgl_camera* camera);
/* Per-pixel seed texture (uNoiseTex, RG32F width x height, src/core/Sampler.cpp:66-80). */
zgl_status zgl_set_noise(zgl_context* ctx, const float* noise_rg, int32_t width, int32_t height);
/* Albedo textures (uTextures + uTexUVScale, src/core/Scene.cpp:256-258; Texture2DArray,
 * src/core/Texture.cpp:134-171): n_layers sRGB8 RGB images of width x height texels (the
 * array's size: every image's texels from (0, 0), the rest of its layer as the caller
 * fills it; row 0 = t 0, i.e. the image's first row as stored), and per layer the
 * (image size / array size) uv scale (2 floats).  A triangle whose matTexIndex holds
 * texIndex k >= 0 in its upper 16 bits takes baseColor from layer k, read as
 * texture2DArray(vec3(fract(uv) * scale_k, k)) at level 0: sRGB-decoded texels (FP64
 * formula rounded to FP32), GL_LINEAR, GL_REPEAT over the array's size
 * (material_loader.glsl:8-97).  n_layers = 0 removes the textures.  Rendering a scene that
 * names a texture not set returns ZGL_E_STATE; the differentiable mode (zgl_render_replay,
 * zgl_adjoint) rejects scenes with textured triangles (ZGL_E_INVALID). */
zgl_status zgl_set_textures(zgl_context* ctx, const uint8_t* rgb8, int32_t n_layers, int32_t width, int32_t height,
                            const float* uv_scale);

/* Closest hit (bvhHit, src/shader/intersection.glsl:395-427) for rays in SoA layout
 * ox[n] oy[n] oz[n] dx[n] dy[n] dz[n].  prim = -1 and t = 1e8 on a miss; bary_bin =
 * 16*min(15,int(16u)) + min(15,int(16v)) of the hit's barycentrics (0xffffffff on miss).
 * Host pointers; synchronous. */
zgl_status zgl_trace_closest(zgl_context* ctx, const float* rays, int64_t n, int32_t* prim, float* t,
                             uint32_t* bary_bin);
/* Same on device pointers, enqueued on the context stream (asynchronous). */
zgl_status zgl_trace_closest_device(zgl_context* ctx, const float* d_rays, int64_t n, int32_t* d_prim,
                                    float* d_t, uint32_t* d_bary_bin);
/* zgl_trace_closest through the counting kernel variant: same outputs, and adds the
 * algorithmic bytes (36 per row visited + 48 per leaf test) to the counters. */
zgl_status zgl_trace_closest_counted(zgl_context* ctx, const float* rays, int64_t n, int32_t* prim, float* t,
                                     uint32_t* bary_bin);
/* Any hit (bvhTest, intersection.glsl:367-393) with per-ray t limits; host pointers. */
zgl_status zgl_trace_any(zgl_context* ctx, const float* rays, const float* tmax, int64_t n, uint8_t* occluded);

/* One or more progressive passes of an integrator (Integrator::renderOnePass). */
typedef struct zgl_pass_params {
    int32_t integrator;           /* ZGL_INTEGRATOR_* */
    int32_t first_pass;           /* uSpp of the first pass (mCurSample) */
    int32_t n_passes;             /* consecutive passes in this call */
    int32_t free_counter;         /* uFreeCounter of the first pass (mFreeCounter) */
    int32_t max_depth;            /* uMaxDepth */
    int32_t russian_roulette;     /* uRussianRoulette */
    int32_t sample_light;         /* uSampleLight (naive only) */
    int32_t light_env_uniform;    /* uLightEnvUniformSample */
    float light_portion;          /* uLightSamplePortion */
    int32_t blocks_one_pass;      /* light: threadBlocksOnePass; triple: LPTBlocksOnePass */
    int32_t loops_per_pass;       /* triple: LPTLoopsPerPass */
    int32_t count_bytes;          /* accumulate algorithmic-byte counters (slower) */
    int32_t regen_threshold;      /* naive, light, triple (persistent kernels): lanes of a wave64
                                     still walking the MTBVH when the wave stops to shade and
                                     reissue rays (1..63).  0 = default: from the triangle count
                                     n, round(1.4 log2 n - 2) clamped to 4..28 for the naive
                                     tracer, round(1.3 log2 n - 7) clamped to 1..24 for the
                                     triple tracer's camera pass, 4 for light paths.  Frames do
                                     not depend on it. */
} zgl_pass_params;
/* ZGL_INTEGRATOR_NAIVE keeps each sample's value (16 bytes) in a context-owned buffer of
 * at most 4 GiB and adds them to the frame in pass order; calls with more passes than fit
 * run in chunks of passes.
 * ZGL_INTEGRATOR_GLOBAL_QUEUE reads first_pass, n_passes, free_counter, max_depth,
 * russian_roulette, sample_light, light_env_uniform and light_portion; every contribution
 * is added to the frame as produced (accumulateImage), so the frame must be cleared
 * (zgl_clear_frame) on reset, as the reference's image clear does.  count_bytes is not
 * supported for it. */

zgl_status zgl_render(zgl_context* ctx, const zgl_pass_params* params);
/* zgl_render for ZGL_INTEGRATOR_NAIVE that allocates no device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The unit counters, Sobol points and
 * per-sample values live in a static 128 MiB device array that each GPU holds until the
 * process ends; passes run in chunks that fit it (at most 256 per chunk).  The frame is
 * bit-identical to zgl_render's.  Rejects (ZGL_E_INVALID) a film whose one pass of
 * sample values (W x H x 16 bytes) exceeds the 133,954,304 bytes left for them
 * (8,372,144 pixels: 3840 x 2160 fits), and any other integrator.  Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (contexts on different streams) race; calls on one stream are ordered. */
zgl_status zgl_render_static(zgl_context* ctx, const zgl_pass_params* params);
/* The converged reference of the naive integrator (path_integ_naive.glsl:35-174): the
 * same estimator with every quantity in FP64 except the FP32 scene data, literals and
 * random numbers, one thread per pixel.  Adds each pixel's sum of the n_passes samples
 * (NaN samples skipped) to rgb_sum, a host array of film_width x film_height x 3 doubles
 * (RGB, row-major), which it reads and writes; returns when it is written.  For
 * measuring the RMSE of the FP32 renderer (whose frames it does not touch); it runs at
 * a fraction of zgl_render's rate.  Only ZGL_INTEGRATOR_NAIVE, without count_bytes. */
zgl_status zgl_render_reference_f64(zgl_context* ctx, const zgl_pass_params* params, double* rgb_sum);
/* Zero the frame (RGBA32F) and the light-splat film (planar R32F, 3W x H). */
zgl_status zgl_clear_frame(zgl_context* ctx);
/* Display transform of the frame (post_proc.glsl:1-59, called by Application::integrate,
 * src/Application.cpp:645-660, without the GUI's preview mode): sum * result_scale
 * clamped to [0, 1e30], tone mapped (0 none, 1 filmic, 2 ACES), raised to 1/2.2.
 * rgba_out (W*H*4 floats, alpha 1) is uOut; rgb8_out (W*H*3 bytes, rows top-down) is
 * what captureImage (Application.cpp:370-378) writes to PNG: each channel clamped to
 * [0, 1] and rounded to round(255 x).  Either output may be NULL. */
zgl_status zgl_post_process(zgl_context* ctx, float result_scale, int32_t tone_mapper, float* rgba_out,
                            int64_t n_rgba, uint8_t* rgb8_out, int64_t n_rgb8);
/* Copy the RGBA32F frame (W*H*4 floats; the running sum of the naive integrator,
 * the film copy for light/triple). Synchronous. */
zgl_status zgl_read_frame(zgl_context* ctx, float* rgba, int64_t n_floats);
/* Copy the planar film (3W x H floats). Synchronous. */
zgl_status zgl_read_film(zgl_context* ctx, float* film, int64_t n_floats);
/* Restore a frame (checkpoint resume). */
zgl_status zgl_write_frame(zgl_context* ctx, const float* rgba, int64_t n_floats);
/* Device pointer of the RGBA32F frame (valid until the film size changes). */
void* zgl_frame_device_ptr(zgl_context* ctx);
/* Counters: [0] algorithmic bytes, [1] MTBVH rows visited, [2] leaf triangle tests,
 * [3] traversals. */
zgl_status zgl_read_counters(zgl_context* ctx, uint64_t* out, int32_t n);
zgl_status zgl_reset_counters(zgl_context* ctx);

/* ---- Differentiable naive path tracer.  No counterpart in the reference (it has no
 * differentiable mode); the estimator is path_integ_naive.glsl:35-174's.
 *
 * Parameters: per material, (baseColor r, g, b, roughness, metallic) as stored in the
 * material record (slots 0, 1, 2, 3, 5; roughness before the loader's remap), used by
 * principled and Lambertian materials; per environment texel, RGB in FP32.
 * Sampling -- directions, lobe, light and texel choices, pdfs, MIS weights, Russian
 * roulette -- stays at the uploaded scene.  BSDF values and environment radiance are
 * taken at the evaluation point set here (NULL restores the uploaded values; at upload
 * the evaluation point is the uploaded scene).  Values for materials of other types
 * are ignored. */
zgl_status zgl_set_eval_params(zgl_context* ctx, const float* material_params /* n_materials x 5 */,
                               const float* env_rgb /* env_height x env_width x 3 */);
/* The estimator at the evaluation point, added into the frame as zgl_render adds
 * (naive integrator only).  At the uploaded parameters it is zgl_render's, bit for bit.
 * Works in chunks of passes like zgl_adjoint: the forward tracer logs each sample's
 * traversal results, and each sample is evaluated from the log.  The context keeps a log
 * that holds all the passes of a call: a zgl_render_replay or zgl_adjoint of the same
 * passes (first_pass, n_passes, free_counter, max_depth and the sampling options) made
 * before the scene, camera, noise or textures change skips the forward tracer. */
zgl_status zgl_render_replay(zgl_context* ctx, const zgl_pass_params* params);
/* Path-replay adjoint: gradients of sum(d_frame . frame sum of these passes) with
 * respect to the parameters at the evaluation point.  d_frame: W*H*4 floats (RGBA
 * layout of zgl_read_frame, alpha ignored).  Outputs are overwritten: d_material_params
 * n_materials x 5, d_env_rgb env_height x env_width x 3.  Does not change the frame.
 * Works in chunks of passes: the forward tracer logs each sample's traversal results
 * (up to 16 queries, 128 bytes per sample, with the sample values at most 8 GiB per
 * chunk) and the replay differentiates from the log. */
/* The traversal log is kept and reused as zgl_render_replay's is. */
zgl_status zgl_adjoint(zgl_context* ctx, const zgl_pass_params* params, const float* d_frame,
                       float* d_material_params, float* d_env_rgb);

/* ---- Batched persistent launch.  No counterpart in the reference (it renders one
 * scene per dispatch).  A batch renders the passes of several contexts (each with its
 * scene, camera, noise and frame uploaded) on one device in ONE persistent launch of the
 * naive integrator, recorded once as a HIP graph and replayed by zgl_batch_launch.  Each
 * frame receives exactly what zgl_render of the same params would write.  Waves take the
 * scenes in the configured order, one scene at a time per wave, pixels per wave from
 * that scene's counter.  The batch runs on its own stream, ordered after the work already
 * enqueued on the contexts' streams when zgl_batch_launch is called (e.g. zgl_clear_frame)
 * and before what they enqueue afterwards (zgl_read_frame waits for the batch).  The batch
 * copies each context's scene, camera, noise and frame pointer at zgl_batch_set: after
 * zgl_upload_scene, zgl_set_camera or zgl_set_noise on a context of the batch, or its
 * zgl_destroy, zgl_batch_launch returns ZGL_E_STATE until zgl_batch_set is called again. */
typedef struct zgl_batch zgl_batch;
zgl_status zgl_batch_create(int32_t device, zgl_batch** out);
void zgl_batch_destroy(zgl_batch* batch);
const char* zgl_batch_last_error(const zgl_batch* batch);
/* The scenes (contexts) and their passes (naive integrator); order = as given.  Allocates
 * the batch's sample values, sum of n_passes x width x height x 16 bytes, and a Sobol
 * table of the union of the pass ranges (at most 65,537 passes of 1 KiB; points of passes
 * past it are computed from the generator matrices).  Rejected (ZGL_E_INVALID): sample
 * values above 64 GiB, and a scene whose work units, ceil(W/8) ceil(H/8) 64 n_passes,
 * exceed 2^31 - 2^24. */
zgl_status zgl_batch_set(zgl_batch* batch, zgl_context* const* contexts, const zgl_pass_params* params, int32_t n);
/* The checks and sizes of zgl_batch_set without a device: film_wh holds n (width, height)
 * pairs.  Returns the sample values' bytes and the Sobol table's first pass and pass count
 * (any output may be NULL). */
zgl_status zgl_batch_plan(const int32_t* film_wh, const zgl_pass_params* params, int32_t n, int64_t* sample_bytes,
                          int32_t* sobol_first, int32_t* sobol_count);
/* regen_threshold 1..63 for every scene, or 0 (the default): each scene's own zgl_render
 * default for the naive tracer; blocks_per_cu 1..max (0 = max resident);
 * order: a permutation of 0..n-1 giving the launch order (NULL keeps the current one). */
zgl_status zgl_batch_configure(zgl_batch* batch, int32_t regen_threshold, int32_t blocks_per_cu,
                               const int32_t* order);
zgl_status zgl_batch_launch(zgl_batch* batch);
zgl_status zgl_batch_synchronize(zgl_batch* batch);
/* Autotuner: times the batch (median of `reps` launches after a warm-up) over scene
 * orders (as given / most work first), regeneration thresholds {per scene (0), 8, ..., 32}
 * and the two
 * highest blocks-per-CU counts, coordinate by coordinate, and keeps the fastest.  It
 * launches the batch repeatedly: frames of passes that do not start at pass 0
 * accumulate every launch. */
zgl_status zgl_batch_tune(zgl_batch* batch, int32_t reps, double* best_ms);
zgl_status zgl_batch_get_config(const zgl_batch* batch, int32_t* regen_threshold, int32_t* blocks_per_cu,
                                int32_t* order, int32_t max_blocks_per_cu[1]);

/* ------------------------------------------------------------------ */
/* MTBVH construction on the device                                     */
/* ------------------------------------------------------------------ */
/* The same output as zgl_build_mtbvh, bit for bit -- BVH::build + buildHitTable
 * (src/accelerator/BVH.cpp:116-345), the reference's CPU build, whose GPU version its
 * README lists as future work (README.md:24) -- built by HIP kernels: node bounds
 * ((2n-1)*6 floats) and six hit tables (6*(2n-1)*3 ints).  Vertices must be finite.
 * Scenes whose triangle centroids reach 1e8 in magnitude (the reference's empty-box
 * value) can make a split leave a child empty; both builds reject them with
 * ZGL_E_INVALID.  A builder keeps its device buffers between builds and is not
 * thread-safe. */
typedef struct zgl_bvh_builder zgl_bvh_builder;
zgl_status zgl_bvh_builder_create(int32_t device, zgl_bvh_builder** out);
void zgl_bvh_builder_destroy(zgl_bvh_builder* builder);
const char* zgl_bvh_builder_last_error(const zgl_bvh_builder* builder);
/* Copy a mesh (n_vertices x 3 floats, n_triangles x 3 indices, 1 <= n_triangles <= 2^27)
 * to the device. */
zgl_status zgl_bvh_builder_upload(zgl_bvh_builder* builder, const float* vertices, int64_t n_vertices,
                                  const uint32_t* indices, int64_t n_triangles);
/* Build from the uploaded mesh; returns when the build is complete.  subtree_size: nodes
 * with fewer triangles are finished by one wave each (0 = default 256, minimum 3); the
 * output does not depend on it. */
zgl_status zgl_bvh_builder_build(zgl_bvh_builder* builder, int32_t subtree_size);
/* Copy the last build's bounds and hit tables (either may be NULL). */
zgl_status zgl_bvh_builder_download(zgl_bvh_builder* builder, float* bounds_out, int32_t* hit_table_out);
/* Statistics of the last build: [0] levels of the multi-workgroup phase, [1] nodes split
 * there, [2] subtrees finished by one wave, [3] tree depth, [4] pointer-doubling rounds. */
zgl_status zgl_bvh_builder_stats(const zgl_bvh_builder* builder, int64_t stats_out[5]);
/* One call: upload, build and download on `device`. */
zgl_status zgl_build_mtbvh_device(int32_t device, const float* vertices, int64_t n_vertices,
                                  const uint32_t* indices, int64_t n_triangles, float* bounds_out,
                                  int32_t* hit_table_out);

#ifdef __cplusplus
}
#endif

#endif /* ZGL_H */
