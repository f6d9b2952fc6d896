// zgl_capi.hip -- the C ABI: context, validated scene upload, pass dispatch.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/zgl.h"
#include "zgl_build.h"
#include "zgl_bvh.h"
#include "zgl_internal.h"

using namespace zgl;

// zgl_render's default for it (regen_threshold 0).  Frames do not depend on it; the
// fastest value grows with the length of the walks, which grows with log2 of the
// triangle count.  Fitted to sweeps on one MI355X over scenes of 36 to 1.3M triangles:
// the naive tracer was fastest at 4 with 36 triangles, 12 with 852, 16-20 with 10k-100k
// and 24-28 with 0.8M-1.3M; the triple tracer's camera pass at 1, 8-12 and 16 with 36,
// 10k and 100k; the light paths at 1-4 at every size.
enum class WalkKind { Naive, TripleCamera, LightPaths };
static int defaultThreshold(WalkKind kind, int bvhSize)
{
    const double l = std::log2((double)std::max(2, (bvhSize + 1) / 2));
    switch (kind) {
    case WalkKind::LightPaths: return 4;
    case WalkKind::TripleCamera: return std::min(24, std::max(1, (int)std::lround(1.3 * l - 7.0)));
    default: return std::min(28, std::max(4, (int)std::lround(1.4 * l - 2.0)));
    }
}
// Largest sample-value plus traversal-log buffer of one chunk of the differentiable mode's passes.
constexpr size_t kAdjointLogBytes = (size_t)8 << 30;
// Largest total of a batch's per-sample value buffers.
constexpr size_t kBatchSampleBytes = (size_t)64 << 30;
// Largest work-unit count of one scene of a batch ((8x8 tiles) x 64 x passes): the
// persistent kernel counts units in int32, and each of its waves may take up to 64
// indices past the end before it stops.
constexpr int64_t kBatchMaxUnits = (int64_t)INT_MAX - (1 << 24);
// Rows of a batch's precomputed Sobol table (1 KiB each); passes outside it compute
// their points from the generator matrices.
constexpr int kBatchSobolRows = 65537;

// Scratch of zgl_render_static, one array per GPU, held until the process ends:
// [0, 256) the persistent schedule's unit counters, then the Sobol points of one chunk
// of at most kStaticChunkPasses passes (+1 row for dims past 256), then the chunk's
// per-sample values (16 B per pixel and pass).
constexpr size_t kStaticScratchBytes = (size_t)128 << 20;
constexpr int kStaticChunkPasses = 256;
constexpr size_t kStaticSobolOffset = 256;
constexpr size_t kStaticResOffset =
    (kStaticSobolOffset + (size_t)(kStaticChunkPasses + 1) * 256 * 4 + 255) & ~(size_t)255;
constexpr size_t kStaticResBytes = kStaticScratchBytes - kStaticResOffset;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

void hipCheck(hipError_t e, const char* what);

// Grows b to at least `bytes` (contents not kept) and copies src into it on stream st.
void bufUpload(DevBuf& b, const void* src, size_t bytes, hipStream_t st, const char* what)
{
    if (b.bytes < bytes || bytes == 0) {
        b.release();
        size_t alloc = bytes ? bytes : 16;
        hipCheck(hipMalloc(&b.p, alloc), what);
        b.bytes = alloc;
    }
    if (bytes) hipCheck(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st), what);
}
void bufEnsure(DevBuf& b, size_t bytes, const char* what)
{
    if (b.bytes >= bytes && b.p) return;
    b.release();
    hipCheck(hipMalloc(&b.p, bytes ? bytes : 16), what);
    b.bytes = bytes ? bytes : 16;
}

struct ZglError {
    zgl_status code;
    std::string msg;
};

void hipCheck(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory) throw ZglError{ZGL_E_OOM, std::string(what) + ": out of device memory"};
    throw ZglError{ZGL_E_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

void invalid(const std::string& m) { throw ZglError{ZGL_E_INVALID, m}; }

// Generation stamps: every call that replaces a context's device buffers or the scene,
// camera or noise a batch copies takes a new stamp, and a batch launch checks them.
std::atomic<uint64_t> gGeneration{0};
uint64_t nextGeneration() { return gGeneration.fetch_add(1) + 1; }
// Contexts not yet destroyed (a batch launch refuses a destroyed context).
std::mutex gLiveMutex;
std::unordered_set<const void*> gLive;

// Device address of `scratch` on each GPU (each GPU holds its own copy of the array),
// resolved once per device by zgl_create with that device current.
std::mutex gScratchMutex;
unsigned char* gScratch[kMaxDevices] = {};

void resolveScratch(int device)
{
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (device < 0 || device >= kMaxDevices || gScratch[device]) return;
    void* ptr = nullptr;
    hipCheck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "static scratch address");
    if (!ptr) hipCheck(hipErrorInvalidSymbol, "static scratch address");
    gScratch[device] = static_cast<unsigned char*>(ptr);
}

unsigned char* scratchOn(int device)
{
    std::lock_guard<std::mutex> lock(gScratchMutex);
    return (device >= 0 && device < kMaxDevices) ? gScratch[device] : nullptr;
}

}  // namespace

struct zgl_context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf rows, tris, indices, vertices, normals, texcoords, matTex, materials;
    DevBuf lightPower, lightAlias, lightProb, env, envAlias, envProb, noise, sobolM, sobolTab;
    DevBuf frame, film, counters, scratchRays, scratchOut, workCounter;
    // differentiable mode: evaluation point, adjoint input and outputs
    DevBuf evalMat, evalEnv, dFrame, dMat, dEnv;
    DevBuf bvhBounds;  // node bounds of a device-built MTBVH
    DevBuf gqQueues, gqCounts;  // global-queue path tracer
    int gqBlocksPerCU = 0;
    DevBuf wfWork;    // wavefront schedule of the naive tracer
    DevBuf tex, texScale, texRaw;  // texture array (linear RGB float4), per-layer uv scale, upload staging
    int maxTexId = -1;             // largest texture id the uploaded scene's triangles use
    DevBuf naiveRes;  // per-sample values of the persistent naive tracer
    DevBuf hitLog;    // per-sample traversal results of the differentiable mode's passes
    // The passes whose traversal results hitLog holds in full and the generation they were
    // logged at (0: none): the results depend only on the uploaded scene and the passes.
    uint64_t logGeneration = 0;
    PassParams logPasses = {};
    DevBuf refSum;    // FP64 reference: per-pixel RGB sums of the passes rendered so far
    int wfTraceBlocksPerCU = 0, wfShadeBlocksPerCU = 0;
    DeviceBvh bvh;     // work memory of the device MTBVH build
    std::vector<float> hostMaterials;  // uploaded material records (16 floats each)
    int numCUs = 0;
    int naiveBlocksPerCU = 0;
    int tripleBlocksPerCU = 0;
    int lightBlocksPerCU = 0, tripleLightBlocksPerCU = 0;
    DevScene scene = {};
    bool hasScene = false, hasCamera = false, hasNoise = false;
    int filmW = 0, filmH = 0;
    uint64_t generation = 0;  // stamp of the last change a batch would have to see
    void touch() { generation = nextGeneration(); }

    void upload(DevBuf& b, const void* src, size_t bytes, const char* what) { bufUpload(b, src, bytes, stream, what); }
    void ensure(DevBuf& b, size_t bytes, const char* what) { bufEnsure(b, bytes, what); }
    void releaseAll()
    {
        DevBuf* all[] = {&rows, &tris, &indices, &vertices, &normals, &texcoords, &matTex, &materials,
                         &lightPower, &lightAlias, &lightProb, &env, &envAlias, &envProb, &noise, &sobolM,
                         &sobolTab, &frame, &film, &counters, &scratchRays, &scratchOut, &workCounter,
                         &evalMat, &evalEnv, &dFrame, &dMat, &dEnv, &bvhBounds, &gqQueues, &gqCounts, &wfWork,
                         &naiveRes, &hitLog, &refSum, &tex, &texScale, &texRaw};
        for (DevBuf* b : all) b->release();
        bvh.release();
    }
};

struct zgl_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<zgl_context*> ctxs;
    std::vector<uint64_t> generations;  // per context: its stamp when the batch was set
    std::vector<zgl_pass_params> params;
    hipEvent_t fence = nullptr;          // orders the batch after / before the contexts' streams
    std::vector<BatchScene> host;  // in launch order
    std::vector<int32_t> order;    // launch position -> index in ctxs
    DevBuf items, counters, sobolTab, res;
    std::vector<size_t> resOffset;  // per context (index in ctxs): float4 offset into res
    int maxPixels = 0;
    int numCUs = 0, maxBlocksPerCU = 1;
    int threshold = 0, blocksPerCU = 0;  // threshold 0: each scene's zgl_render default
    int sobolFirst = 0, sobolCount = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void upload(DevBuf& b, const void* src, size_t bytes, const char* what) { bufUpload(b, src, bytes, stream, what); }
    void ensure(DevBuf& b, size_t bytes, const char* what) { bufEnsure(b, bytes, what); }
    void dropGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

template <class T, class F> zgl_status guarded(T* ctx, F&& f)
{
    if (!ctx) return ZGL_E_INVALID;
    try {
        hipCheck(hipSetDevice(ctx->device), "hipSetDevice");
        f();
        ctx->err.clear();
        return ZGL_OK;
    } catch (const ZglError& e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->err = "host allocation failed";
        return ZGL_E_OOM;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return ZGL_E_INVALID;
    }
}

template <class F> zgl_status guardedHost(F&& f)
{
    try {
        f();
        return ZGL_OK;
    } catch (const ZglError& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        return ZGL_E_OOM;
    } catch (...) {
        return ZGL_E_INVALID;
    }
}

}  // namespace


extern "C" {

// ---------------------------------------------------------------- host producers
zgl_status zgl_build_mtbvh(const float* vertices, int64_t n_vertices, const uint32_t* indices, int64_t n_triangles,
                           float* bounds_out, int32_t* hit_table_out)
{
    if (!vertices || !indices || !bounds_out || !hit_table_out || n_triangles < 1 || n_vertices < 1 ||
        n_triangles > (1LL << 29))
        return ZGL_E_INVALID;
    for (int64_t i = 0; i < n_triangles * 3; i++)
        if ((int64_t)indices[i] >= n_vertices) return ZGL_E_INVALID;
    return guardedHost([&] {
        std::vector<float> b;
        std::vector<int32_t> h;
        buildMtbvh(vertices, indices, n_triangles, b, h);
        std::memcpy(bounds_out, b.data(), b.size() * sizeof(float));
        std::memcpy(hit_table_out, h.data(), h.size() * sizeof(int32_t));
    });
}

zgl_status zgl_alias_build(const float* pdf, int32_t n, int32_t* alias_out, float* prob_out)
{
    if (n < 0 || (n > 0 && (!pdf || !alias_out || !prob_out))) return ZGL_E_INVALID;
    return guardedHost([&] { aliasBuild(pdf, n, alias_out, prob_out); });
}

zgl_status zgl_env_alias_build(const float* rgb, int32_t width, int32_t height, int32_t* alias_out,
                               float* prob_out, float* sum_out)
{
    if (!rgb || !alias_out || !prob_out || !sum_out || width < 1 || height < 1) return ZGL_E_INVALID;
    return guardedHost([&] { *sum_out = envAliasBuild(rgb, width, height, alias_out, prob_out); });
}

uint32_t zgl_sobol_sample(uint32_t index, int32_t dim, uint32_t scramble)
{
    if (dim < 0 || dim >= 256) return 0;
    return sobolSample(index, dim, scramble);
}

void zgl_sobol_matrices(uint32_t* out)
{
    if (out) std::memcpy(out, sobolMatrices(), 256 * 32 * sizeof(uint32_t));
}

void zgl_float_to_half(const float* in, uint16_t* out, int64_t n)
{
    if (!in || !out) return;
    for (int64_t i = 0; i < n; i++) out[i] = floatToHalf(in[i]);
}

// ---------------------------------------------------------------- context
int32_t zgl_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

zgl_status zgl_create(int32_t device, zgl_context** out)
{
    if (!out) return ZGL_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return ZGL_E_NODEVICE;
    if (device < 0 || device >= n) return ZGL_E_INVALID;
    zgl_context* ctx = new (std::nothrow) zgl_context();
    if (!ctx) return ZGL_E_OOM;
    ctx->device = device;
    zgl_status st = guarded(ctx, [&] {
        hipCheck(hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking), "hipStreamCreate");
        ctx->stream = ctx->own;
        resolveScratch(device);
        ctx->upload(ctx->sobolM, sobolMatrices(), 256 * 32 * sizeof(uint32_t), "sobol matrices");
        ctx->ensure(ctx->counters, NumCounters * sizeof(unsigned long long), "counters");
        ctx->ensure(ctx->workCounter, 256, "work counter");
        hipCheck(hipDeviceGetAttribute(&ctx->numCUs, hipDeviceAttributeMultiprocessorCount, device), "attribute");
        hipCheck(hipMemsetAsync(ctx->counters.p, 0, NumCounters * sizeof(unsigned long long), ctx->stream),
                 "counters");
        hipCheck(hipStreamSynchronize(ctx->stream), "sync");
        ctx->touch();
        std::lock_guard<std::mutex> lock(gLiveMutex);
        gLive.insert(ctx);
    });
    if (st != ZGL_OK) {
        ctx->releaseAll();
        if (ctx->own) (void)hipStreamDestroy(ctx->own);
        delete ctx;
        return st;
    }
    *out = ctx;
    return ZGL_OK;
}

void zgl_destroy(zgl_context* ctx)
{
    if (!ctx) return;
    {
        std::lock_guard<std::mutex> lock(gLiveMutex);
        gLive.erase(ctx);
    }
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->releaseAll();
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
}

const char* zgl_last_error(const zgl_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

zgl_status zgl_set_stream(zgl_context* ctx, void* s)
{
    return guarded(ctx, [&] {
        hipCheck(hipStreamSynchronize(ctx->stream), "sync");
        ctx->stream = s ? (hipStream_t)s : ctx->own;
    });
}

void* zgl_get_stream(zgl_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

zgl_status zgl_synchronize(zgl_context* ctx)
{
    return guarded(ctx, [&] { hipCheck(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"); });
}

// ---------------------------------------------------------------- scene upload
zgl_status zgl_upload_scene(zgl_context* ctx, const zgl_scene_buffers* sb)
{
    return guarded(ctx, [&] {
        ctx->touch();
        if (!sb) invalid("null scene");
        const int64_t nV = sb->n_vertices, nT = sb->n_triangles;
        if (nV < 3 || nT < 1 || !sb->vertices || !sb->normals || !sb->indices) invalid("empty geometry");
        // absolute MTBVH row indices (6 tables x (2 nT - 1) rows) must fit in int32
        if (nT > (1LL << 27) || nV > (1LL << 30)) invalid("scene too large");
        const bool deviceBvh = !sb->bounds && !sb->hit_table;
        if (sb->bvh_size != 2 * nT - 1) invalid("bvh_size must be 2*n_triangles-1");
        if (!deviceBvh && (!sb->bounds || !sb->hit_table)) invalid("bounds and hit_table: both or neither");
        if (sb->obj_prim_count < 0 || sb->obj_prim_count > nT) invalid("obj_prim_count out of range");
        if (sb->n_light_triangles != nT - sb->obj_prim_count) invalid("n_light_triangles != n_triangles - obj_prim_count");
        if (sb->n_texcoords < 0 || sb->n_texcoords > nV || (sb->n_texcoords > 0 && !sb->texcoords))
            invalid("texcoords");
        if (sb->n_materials < 1 || !sb->materials) invalid("materials");
        if (sb->obj_prim_count > 0 && !sb->mat_tex_indices) invalid("mat_tex_indices");
        if (sb->env_width < 1 || sb->env_height < 1 || !sb->env_rgb16f || !sb->env_alias || !sb->env_prob)
            invalid("environment map");
        for (int64_t i = 0; i < 3 * nT; i++)
            if (sb->indices[i] < 0 || sb->indices[i] >= nV) invalid("index out of range");
        int maxTexId = -1;
        for (int64_t i = 0; i < sb->obj_prim_count; i++) {
            if ((sb->mat_tex_indices[i] & 0xffff) >= sb->n_materials) invalid("material index out of range");
            const int texId = sb->mat_tex_indices[i] >> 16;  // (texIndex << 16 | matIndex), texIndex -1: none
            if (texId < -1) invalid("texture index out of range");
            maxTexId = std::max(maxTexId, texId);
        }
        const int64_t B = sb->bvh_size;
        // Finite bounds: the box test of general rays relies on NaN-free slab distances
        // (the device build checks the vertices instead).
        for (int64_t i = 0; !deviceBvh && i < B * 6; i++)
            if (!(std::fabs(sb->bounds[i]) <= 3.4028235e38f)) invalid("non-finite MTBVH bounds");
        // Every row must point forward so that every traversal terminates.
        for (int t = 0; t < 6 && !deviceBvh; t++)
            for (int64_t k = 0; k < B; k++) {
                const int32_t* r = sb->hit_table + (t * B + k) * 3;
                if (r[0] < 0 || r[0] >= B || r[1] < -1 || r[1] >= nT || r[2] <= k || r[2] > B)
                    invalid("malformed hit table row");
            }
        const int nL = sb->n_light_triangles;
        if (nL > 0) {
            if (!sb->light_power || !sb->light_alias || !sb->light_prob) invalid("light tables");
            for (int i = 0; i < nL; i++)
                if (sb->light_alias[i] < 0 || sb->light_alias[i] >= nL) invalid("light alias out of range");
        }
        const int w = sb->env_width, h = sb->env_height;
        for (int i = 0; i < h; i++) {
            for (int j = 0; j < w; j++) {
                int a = sb->env_alias[(size_t)i * (w + 1) + j];
                if (a < 0 || a >= w) invalid("env alias out of range");
            }
            int a = sb->env_alias[(size_t)i * (w + 1) + w];
            if (a < 0 || a >= h) invalid("env marginal alias out of range");
        }

        // MTBVH row records: {pMin.xyz, pMax.x} {pMax.yz, prim, miss}, the miss link
        // made absolute (table * B + miss) so that a walk needs no per-table base
        std::vector<float> rows(deviceBvh ? 0 : (size_t)6 * B * 8);
        for (int t = 0; t < 6 && !deviceBvh; t++)
            for (int64_t k = 0; k < B; k++) {
                const int32_t* r = sb->hit_table + (t * B + k) * 3;
                const float* bb = sb->bounds + (size_t)r[0] * 6;
                float* o = rows.data() + ((size_t)t * B + k) * 8;
                o[0] = bb[0]; o[1] = bb[1]; o[2] = bb[2]; o[3] = bb[3];
                o[4] = bb[4]; o[5] = bb[5];
                const int32_t miss = (int32_t)(t * B) + r[2];
                std::memcpy(o + 6, &r[1], 4);
                std::memcpy(o + 7, &miss, 4);
            }
        std::vector<float> tris((size_t)nT * 12, 0.0f);
        for (int64_t i = 0; i < nT; i++) {
            float* o = tris.data() + i * 12;
            for (int v = 0; v < 3; v++)
                for (int c = 0; c < 3; c++) o[v * 3 + c] = sb->vertices[(size_t)sb->indices[3 * i + v] * 3 + c];
        }
        std::vector<uint16_t> env((size_t)w * h * 4, 0);
        for (size_t i = 0; i < (size_t)w * h; i++)
            for (int c = 0; c < 3; c++) env[i * 4 + c] = sb->env_rgb16f[i * 3 + c];

        if (!deviceBvh) ctx->upload(ctx->rows, rows.data(), rows.size() * 4, "rows");
        ctx->upload(ctx->tris, tris.data(), tris.size() * 4, "triangles");
        ctx->upload(ctx->indices, sb->indices, (size_t)nT * 12, "indices");
        ctx->upload(ctx->vertices, sb->vertices, (size_t)nV * 12, "vertices");
        if (deviceBvh) {
            // indices were checked to lie in [0, n_vertices): the int32 buffer reads as uint32
            ctx->ensure(ctx->rows, (size_t)6 * B * 32, "rows");
            ctx->ensure(ctx->bvhBounds, (size_t)B * 24, "MTBVH bounds");
            std::string msg;
            const int rc = ctx->bvh.build((const float*)ctx->vertices.p, nV, (const uint32_t*)ctx->indices.p, nT,
                                          kDefaultBvhSubtree, (float*)ctx->bvhBounds.p, nullptr, ctx->rows.p,
                                          ctx->stream, msg);
            if (rc) throw ZglError{rc == 1 ? ZGL_E_INVALID : (rc == 3 ? ZGL_E_OOM : ZGL_E_HIP), msg};
        }
        ctx->upload(ctx->normals, sb->normals, (size_t)nV * 12, "normals");
        ctx->upload(ctx->texcoords, sb->texcoords, (size_t)sb->n_texcoords * 8, "texcoords");
        ctx->upload(ctx->matTex, sb->mat_tex_indices, (size_t)sb->obj_prim_count * 4, "matTexIndices");
        ctx->maxTexId = maxTexId;
        ctx->upload(ctx->materials, sb->materials, (size_t)sb->n_materials * 64, "materials");
        ctx->upload(ctx->lightPower, sb->light_power, (size_t)nL * 12, "lightPower");
        ctx->upload(ctx->lightAlias, sb->light_alias, (size_t)nL * 4, "lightAlias");
        ctx->upload(ctx->lightProb, sb->light_prob, (size_t)nL * 4, "lightProb");
        ctx->upload(ctx->env, env.data(), env.size() * 2, "env");
        ctx->upload(ctx->envAlias, sb->env_alias, (size_t)(w + 1) * h * 4, "envAlias");
        ctx->upload(ctx->envProb, sb->env_prob, (size_t)(w + 1) * h * 4, "envProb");
        hipCheck(hipStreamSynchronize(ctx->stream), "scene upload");

        DevScene& s = ctx->scene;
        s.rows = (const float4*)ctx->rows.p;
        s.tris = (const float4*)ctx->tris.p;
        s.indices = (const int32_t*)ctx->indices.p;
        s.vertices = (const float*)ctx->vertices.p;
        s.normals = (const float*)ctx->normals.p;
        s.texcoords = (const float*)ctx->texcoords.p;
        s.matTexIndices = (const int32_t*)ctx->matTex.p;
        s.materials = (const float4*)ctx->materials.p;
        s.lightPower = (const float*)ctx->lightPower.p;
        s.lightAlias = (const int32_t*)ctx->lightAlias.p;
        s.lightProb = (const float*)ctx->lightProb.p;
        s.env = (const ushort4*)ctx->env.p;
        s.envAlias = (const int32_t*)ctx->envAlias.p;
        s.envProb = (const float*)ctx->envProb.p;
        s.sobolM = (const uint32_t*)ctx->sobolM.p;
        s.nVertices = nV;
        s.nTexcoords = (int32_t)sb->n_texcoords;
        s.nTriangles = (int32_t)nT;
        s.objPrimCount = sb->obj_prim_count;
        s.bvhSize = sb->bvh_size;
        s.nMaterials = sb->n_materials;
        s.nLightTriangles = nL;
        s.lightSum = sb->light_sum;
        s.envW = w;
        s.envH = h;
        s.envSum = sb->env_sum;
        sincosZ(sb->env_rotation, &s.envRotSin, &s.envRotCos);
        sincosZ(-sb->env_rotation, &s.envRotSinN, &s.envRotCosN);
        s.envRotation = sb->env_rotation;
        s.sampler = sb->sampler ? 1 : 0;
        // evaluation point of the differentiable mode: the uploaded values
        ctx->hostMaterials.assign(sb->materials, sb->materials + (size_t)sb->n_materials * 16);
        ctx->upload(ctx->evalMat, sb->materials, (size_t)sb->n_materials * 64, "eval materials");
        ctx->ensure(ctx->evalEnv, (size_t)w * h * 16, "eval env");
        hipCheck(launchEnvExpand(s.env, (float4*)ctx->evalEnv.p, (int64_t)w * h, ctx->stream), "eval env");
        hipCheck(hipStreamSynchronize(ctx->stream), "scene upload");
        ctx->hasScene = true;
    });
}

zgl_status zgl_set_camera(zgl_context* ctx, const zgl_camera* c)
{
    return guarded(ctx, [&] {
        ctx->touch();
        if (!c) invalid("null camera");
        if (c->film_width < 1 || c->film_height < 1 || c->film_width > 16384 || c->film_height > 16384)
            invalid("film size");
        DevScene& s = ctx->scene;
        s.camF = v3(c->front[0], c->front[1], c->front[2]);
        s.camR = v3(c->right[0], c->right[1], c->right[2]);
        s.camU = v3(c->up[0], c->up[1], c->up[2]);
        s.camPos = v3(c->pos[0], c->pos[1], c->pos[2]);
        s.camMatInv.c0 = v3(c->mat_inv[0], c->mat_inv[1], c->mat_inv[2]);
        s.camMatInv.c1 = v3(c->mat_inv[3], c->mat_inv[4], c->mat_inv[5]);
        s.camMatInv.c2 = v3(c->mat_inv[6], c->mat_inv[7], c->mat_inv[8]);
        s.tanFOV = c->tan_fov;
        s.camAsp = c->aspect;
        s.lensRadius = c->lens_radius;
        s.focalDist = c->focal_dist;
        if (c->film_width != ctx->filmW || c->film_height != ctx->filmH || !ctx->frame.p) {
            const size_t px = (size_t)c->film_width * c->film_height;
            ctx->frame.release();
            ctx->film.release();
            ctx->ensure(ctx->frame, px * 16, "frame");
            ctx->ensure(ctx->film, px * 12, "film");
            hipCheck(launchFill((float*)ctx->frame.p, (int64_t)px * 4, 0.0f, ctx->stream), "clear");
            hipCheck(launchFill((float*)ctx->film.p, (int64_t)px * 3, 0.0f, ctx->stream), "clear");
            ctx->filmW = c->film_width;
            ctx->filmH = c->film_height;
        }
        s.filmW = ctx->filmW;
        s.filmH = ctx->filmH;
        ctx->hasCamera = true;
    });
}

zgl_status zgl_set_noise(zgl_context* ctx, const float* noise, int32_t w, int32_t h)
{
    return guarded(ctx, [&] {
        ctx->touch();
        if (!noise || w < 1 || h < 1 || (int64_t)w * h > (1LL << 28)) invalid("noise texture");
        ctx->upload(ctx->noise, noise, (size_t)w * h * 8, "noise");
        hipCheck(hipStreamSynchronize(ctx->stream), "noise upload");
        ctx->scene.noise = (const float2*)ctx->noise.p;
        ctx->scene.noiseW = w;
        ctx->scene.noiseH = h;
        ctx->hasNoise = true;
    });
}

// The sRGB8 texture array of Scene::createGLContext (src/core/Scene.cpp:256-258;
// Texture2DArray, src/core/Texture.cpp:134-171), decoded to linear FP32 on upload.
zgl_status zgl_set_textures(zgl_context* ctx, const uint8_t* rgb8, int32_t n_layers, int32_t width, int32_t height,
                            const float* uv_scale)
{
    return guarded(ctx, [&] {
        ctx->touch();
        if (n_layers == 0) {
            ctx->scene.tex = nullptr;
            ctx->scene.texScale = nullptr;
            ctx->scene.nTextures = ctx->scene.texW = ctx->scene.texH = 0;
            return;
        }
        if (!rgb8 || !uv_scale || n_layers < 0 || n_layers > 32767 || width < 1 || height < 1 || width > 16384 ||
            height > 16384 || (int64_t)n_layers * width * height > (1LL << 28))
            invalid("textures: layers / size");
        for (int i = 0; i < 2 * n_layers; i++)
            if (!(std::fabs(uv_scale[i]) <= 3.4028235e38f)) invalid("textures: non-finite uv scale");
        const size_t n = (size_t)n_layers * width * height;
        // the GL_SRGB decode of each 8-bit value, in FP64 rounded to FP32
        float lut[256];
        for (int b = 0; b < 256; b++) {
            const double c = b / 255.0;
            lut[b] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
        }
        const size_t rawBytes = (n * 3 + 255) & ~(size_t)255;
        ctx->ensure(ctx->texRaw, rawBytes + sizeof(lut), "texture staging");
        hipCheck(hipMemcpyAsync(ctx->texRaw.p, rgb8, n * 3, hipMemcpyHostToDevice, ctx->stream), "textures");
        unsigned char* lutDev = static_cast<unsigned char*>(ctx->texRaw.p) + rawBytes;
        hipCheck(hipMemcpyAsync(lutDev, lut, sizeof(lut), hipMemcpyHostToDevice, ctx->stream), "textures");
        ctx->ensure(ctx->tex, n * 16, "textures");
        hipCheck(launchSrgbExpand((const uint8_t*)ctx->texRaw.p, (const float*)lutDev, (float4*)ctx->tex.p, (int64_t)n,
                                  ctx->stream),
                 "textures");
        ctx->upload(ctx->texScale, uv_scale, (size_t)n_layers * 8, "texture uv scale");
        hipCheck(hipStreamSynchronize(ctx->stream), "textures");
        ctx->scene.tex = (const float4*)ctx->tex.p;
        ctx->scene.texScale = (const float2*)ctx->texScale.p;
        ctx->scene.texW = width;
        ctx->scene.texH = height;
        ctx->scene.nTextures = n_layers;
    });
}

// The scene's triangles name textures that have been set.
static void checkTextures(const zgl_context* ctx)
{
    if (ctx->maxTexId >= ctx->scene.nTextures)
        throw ZglError{ZGL_E_STATE, "the scene uses texture " + std::to_string(ctx->maxTexId) + " but " +
                                        std::to_string(ctx->scene.nTextures) + " are set (zgl_set_textures)"};
}

// ---------------------------------------------------------------- tracing
static void checkRays(zgl_context* ctx, int64_t n)
{
    if (!ctx->hasScene) throw ZglError{ZGL_E_STATE, "no scene uploaded"};
    if (n < 0 || n > (1LL << 31)) invalid("ray count");
}

zgl_status zgl_trace_closest_device(zgl_context* ctx, const float* d_rays, int64_t n, int32_t* d_prim, float* d_t,
                                    uint32_t* d_bin)
{
    return guarded(ctx, [&] {
        checkRays(ctx, n);
        if (n && (!d_rays || !d_prim || !d_t || !d_bin)) invalid("null device pointer");
        hipCheck(launchTraceClosest(ctx->scene, d_rays, n, d_prim, d_t, d_bin, nullptr, ctx->stream),
                 "trace_closest");
    });
}

static zgl_status traceClosestHost(zgl_context* ctx, const float* rays, int64_t n, int32_t* prim, float* t,
                                   uint32_t* bin, bool count)
{
    return guarded(ctx, [&] {
        checkRays(ctx, n);
        if (n == 0) return;
        if (!rays || !prim || !t || !bin) invalid("null pointer");
        ctx->upload(ctx->scratchRays, rays, (size_t)n * 24, "rays");
        ctx->ensure(ctx->scratchOut, (size_t)n * 12, "trace output");
        int32_t* dp = (int32_t*)ctx->scratchOut.p;
        float* dt = (float*)(dp + n);
        uint32_t* db = (uint32_t*)(dt + n);
        hipCheck(launchTraceClosest(ctx->scene, (const float*)ctx->scratchRays.p, n, dp, dt, db,
                                    count ? (unsigned long long*)ctx->counters.p : nullptr, ctx->stream),
                 "trace_closest");
        hipCheck(hipMemcpyAsync(prim, dp, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy");
        hipCheck(hipMemcpyAsync(t, dt, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy");
        hipCheck(hipMemcpyAsync(bin, db, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy");
        hipCheck(hipStreamSynchronize(ctx->stream), "trace_closest");
    });
}

zgl_status zgl_trace_closest(zgl_context* ctx, const float* rays, int64_t n, int32_t* prim, float* t,
                             uint32_t* bin)
{
    return traceClosestHost(ctx, rays, n, prim, t, bin, false);
}

zgl_status zgl_trace_closest_counted(zgl_context* ctx, const float* rays, int64_t n, int32_t* prim, float* t,
                                     uint32_t* bin)
{
    return traceClosestHost(ctx, rays, n, prim, t, bin, true);
}

zgl_status zgl_trace_any(zgl_context* ctx, const float* rays, const float* tmax, int64_t n, uint8_t* occ)
{
    return guarded(ctx, [&] {
        checkRays(ctx, n);
        if (n == 0) return;
        if (!rays || !tmax || !occ) invalid("null pointer");
        std::vector<float> buf((size_t)n * 7);
        std::memcpy(buf.data(), rays, (size_t)n * 24);
        std::memcpy(buf.data() + 6 * n, tmax, (size_t)n * 4);
        ctx->upload(ctx->scratchRays, buf.data(), buf.size() * 4, "rays");
        ctx->ensure(ctx->scratchOut, (size_t)n, "trace output");
        const float* dr = (const float*)ctx->scratchRays.p;
        hipCheck(launchTraceAny(ctx->scene, dr, dr + 6 * n, n, (uint8_t*)ctx->scratchOut.p, ctx->stream),
                 "trace_any");
        hipCheck(hipMemcpyAsync(occ, ctx->scratchOut.p, (size_t)n, hipMemcpyDeviceToHost, ctx->stream), "copy");
        hipCheck(hipStreamSynchronize(ctx->stream), "trace_any");
    });
}

// ---------------------------------------------------------------- rendering
// Validates a pass request.
static void checkPasses(zgl_context* ctx, const zgl_pass_params* p)
{
    if (!p) invalid("null params");
    if (!ctx->hasScene || !ctx->hasCamera || !ctx->hasNoise)
        throw ZglError{ZGL_E_STATE, "scene, camera and noise must be set before rendering"};
    if (p->n_passes < 1 || p->n_passes > 65536 || p->first_pass < 0 || p->first_pass > (1 << 24) ||
        p->max_depth < 0 || p->max_depth > 1024)
        invalid("pass range / depth");
    if (p->integrator < 0 || p->integrator > 4) invalid("integrator");
    if (p->integrator == 4 && p->count_bytes) invalid("count_bytes is not supported by the wavefront schedule");
    if (p->integrator == ZGL_INTEGRATOR_GLOBAL_QUEUE && p->count_bytes)
        invalid("count_bytes is not supported by the global-queue integrator");
    if ((p->integrator == ZGL_INTEGRATOR_LIGHT || p->integrator == ZGL_INTEGRATOR_TRIPLE) &&
        ctx->scene.nLightTriangles < 1)
        invalid("light and triple integrators need light triangles");
    if (p->regen_threshold < 0 || p->regen_threshold > 63) invalid("regen_threshold must be in [0, 63]");
    if ((p->integrator == ZGL_INTEGRATOR_LIGHT || p->integrator == ZGL_INTEGRATOR_TRIPLE) &&
        (p->blocks_one_pass < 1 || p->blocks_one_pass > 65536))
        invalid("blocks_one_pass");
    if (p->integrator == ZGL_INTEGRATOR_TRIPLE && (p->loops_per_pass < 1 || p->loops_per_pass > 1024))
        invalid("loops_per_pass");
    checkTextures(ctx);
}

// The kernels' pass parameters of a request.
static PassParams passParams(const zgl_pass_params* p)
{
    PassParams pp;
    pp.firstPass = p->first_pass;
    pp.nPasses = p->n_passes;
    pp.freeCounter = p->free_counter;
    pp.maxDepth = p->max_depth;
    pp.russianRoulette = p->russian_roulette;
    pp.sampleLight = p->sample_light;
    pp.lightEnvUniform = p->light_env_uniform;
    pp.lightPortion = p->light_portion;
    pp.loopsPerPass = p->loops_per_pass > 0 ? p->loops_per_pass : 1;
    pp.pathsPerPass = 1536 * (p->blocks_one_pass > 0 ? p->blocks_one_pass : 1);
    pp.lptScale = 0.0f;
    return pp;
}

// Validates a pass request and builds the per-launch scene view and parameters.
static void preparePasses(zgl_context* ctx, const zgl_pass_params* p, DevScene& s, PassParams& pp)
{
    checkPasses(ctx, p);
    // Sobol points of the passes in this launch, plus one for dims spilling past 256.
    const int first = p->first_pass, count = p->n_passes + 1;
    ctx->ensure(ctx->sobolTab, (size_t)count * 256 * 4, "sobol table");
    hipCheck(launchSobolTable((const uint32_t*)ctx->sobolM.p, (uint32_t*)ctx->sobolTab.p, first, count,
                              ctx->stream),
             "sobol table");
    s = ctx->scene;
    s.sobolTab = (const uint32_t*)ctx->sobolTab.p;
    s.sobolTabFirst = first;
    s.sobolTabCount = count;
    pp = passParams(p);
}

zgl_status zgl_render(zgl_context* ctx, const zgl_pass_params* p)
{
    return guarded(ctx, [&] {
        DevScene s;
        PassParams pp;
        preparePasses(ctx, p, s, pp);
        unsigned long long* cnt = p->count_bytes ? (unsigned long long*)ctx->counters.p : nullptr;
        float4* frame = (float4*)ctx->frame.p;
        float* film = (float*)ctx->film.p;
        if (p->integrator == ZGL_INTEGRATOR_NAIVE) {
            if (!ctx->naiveBlocksPerCU) ctx->naiveBlocksPerCU = naivePersistentBlocksPerCU();
            const int threshold =
                p->regen_threshold > 0 ? p->regen_threshold : defaultThreshold(WalkKind::Naive, s.bvhSize);
            // per-sample values of up to 4 GB worth of passes (121 passes at 1080p)
            const size_t perPass = (size_t)ctx->filmW * ctx->filmH * 16;
            const int resPasses = (int)std::max<size_t>(1, std::min<size_t>(pp.nPasses, ((size_t)4 << 30) / perPass));
            ctx->ensure(ctx->naiveRes, perPass * resPasses, "sample values");
            hipCheck(launchNaivePersistent(s, pp, frame, (float4*)ctx->naiveRes.p, resPasses, (int*)ctx->workCounter.p,
                                           threshold, ctx->numCUs * ctx->naiveBlocksPerCU, cnt, ctx->stream),
                     "naive pass");
        } else if (p->integrator == ZGL_INTEGRATOR_NAIVE_WAVEFRONT) {
            if (!ctx->wfTraceBlocksPerCU) wavefrontBlocks(ctx->wfTraceBlocksPerCU, ctx->wfShadeBlocksPerCU);
            // up to 24 GB of path state: 44 passes of 1080p at once
            const int batch = wavefrontBatch(ctx->filmW, ctx->filmH, pp.nPasses, (size_t)24 << 30);
            ctx->ensure(ctx->wfWork, wavefrontBytes(ctx->filmW, ctx->filmH, batch) + (size_t)3 * (pp.maxDepth + 2) * 4,
                        "wavefront work");
            hipCheck(launchNaiveWavefront(s, pp, frame, ctx->wfWork.p, batch, ctx->numCUs * ctx->wfTraceBlocksPerCU,
                                          ctx->numCUs * ctx->wfShadeBlocksPerCU, ctx->stream),
                     "wavefront pass");
        } else if (p->integrator == ZGL_INTEGRATOR_GLOBAL_QUEUE) {
            if (!ctx->gqBlocksPerCU) ctx->gqBlocksPerCU = globalQueueBlocks();
            ctx->ensure(ctx->gqQueues, globalQueueBytes(ctx->filmW, ctx->filmH), "global queue");
            ctx->ensure(ctx->gqCounts, (size_t)(pp.maxDepth + 1) * 4, "global queue counts");
            hipCheck(launchGlobalQueue(s, pp, frame, ctx->gqQueues.p, (int*)ctx->gqCounts.p,
                                       ctx->numCUs * ctx->gqBlocksPerCU, ctx->stream),
                     "global-queue pass");
        } else if (p->integrator == ZGL_INTEGRATOR_LIGHT) {
            s.sampler = 0;  // LightPath.cpp:49 forces the independent sampler
            if (cnt) {
                // the byte-counting variant: one thread per light path
                hipCheck(launchLight(s, pp, film, cnt, ctx->stream), "light pass");
            } else {
                if (!ctx->lightBlocksPerCU) ctx->lightBlocksPerCU = lightPersistentBlocksPerCU(false);
                const int threshold =
                    p->regen_threshold > 0 ? p->regen_threshold : defaultThreshold(WalkKind::LightPaths, s.bvhSize);
                hipCheck(launchLightPersistent(s, pp, false, film, (int*)ctx->workCounter.p, threshold,
                                               ctx->numCUs * ctx->lightBlocksPerCU, ctx->stream),
                         "light pass");
            }
            hipCheck(launchFilmToFrame(film, frame, ctx->filmW, ctx->filmH, ctx->stream), "film copy");
        } else {
            // uScale = W*H / (blocks * loops * 1536)  (TriplePath.cpp:76-77)
            pp.lptScale = (float)(ctx->filmW * ctx->filmH) / (float)(p->blocks_one_pass * pp.loopsPerPass * 1536);
            if (cnt) {
                // the byte-counting variant of the camera pass: one thread per pixel
                hipCheck(launchTriplePT(s, pp, film, cnt, ctx->stream), "triple camera pass");
            } else {
                if (!ctx->tripleBlocksPerCU) ctx->tripleBlocksPerCU = triplePersistentBlocksPerCU();
                const size_t perPass = (size_t)ctx->filmW * ctx->filmH * 16;
                const int resPasses =
                    (int)std::max<size_t>(1, std::min<size_t>(pp.nPasses, ((size_t)4 << 30) / perPass));
                ctx->ensure(ctx->naiveRes, perPass * resPasses, "sample values");
                const int threshold =
                    p->regen_threshold > 0 ? p->regen_threshold : defaultThreshold(WalkKind::TripleCamera, s.bvhSize);
                hipCheck(launchTriplePersistent(s, pp, film, (float4*)ctx->naiveRes.p, resPasses,
                                                (int*)ctx->workCounter.p, threshold,
                                                ctx->numCUs * ctx->tripleBlocksPerCU, ctx->stream),
                         "triple camera pass");
            }
            DevScene sl = s;
            sl.sampler = 0;  // TriplePath.cpp:72
            if (cnt) {
                hipCheck(launchTripleLPT(sl, pp, film, cnt, ctx->stream), "triple light pass");
            } else {
                if (!ctx->tripleLightBlocksPerCU) ctx->tripleLightBlocksPerCU = lightPersistentBlocksPerCU(true);
                const int threshold =
                    p->regen_threshold > 0 ? p->regen_threshold : defaultThreshold(WalkKind::LightPaths, s.bvhSize);
                hipCheck(launchLightPersistent(sl, pp, true, film, (int*)ctx->workCounter.p, threshold,
                                               ctx->numCUs * ctx->tripleLightBlocksPerCU, ctx->stream),
                         "triple light pass");
            }
            hipCheck(launchFilmToFrame(film, frame, ctx->filmW, ctx->filmH, ctx->stream), "film copy");
        }
    });
}

// zgl_render for ZGL_INTEGRATOR_NAIVE with every buffer it writes besides the frame (unit
// counters, Sobol points, per-sample values) in the GPU's static `scratch` array, so the
// call allocates nothing. Passes run in chunks that fit the array; each chunk builds its
// own Sobol rows, which hold the values the one-table form reads.
zgl_status zgl_render_static(zgl_context* ctx, const zgl_pass_params* p)
{
    return guarded(ctx, [&] {
        checkPasses(ctx, p);
        if (p->integrator != ZGL_INTEGRATOR_NAIVE) invalid("zgl_render_static runs the naive integrator only");
        unsigned char* base = scratchOn(ctx->device);
        if (!base) throw ZglError{ZGL_E_STATE, "static scratch not resolved on this device"};
        const size_t perPass = (size_t)ctx->filmW * ctx->filmH * 16;
        if (perPass > kStaticResBytes)
            invalid("zgl_render_static: a " + std::to_string(ctx->filmW) + "x" + std::to_string(ctx->filmH) +
                    " film needs " + std::to_string(perPass) + " bytes of sample values per pass; the static scratch " +
                    "holds " + std::to_string(kStaticResBytes));
        const int chunk = (int)std::min<size_t>({(size_t)p->n_passes, kStaticResBytes / perPass,
                                                 (size_t)kStaticChunkPasses});
        if (!ctx->naiveBlocksPerCU) ctx->naiveBlocksPerCU = naivePersistentBlocksPerCU();
        const int threshold =
            p->regen_threshold > 0 ? p->regen_threshold : defaultThreshold(WalkKind::Naive, ctx->scene.bvhSize);
        int* counter = reinterpret_cast<int*>(base);
        uint32_t* sobol = reinterpret_cast<uint32_t*>(base + kStaticSobolOffset);
        float4* res = reinterpret_cast<float4*>(base + kStaticResOffset);
        unsigned long long* cnt = p->count_bytes ? (unsigned long long*)ctx->counters.p : nullptr;
        const PassParams pp = passParams(p);
        DevScene s = ctx->scene;
        s.sobolTab = sobol;
        for (int k0 = 0; k0 < pp.nPasses; k0 += chunk) {
            PassParams q = pp;
            q.firstPass = pp.firstPass + k0;
            q.freeCounter = pp.freeCounter + k0;
            q.nPasses = std::min(chunk, pp.nPasses - k0);
            s.sobolTabFirst = q.firstPass;
            s.sobolTabCount = q.nPasses + 1;
            hipCheck(launchSobolTable((const uint32_t*)ctx->sobolM.p, sobol, q.firstPass, q.nPasses + 1, ctx->stream),
                     "sobol table");
            hipCheck(launchNaivePersistent(s, q, (float4*)ctx->frame.p, res, q.nPasses, counter, threshold,
                                           ctx->numCUs * ctx->naiveBlocksPerCU, cnt, ctx->stream),
                     "naive pass");
        }
    });
}

// The naive integrator's passes in FP64 (the converged reference), added to a host array
// of per-pixel RGB sums.
zgl_status zgl_render_reference_f64(zgl_context* ctx, const zgl_pass_params* p, double* rgb_sum)
{
    return guarded(ctx, [&] {
        if (!rgb_sum) invalid("rgb_sum is NULL");
        if (p && p->integrator != ZGL_INTEGRATOR_NAIVE) invalid("the FP64 reference renders the naive integrator only");
        if (p && p->count_bytes) invalid("the FP64 reference does not count bytes");
        DevScene s;
        PassParams pp;
        preparePasses(ctx, p, s, pp);
        const size_t bytes = (size_t)ctx->filmW * ctx->filmH * 3 * sizeof(double);
        ctx->ensure(ctx->refSum, bytes, "FP64 reference sums");
        hipCheck(hipMemcpyAsync(ctx->refSum.p, rgb_sum, bytes, hipMemcpyHostToDevice, ctx->stream), "reference upload");
        // launches of at most 64 passes (about 1.8 s at 1080p), each adding to the sums
        for (int k = 0; k < pp.nPasses; k += 64) {
            PassParams q = pp;
            q.firstPass = pp.firstPass + k;
            q.freeCounter = pp.freeCounter + k;
            q.nPasses = std::min(64, pp.nPasses - k);
            hipCheck(launchNaiveF64(s, q, (double*)ctx->refSum.p, ctx->stream), "FP64 reference pass");
        }
        hipCheck(hipMemcpyAsync(rgb_sum, ctx->refSum.p, bytes, hipMemcpyDeviceToHost, ctx->stream), "reference download");
        hipCheck(hipStreamSynchronize(ctx->stream), "FP64 reference");
    });
}

// ---------------------------------------------------------------- differentiable mode
zgl_status zgl_set_eval_params(zgl_context* ctx, const float* material_params, const float* env_rgb)
{
    return guarded(ctx, [&] {
        if (!ctx->hasScene) throw ZglError{ZGL_E_STATE, "no scene uploaded"};
        const DevScene& s = ctx->scene;
        std::vector<float> rec(ctx->hostMaterials);
        if (material_params) {
            for (int m = 0; m < s.nMaterials; m++) {
                const float* q = material_params + (size_t)m * 5;
                for (int k = 0; k < 5; k++)
                    if (!(std::fabs(q[k]) <= 3.4028235e38f)) invalid("non-finite material parameter");
                float* r = rec.data() + (size_t)m * 16;
                r[0] = q[0]; r[1] = q[1]; r[2] = q[2]; r[3] = q[3]; r[5] = q[4];
            }
        }
        ctx->upload(ctx->evalMat, rec.data(), rec.size() * 4, "eval materials");
        const int64_t n = (int64_t)s.envW * s.envH;
        if (env_rgb) {
            std::vector<float> env((size_t)n * 4, 0.0f);
            for (int64_t i = 0; i < n; i++)
                for (int c = 0; c < 3; c++) env[i * 4 + c] = env_rgb[i * 3 + c];
            ctx->upload(ctx->evalEnv, env.data(), env.size() * 4, "eval env");
        } else {
            hipCheck(launchEnvExpand(s.env, (float4*)ctx->evalEnv.p, n, ctx->stream), "eval env");
        }
        hipCheck(hipStreamSynchronize(ctx->stream), "eval params");
    });
}

}  // extern "C"

// Whether hitLog holds the traversal results of exactly these passes of the current scene.
static bool logHolds(const zgl_context* ctx, const PassParams& pp)
{
    const PassParams& l = ctx->logPasses;
    return ctx->logGeneration != 0 && ctx->logGeneration == ctx->generation && l.firstPass == pp.firstPass &&
           l.nPasses == pp.nPasses && l.freeCounter == pp.freeCounter && l.maxDepth == pp.maxDepth &&
           l.russianRoulette == pp.russianRoulette && l.sampleLight == pp.sampleLight &&
           l.lightEnvUniform == pp.lightEnvUniform && l.lightPortion == pp.lightPortion;
}

// The differentiable mode's passes in chunks whose traversal log fits kAdjointLogBytes:
// for each chunk the persistent forward kernel logs every sample's traversal results (the
// rays depend only on the uploaded scene, which all sampling uses), then `replay` runs the
// chunk's samples from the log.  A request that fits one chunk and whose passes the log
// already holds (a zgl_adjoint after zgl_render_replay of the same passes) skips the
// forward kernel; such a log is kept for the next call.
template <class Replay>
static void loggedPasses(zgl_context* ctx, const DevScene& s, const PassParams& pp, int regenThreshold,
                         Replay replay)
{
    if (!ctx->naiveBlocksPerCU) ctx->naiveBlocksPerCU = naivePersistentBlocksPerCU();
    const size_t px = (size_t)ctx->filmW * ctx->filmH;
    const size_t perPass = px * (16 + (size_t)kHitLogQueries * 8);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(pp.nPasses, kAdjointLogBytes / perPass));
    const bool reuse = chunk == pp.nPasses && logHolds(ctx, pp);
    ctx->logGeneration = 0;
    ctx->ensure(ctx->naiveRes, px * 16 * chunk, "sample values");
    ctx->ensure(ctx->hitLog, px * kHitLogQueries * 8 * chunk, "traversal log");
    const int threshold = regenThreshold > 0 ? regenThreshold : defaultThreshold(WalkKind::Naive, s.bvhSize);
    for (int k0 = 0; k0 < pp.nPasses; k0 += chunk) {
        PassParams q = pp;
        q.firstPass = pp.firstPass + k0;
        q.freeCounter = pp.freeCounter + k0;
        q.nPasses = std::min(chunk, pp.nPasses - k0);
        if (!reuse)
            hipCheck(launchNaiveLogged(s, q, (float4*)ctx->naiveRes.p, (int2*)ctx->hitLog.p, (int*)ctx->workCounter.p,
                                       threshold, ctx->numCUs * ctx->naiveBlocksPerCU, ctx->stream),
                     "logged forward pass");
        replay(q);
    }
    if (chunk == pp.nPasses) {
        ctx->logPasses = pp;
        ctx->logGeneration = ctx->generation;
    }
}

extern "C" {

zgl_status zgl_render_replay(zgl_context* ctx, const zgl_pass_params* p)
{
    return guarded(ctx, [&] {
        DevScene s;
        PassParams pp;
        preparePasses(ctx, p, s, pp);
        if (p->integrator != ZGL_INTEGRATOR_NAIVE) invalid("the replay estimator is the naive integrator's");
        if (ctx->maxTexId >= 0) invalid("the differentiable mode does not take textured materials");
        // each chunk's samples are added to the frame in pass order, after the previous chunk's
        loggedPasses(ctx, s, pp, p->regen_threshold, [&](const PassParams& q) {
            hipCheck(launchNaiveReplay(s, q, (const float4*)ctx->evalMat.p, (const float4*)ctx->evalEnv.p,
                                       (float4*)ctx->frame.p, nullptr, nullptr, nullptr, 1, nullptr,
                                       (const int2*)ctx->hitLog.p, ctx->stream),
                     "replay pass");
        });
    });
}

zgl_status zgl_adjoint(zgl_context* ctx, const zgl_pass_params* p, const float* d_frame, float* d_material_params,
                       float* d_env_rgb)
{
    return guarded(ctx, [&] {
        if (ctx->maxTexId >= 0) invalid("the differentiable mode does not take textured materials");
        DevScene s;
        PassParams pp;
        preparePasses(ctx, p, s, pp);
        if (p->integrator != ZGL_INTEGRATOR_NAIVE) invalid("the adjoint is the naive integrator's");
        if (!d_frame || !d_material_params || !d_env_rgb) invalid("null adjoint buffer");
        const size_t px = (size_t)ctx->filmW * ctx->filmH;
        const size_t nm = (size_t)s.nMaterials * 5, ne = (size_t)s.envW * s.envH * 3;
        ctx->upload(ctx->dFrame, d_frame, px * 16, "adjoint input");
        // FP64 accumulators, zeroed as float pairs (all-zero bits are 0.0 in both).  The
        // texel gradients get up to 16 interleaved copies (at most 256 MB), summed after.
        int copies = 16;
        while (copies > 1 && (size_t)copies * ne * 8 > ((size_t)256 << 20)) copies >>= 1;
        ctx->ensure(ctx->dMat, nm * 8, "material gradients");
        ctx->ensure(ctx->dEnv, (size_t)(copies + 1) * ne * 8, "environment gradients");
        double* envCopies = (double*)ctx->dEnv.p;
        double* envSum = envCopies + (size_t)copies * ne;
        hipCheck(launchFill((float*)ctx->dMat.p, (int64_t)nm * 2, 0.0f, ctx->stream), "clear");
        hipCheck(launchFill((float*)envCopies, (int64_t)ne * 2 * copies, 0.0f, ctx->stream), "clear");
        // the chunks' gradients, differentiated from the traversal log
        loggedPasses(ctx, s, pp, p->regen_threshold, [&](const PassParams& q) {
            hipCheck(launchNaiveReplay(s, q, (const float4*)ctx->evalMat.p, (const float4*)ctx->evalEnv.p,
                                       (float4*)ctx->frame.p, (const float4*)ctx->dFrame.p, (double*)ctx->dMat.p,
                                       envCopies, copies, envSum, (const int2*)ctx->hitLog.p, ctx->stream),
                     "adjoint pass");
        });
        std::vector<double> gm(nm), ge(ne);
        hipCheck(hipMemcpyAsync(gm.data(), ctx->dMat.p, nm * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
        hipCheck(hipMemcpyAsync(ge.data(), envSum, ne * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
        hipCheck(hipStreamSynchronize(ctx->stream), "adjoint");
        for (size_t i = 0; i < nm; i++) d_material_params[i] = (float)gm[i];
        for (size_t i = 0; i < ne; i++) d_env_rgb[i] = (float)ge[i];
    });
}

zgl_status zgl_post_process(zgl_context* ctx, float result_scale, int32_t tone_mapper, float* rgba_out,
                            int64_t n_rgba, uint8_t* rgb8_out, int64_t n_rgb8)
{
    return guarded(ctx, [&] {
        if (!ctx->hasCamera) throw ZglError{ZGL_E_STATE, "no camera"};
        if (tone_mapper < 0 || tone_mapper > 2) invalid("tone_mapper must be 0 (none), 1 (filmic) or 2 (ACES)");
        const int64_t px = (int64_t)ctx->filmW * ctx->filmH;
        if (rgba_out && n_rgba < px * 4) invalid("rgba buffer too small");
        if (rgb8_out && n_rgb8 < px * 3) invalid("rgb8 buffer too small");
        ctx->ensure(ctx->scratchOut, (size_t)px * 19, "post-process output");
        float4* dOut = (float4*)ctx->scratchOut.p;
        unsigned char* dRgb = (unsigned char*)ctx->scratchOut.p + (size_t)px * 16;
        hipCheck(launchPostProcess((const float4*)ctx->frame.p, rgba_out ? dOut : nullptr, rgb8_out ? dRgb : nullptr,
                                   ctx->filmW, ctx->filmH, result_scale, tone_mapper, ctx->stream),
                 "post-process");
        if (rgba_out)
            hipCheck(hipMemcpyAsync(rgba_out, dOut, (size_t)px * 16, hipMemcpyDeviceToHost, ctx->stream), "copy");
        if (rgb8_out)
            hipCheck(hipMemcpyAsync(rgb8_out, dRgb, (size_t)px * 3, hipMemcpyDeviceToHost, ctx->stream), "copy");
        hipCheck(hipStreamSynchronize(ctx->stream), "post-process");
    });
}

zgl_status zgl_clear_frame(zgl_context* ctx)
{
    return guarded(ctx, [&] {
        if (!ctx->hasCamera) throw ZglError{ZGL_E_STATE, "no camera"};
        const int64_t px = (int64_t)ctx->filmW * ctx->filmH;
        hipCheck(launchFill((float*)ctx->frame.p, px * 4, 0.0f, ctx->stream), "clear");
        hipCheck(launchFill((float*)ctx->film.p, px * 3, 0.0f, ctx->stream), "clear");
    });
}

zgl_status zgl_read_frame(zgl_context* ctx, float* rgba, int64_t n)
{
    return guarded(ctx, [&] {
        if (!ctx->hasCamera) throw ZglError{ZGL_E_STATE, "no camera"};
        const int64_t need = (int64_t)ctx->filmW * ctx->filmH * 4;
        if (!rgba || n < need) invalid("frame buffer too small");
        hipCheck(hipMemcpyAsync(rgba, ctx->frame.p, (size_t)need * 4, hipMemcpyDeviceToHost, ctx->stream), "read");
        hipCheck(hipStreamSynchronize(ctx->stream), "read frame");
    });
}

zgl_status zgl_read_film(zgl_context* ctx, float* film, int64_t n)
{
    return guarded(ctx, [&] {
        if (!ctx->hasCamera) throw ZglError{ZGL_E_STATE, "no camera"};
        const int64_t need = (int64_t)ctx->filmW * ctx->filmH * 3;
        if (!film || n < need) invalid("film buffer too small");
        hipCheck(hipMemcpyAsync(film, ctx->film.p, (size_t)need * 4, hipMemcpyDeviceToHost, ctx->stream), "read");
        hipCheck(hipStreamSynchronize(ctx->stream), "read film");
    });
}

zgl_status zgl_write_frame(zgl_context* ctx, const float* rgba, int64_t n)
{
    return guarded(ctx, [&] {
        if (!ctx->hasCamera) throw ZglError{ZGL_E_STATE, "no camera"};
        const int64_t need = (int64_t)ctx->filmW * ctx->filmH * 4;
        if (!rgba || n != need) invalid("frame size mismatch");
        hipCheck(hipMemcpyAsync(ctx->frame.p, rgba, (size_t)need * 4, hipMemcpyHostToDevice, ctx->stream), "write");
        hipCheck(hipStreamSynchronize(ctx->stream), "write frame");
    });
}

void* zgl_frame_device_ptr(zgl_context* ctx) { return ctx ? ctx->frame.p : nullptr; }

zgl_status zgl_read_counters(zgl_context* ctx, uint64_t* out, int32_t n)
{
    return guarded(ctx, [&] {
        if (!out || n < 1 || n > NumCounters) invalid("counter buffer");
        hipCheck(hipMemcpyAsync(out, ctx->counters.p, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream), "read");
        hipCheck(hipStreamSynchronize(ctx->stream), "read counters");
    });
}

zgl_status zgl_reset_counters(zgl_context* ctx)
{
    return guarded(ctx, [&] {
        hipCheck(hipMemsetAsync(ctx->counters.p, 0, NumCounters * 8, ctx->stream), "reset counters");
        hipCheck(hipStreamSynchronize(ctx->stream), "reset counters");
    });
}

// ---------------------------------------------------------------- batched launch
zgl_status zgl_batch_create(int32_t device, zgl_batch** out)
{
    if (!out) return ZGL_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return ZGL_E_NODEVICE;
    if (device < 0 || device >= n) return ZGL_E_INVALID;
    zgl_batch* b = new (std::nothrow) zgl_batch();
    if (!b) return ZGL_E_OOM;
    b->device = device;
    zgl_status st = guarded(b, [&] {
        hipCheck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        hipCheck(hipEventCreateWithFlags(&b->fence, hipEventDisableTiming), "hipEventCreate");
        hipCheck(hipDeviceGetAttribute(&b->numCUs, hipDeviceAttributeMultiprocessorCount, device), "attribute");
        b->maxBlocksPerCU = naiveBatchBlocksPerCU();
        b->blocksPerCU = b->maxBlocksPerCU;
    });
    if (st != ZGL_OK) {
        if (b->fence) (void)hipEventDestroy(b->fence);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return st;
    }
    *out = b;
    return ZGL_OK;
}

void zgl_batch_destroy(zgl_batch* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->dropGraph();
    b->items.release();
    b->counters.release();
    b->sobolTab.release();
    b->res.release();
    if (b->fence) (void)hipEventDestroy(b->fence);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* zgl_batch_last_error(const zgl_batch* b) { return b ? b->err.c_str() : "null batch"; }

// Checks a batch description: pass ranges, depth, each scene's work-unit count and the
// total of sample values; returns the sample values' bytes and the Sobol table's range.
struct BatchPlan {
    size_t sampleBytes = 0;
    int sobolFirst = 0, sobolCount = 0;
};
static BatchPlan batchPlan(const int32_t* filmWH, const zgl_pass_params* params, int n)
{
    if (!filmWH || !params || n < 1 || n > 65536) invalid("batch: contexts / params / count");
    BatchPlan plan;
    int first = INT_MAX;
    int64_t end = 0;
    size_t total = 0;
    for (int i = 0; i < n; i++) {
        const zgl_pass_params* p = &params[i];
        const int w = filmWH[2 * i], h = filmWH[2 * i + 1];
        if (w < 1 || h < 1 || w > 16384 || h > 16384) invalid("batch: film size");
        if (p->integrator != ZGL_INTEGRATOR_NAIVE) invalid("batch: the batched launch runs the naive integrator");
        if (p->n_passes < 1 || p->n_passes > 65536 || p->first_pass < 0 || p->first_pass > (1 << 24) ||
            p->max_depth < 0 || p->max_depth > 1024)
            invalid("batch: pass range / depth");
        const int64_t units = (int64_t)((w + 7) / 8) * ((h + 7) / 8) * 64 * p->n_passes;
        if (units > kBatchMaxUnits)
            invalid("batch: a scene's (8x8 tile, pass) work units exceed 2^31 - 2^24; split its passes");
        first = std::min(first, p->first_pass);
        end = std::max(end, (int64_t)p->first_pass + p->n_passes);
        total += (size_t)p->n_passes * (size_t)w * (size_t)h;
    }
    if (total * 16 > kBatchSampleBytes) invalid("batch: sample values exceed 64 GiB; split the batch");
    plan.sampleBytes = total * 16;
    plan.sobolFirst = first;
    plan.sobolCount = (int)std::min<int64_t>(end - first + 1, kBatchSobolRows);
    return plan;
}

// Builds the device table of the batch in launch order (b->order) and drops the graph.
static void batchUpload(zgl_batch* b)
{
    const int n = (int)b->ctxs.size();
    b->host.resize(n);
    b->generations.resize(n);
    for (int i = 0; i < n; i++) b->generations[i] = b->ctxs[i]->generation;
    for (int k = 0; k < n; k++) {
        const int i = b->order[k];
        zgl_context* ctx = b->ctxs[i];
        const zgl_pass_params* p = &b->params[i];
        BatchScene& it = b->host[k];
        it.s = ctx->scene;
        it.threshold = b->threshold > 0 ? b->threshold : defaultThreshold(WalkKind::Naive, it.s.bvhSize);
        it.s.sobolTab = (const uint32_t*)b->sobolTab.p;
        it.s.sobolTabFirst = b->sobolFirst;
        it.s.sobolTabCount = b->sobolCount;
        it.prm = PassParams{};
        it.prm.firstPass = p->first_pass;
        it.prm.nPasses = p->n_passes;
        it.prm.freeCounter = p->free_counter;
        it.prm.maxDepth = p->max_depth;
        it.prm.russianRoulette = p->russian_roulette;
        it.prm.sampleLight = p->sample_light;
        it.prm.lightEnvUniform = p->light_env_uniform;
        it.prm.lightPortion = p->light_portion;
        it.prm.loopsPerPass = 1;
        it.prm.pathsPerPass = 1536;
        it.prm.lptScale = 0.0f;
        it.frame = (float4*)ctx->frame.p;
        it.res = (float4*)b->res.p + b->resOffset[i];
    }
    b->dropGraph();
    hipCheck(hipStreamSynchronize(b->stream), "batch");
    b->upload(b->items, b->host.data(), sizeof(BatchScene) * n, "batch table");
    hipCheck(hipStreamSynchronize(b->stream), "batch table");
}

zgl_status zgl_batch_set(zgl_batch* b, zgl_context* const* ctxs, const zgl_pass_params* params, int32_t n)
{
    return guarded(b, [&] {
        if (!ctxs || !params || n < 1 || n > 65536) invalid("batch: contexts / params / count");
        std::vector<int32_t> wh(2 * (size_t)n);
        for (int i = 0; i < n; i++) {
            zgl_context* ctx = ctxs[i];
            if (!ctx) invalid("batch: null context");
            if (ctx->device != b->device) invalid("batch: context on another device");
            if (!ctx->hasScene || !ctx->hasCamera || !ctx->hasNoise)
                throw ZglError{ZGL_E_STATE, "batch: scene, camera and noise must be set in every context"};
            checkTextures(ctx);
            for (int j = 0; j < i; j++)
                if (ctxs[j] == ctx) invalid("batch: a context appears twice");
            wh[2 * i] = ctx->filmW;
            wh[2 * i + 1] = ctx->filmH;
        }
        const BatchPlan plan = batchPlan(wh.data(), params, n);
        b->ctxs.assign(ctxs, ctxs + n);
        b->params.assign(params, params + n);
        b->order.resize(n);
        for (int i = 0; i < n; i++) b->order[i] = i;
        // one Sobol table for the union of the pass ranges (plus one for dims past 256),
        // at most kBatchSobolRows rows
        b->sobolFirst = plan.sobolFirst;
        b->sobolCount = plan.sobolCount;
        b->ensure(b->sobolTab, (size_t)b->sobolCount * 256 * 4, "sobol table");
        hipCheck(launchSobolTable((const uint32_t*)ctxs[0]->sobolM.p, (uint32_t*)b->sobolTab.p, b->sobolFirst,
                                  b->sobolCount, b->stream),
                 "sobol table");
        b->ensure(b->counters, (size_t)n * kUnitRegions * 4, "batch counters");
        // every scene's sample values (n_passes x W x H float4), added to its frame in
        // pass order after the persistent kernel
        b->resOffset.resize(n);
        size_t total = 0;
        b->maxPixels = 0;
        for (int i = 0; i < n; i++) {
            const int pixels = ctxs[i]->filmW * ctxs[i]->filmH;
            b->resOffset[i] = total;
            total += (size_t)params[i].n_passes * pixels;
            b->maxPixels = std::max(b->maxPixels, pixels);
        }
        b->ensure(b->res, total * 16, "batch sample values");
        batchUpload(b);
    });
}

zgl_status zgl_batch_configure(zgl_batch* b, int32_t regen_threshold, int32_t blocks_per_cu, const int32_t* order)
{
    return guarded(b, [&] {
        const int n = (int)b->ctxs.size();
        if (n == 0) throw ZglError{ZGL_E_STATE, "batch: no scenes set"};
        if (regen_threshold < 0 || regen_threshold > 63) invalid("regen_threshold must be in [0, 63]");
        if (blocks_per_cu < 0 || blocks_per_cu > b->maxBlocksPerCU) invalid("blocks_per_cu out of range");
        if (order) {
            std::vector<char> seen(n, 0);
            for (int k = 0; k < n; k++) {
                if (order[k] < 0 || order[k] >= n || seen[order[k]]) invalid("order must be a permutation");
                seen[order[k]] = 1;
            }
            b->order.assign(order, order + n);
        }
        b->threshold = regen_threshold;
        b->blocksPerCU = blocks_per_cu > 0 ? blocks_per_cu : b->maxBlocksPerCU;
        batchUpload(b);
    });
}

zgl_status zgl_batch_plan(const int32_t* film_wh, const zgl_pass_params* params, int32_t n, int64_t* sample_bytes,
                          int32_t* sobol_first, int32_t* sobol_count)
{
    return guardedHost([&] {
        const BatchPlan plan = batchPlan(film_wh, params, n);
        if (sample_bytes) *sample_bytes = (int64_t)plan.sampleBytes;
        if (sobol_first) *sobol_first = plan.sobolFirst;
        if (sobol_count) *sobol_count = plan.sobolCount;
    });
}

zgl_status zgl_batch_launch(zgl_batch* b)
{
    return guarded(b, [&] {
        const int n = (int)b->ctxs.size();
        if (n == 0) throw ZglError{ZGL_E_STATE, "batch: no scenes set"};
        {
            std::lock_guard<std::mutex> lock(gLiveMutex);
            for (int i = 0; i < n; i++)
                if (!gLive.count(b->ctxs[i]) || b->ctxs[i]->generation != b->generations[i])
                    throw ZglError{ZGL_E_STATE, "batch: a context was destroyed or its scene, camera or noise "
                                                "changed since zgl_batch_set; call zgl_batch_set again"};
        }
        // after the work already enqueued on the contexts' streams (a frame clear, a
        // restored frame) ...
        for (zgl_context* ctx : b->ctxs) {
            hipCheck(hipEventRecord(b->fence, ctx->stream), "batch order");
            hipCheck(hipStreamWaitEvent(b->stream, b->fence, 0), "batch order");
        }
        if (!b->exec) {
            // the whole launch (counter reset, persistent kernel, accumulation) as one graph
            hipCheck(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "capture");
            const hipError_t e = launchNaiveBatch((const BatchScene*)b->items.p, n, b->maxPixels, (int*)b->counters.p,
                                                  b->numCUs * b->blocksPerCU, b->stream);
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            hipCheck(e, "batch capture");
            hipCheck(e2, "batch capture");
            b->graph = g;
            hipCheck(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
        }
        hipCheck(hipGraphLaunch(b->exec, b->stream), "graph launch");
        // ... and before what the contexts enqueue next (a frame read waits for the batch)
        hipCheck(hipEventRecord(b->fence, b->stream), "batch order");
        for (zgl_context* ctx : b->ctxs) hipCheck(hipStreamWaitEvent(ctx->stream, b->fence, 0), "batch order");
    });
}

zgl_status zgl_batch_synchronize(zgl_batch* b)
{
    return guarded(b, [&] { hipCheck(hipStreamSynchronize(b->stream), "batch"); });
}

// Median time of `reps` graph launches after one warm-up launch, in ms.
static double batchTime(zgl_batch* b, int reps)
{
    hipEvent_t e0, e1;
    hipCheck(hipEventCreate(&e0), "event");
    hipCheck(hipEventCreate(&e1), "event");
    std::vector<double> ms;
    for (int r = 0; r <= reps; r++) {
        hipCheck(hipEventRecord(e0, b->stream), "event");
        if (zgl_batch_launch(b) != ZGL_OK) throw ZglError{ZGL_E_HIP, b->err};
        hipCheck(hipEventRecord(e1, b->stream), "event");
        hipCheck(hipEventSynchronize(e1), "event");
        float t = 0.0f;
        hipCheck(hipEventElapsedTime(&t, e0, e1), "event");
        if (r > 0) ms.push_back(t);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
}

zgl_status zgl_batch_tune(zgl_batch* b, int32_t reps, double* best_ms)
{
    return guarded(b, [&] {
        const int n = (int)b->ctxs.size();
        if (n == 0) throw ZglError{ZGL_E_STATE, "batch: no scenes set"};
        if (reps < 1 || reps > 100) invalid("reps must be in [1, 100]");
        auto config = [&](int th, int bpc, const std::vector<int32_t>& ord) {
            if (zgl_batch_configure(b, th, bpc, ord.data()) != ZGL_OK) throw ZglError{ZGL_E_INVALID, b->err};
        };
        // candidate scene orders: as given, and most work first (pixels x passes x depth)
        std::vector<int32_t> given(n), heavy(n);
        for (int i = 0; i < n; i++) given[i] = heavy[i] = i;
        auto work = [&](int i) {
            const zgl_context* c = b->ctxs[i];
            return (double)c->filmW * c->filmH * b->params[i].n_passes * (b->params[i].max_depth + 1);
        };
        std::stable_sort(heavy.begin(), heavy.end(), [&](int a, int c) { return work(a) > work(c); });
        int th = b->threshold, bpc = b->blocksPerCU;
        std::vector<int32_t> ord = b->order;
        double best = 1e30;
        for (const auto& o : {given, heavy}) {
            config(th, bpc, o);
            const double t = batchTime(b, reps);
            if (t < best) { best = t; ord = o; }
        }
        for (int t : {0, 8, 12, 16, 20, 24, 32}) {  // 0: each scene's own default
            config(t, bpc, ord);
            const double tm = batchTime(b, reps);
            if (tm < best) { best = tm; th = t; }
        }
        for (int c = b->maxBlocksPerCU; c >= std::max(1, b->maxBlocksPerCU - 1); c--) {
            config(th, c, ord);
            const double tm = batchTime(b, reps);
            if (tm < best) { best = tm; bpc = c; }
        }
        config(th, bpc, ord);
        if (best_ms) *best_ms = best;
    });
}

zgl_status zgl_batch_get_config(const zgl_batch* b, int32_t* regen_threshold, int32_t* blocks_per_cu,
                                int32_t* order, int32_t max_blocks_per_cu[1])
{
    if (!b) return ZGL_E_INVALID;
    if (regen_threshold) *regen_threshold = b->threshold;
    if (blocks_per_cu) *blocks_per_cu = b->blocksPerCU;
    if (order) std::memcpy(order, b->order.data(), b->order.size() * 4);
    if (max_blocks_per_cu) max_blocks_per_cu[0] = b->maxBlocksPerCU;
    return ZGL_OK;
}

// ---------------------------------------------------------------- device MTBVH build
}  // extern "C"

struct zgl_bvh_builder {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf vertices, indices, bounds, hit;
    int64_t nV = 0, nT = 0;
    bool built = false;
    DeviceBvh bvh;
    ~zgl_bvh_builder()
    {
        DevBuf* all[] = {&vertices, &indices, &bounds, &hit};
        for (DevBuf* b : all) b->release();
        bvh.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

zgl_status zgl_bvh_builder_create(int32_t device, zgl_bvh_builder** out)
{
    if (!out) return ZGL_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return ZGL_E_NODEVICE;
    if (device < 0 || device >= n) return ZGL_E_INVALID;
    zgl_bvh_builder* b = new (std::nothrow) zgl_bvh_builder;
    if (!b) return ZGL_E_OOM;
    b->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return ZGL_E_HIP;
    }
    *out = b;
    return ZGL_OK;
}

void zgl_bvh_builder_destroy(zgl_bvh_builder* b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    delete b;
}

const char* zgl_bvh_builder_last_error(const zgl_bvh_builder* b) { return b ? b->err.c_str() : "null builder"; }

zgl_status zgl_bvh_builder_upload(zgl_bvh_builder* b, const float* vertices, int64_t n_vertices,
                                  const uint32_t* indices, int64_t n_triangles)
{
    return guarded(b, [&] {
        if (!vertices || !indices || n_vertices < 1 || n_triangles < 1) invalid("empty mesh");
        if (n_triangles > (1LL << 27) || n_vertices > (1LL << 30)) invalid("mesh too large");
        bufUpload(b->vertices, vertices, (size_t)n_vertices * 12, b->stream, "vertices");
        bufUpload(b->indices, indices, (size_t)n_triangles * 12, b->stream, "indices");
        hipCheck(hipStreamSynchronize(b->stream), "mesh upload");
        b->nV = n_vertices;
        b->nT = n_triangles;
        b->built = false;
    });
}

zgl_status zgl_bvh_builder_build(zgl_bvh_builder* b, int32_t subtree_size)
{
    return guarded(b, [&] {
        if (b->nT < 1) throw ZglError{ZGL_E_STATE, "build before upload"};
        if (subtree_size < 0) invalid("subtree_size");
        const int64_t B = 2 * b->nT - 1;
        bufEnsure(b->bounds, (size_t)B * 24, "MTBVH bounds");
        bufEnsure(b->hit, (size_t)B * 72, "hit tables");
        b->built = false;
        std::string msg;
        const int rc = b->bvh.build((const float*)b->vertices.p, b->nV, (const uint32_t*)b->indices.p, b->nT,
                                    subtree_size ? subtree_size : kDefaultBvhSubtree, (float*)b->bounds.p,
                                    (int32_t*)b->hit.p, nullptr, b->stream, msg);
        if (rc) throw ZglError{rc == 1 ? ZGL_E_INVALID : (rc == 3 ? ZGL_E_OOM : ZGL_E_HIP), msg};
        hipCheck(hipStreamSynchronize(b->stream), "MTBVH build");
        b->built = true;
    });
}

zgl_status zgl_bvh_builder_download(zgl_bvh_builder* b, float* bounds_out, int32_t* hit_table_out)
{
    return guarded(b, [&] {
        if (!b->built) throw ZglError{ZGL_E_STATE, "download before a successful build"};
        const int64_t B = 2 * b->nT - 1;
        if (bounds_out)
            hipCheck(hipMemcpyAsync(bounds_out, b->bounds.p, (size_t)B * 24, hipMemcpyDeviceToHost, b->stream),
                     "download bounds");
        if (hit_table_out)
            hipCheck(hipMemcpyAsync(hit_table_out, b->hit.p, (size_t)B * 72, hipMemcpyDeviceToHost, b->stream),
                     "download hit tables");
        hipCheck(hipStreamSynchronize(b->stream), "download");
    });
}

zgl_status zgl_bvh_builder_stats(const zgl_bvh_builder* b, int64_t stats_out[5])
{
    if (!b || !stats_out) return ZGL_E_INVALID;
    const DeviceBvhStats& s = b->bvh.stats;
    stats_out[0] = s.levels;
    stats_out[1] = s.bigNodes;
    stats_out[2] = s.smallTasks;
    stats_out[3] = s.maxDepth;
    stats_out[4] = s.jumps;
    return ZGL_OK;
}

zgl_status zgl_build_mtbvh_device(int32_t device, const float* vertices, int64_t n_vertices,
                                  const uint32_t* indices, int64_t n_triangles, float* bounds_out,
                                  int32_t* hit_table_out)
{
    zgl_bvh_builder* b = nullptr;
    zgl_status st = zgl_bvh_builder_create(device, &b);
    if (st != ZGL_OK) return st;
    st = zgl_bvh_builder_upload(b, vertices, n_vertices, indices, n_triangles);
    if (st == ZGL_OK) st = zgl_bvh_builder_build(b, 0);
    if (st == ZGL_OK) st = zgl_bvh_builder_download(b, bounds_out, hit_table_out);
    zgl_bvh_builder_destroy(b);
    return st;
}

}  // extern "C"
