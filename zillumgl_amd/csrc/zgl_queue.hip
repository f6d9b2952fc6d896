// zgl_queue.hip -- the global-queue (wavefront) path tracer.
//
// GlobalQueuePathIntegrator (src/integrator/GlobalQueuePath.cpp:100-155) renders a pass
// as a primary dispatch that pushes every surface hit into a queue
// (pt_global_queue_primary.glsl:40-86), then one streaming dispatch per depth that pops
// the queue, shades the vertex, traces the extension ray and pushes the next vertex
// (pt_global_queue_streaming.glsl:89-208).  Its estimator differs from the naive path
// tracer: the vertex of depth d takes Sobol dims 4 + 9 d onwards (the NEE selector and
// four light dims are drawn whether or not NEE is on), the pixel's seeds restart at every
// depth, and each contribution is added to the frame when it is produced, without the
// naive tracer's NaN test.
//
// Schedule here.  The reference reads the queue's head and tail back to the host before
// every depth and sizes the dispatch from them.  This version keeps the counts on the
// device: each depth has its own compacted queue (two buffers of W*H items used in turn)
// and its own counter, the extension kernel runs an occupancy-sized grid that reads its
// input count itself, and pushes are wave-aggregated (one atomic per wave).  The host
// enqueues passes x (depth + 2) launches with no synchronisation.  A pixel has at most
// one item per depth, so the read-modify-write of its frame texel is race-free and the
// frame does not depend on the order in which items land in a queue.
#include "zgl_internal.h"

namespace zgl {

namespace {

constexpr int kQBlock = 256;

struct __align__(16) QueueItem {
    float4 posWox;   // position, wo.x
    float4 tputWoy;  // throughput, wo.y
    int4 pixWozId;   // pixel index, wo.z bits, primitive id, unused
};

__device__ __forceinline__ void pixelOfTileIndex(int idx, int W, int& x, int& y)
{
    const int tilesX = (W + 7) / 8;
    const int tile = idx >> 6, in = idx & 63;
    x = (tile % tilesX) * 8 + (in & 7);
    y = (tile / tilesX) * 8 + (in >> 3);
}

// accumulateImage (pt_global_queue_primary.glsl:24-28): rgb += c, alpha = 1.
__device__ __forceinline__ void accumulate(float4* frame, int pix, vec3 c)
{
    const float4 last = frame[pix];
    frame[pix] = make_float4(last.x + c.x, last.y + c.y, last.z + c.z, 1.0f);
}

// All lanes of the wave call this; lanes with push append their item.  One atomic per wave.
__device__ __forceinline__ void pushItem(QueueItem* q, int* count, bool push, const QueueItem& it)
{
    const unsigned long long m = __ballot(push);
    if (m == 0ull) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader, 64);
    if (push) q[base + __popcll(m & ((1ull << lane) - 1ull))] = it;
}

__device__ __forceinline__ Rng seedPixel(const DevScene& s, int x, int y, int spp, uint32_t fc)
{
    const vec2 scr = v2((float)x, (float)y) / v2((float)s.filmW, (float)s.filmH);
    const PixelSeed ps = pixelSeed(s, scr);
    Rng r;
    r.randSeed = ps.tx * fc + ps.ty;
    r.sampleSeed = ps.tx * ps.ty;
    r.sampleOffset = sampleOffsetOf(spp);
    r.dim = 0;
    return r;
}

__device__ __forceinline__ QueueItem makeItem(vec3 pos, vec3 tp, vec3 wo, int pix, int id)
{
    QueueItem it;
    it.posWox = make_float4(pos.x, pos.y, pos.z, wo.x);
    it.tputWoy = make_float4(tp.x, tp.y, tp.z, wo.y);
    it.pixWozId = make_int4(pix, __float_as_int(wo.z), id, 0);
    return it;
}

__global__ __launch_bounds__(kQBlock) void k_gq_reset(int* counts, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) counts[i] = 0;
}

// One thread per pixel, pixels in 8x8 tiles.
__global__ __launch_bounds__(kQBlock) void k_gq_primary(DevScene s, int spp, uint32_t fc, float4* __restrict__ frame,
                                                        QueueItem* __restrict__ q, int* __restrict__ count)
{
    const int W = s.filmW, H = s.filmH;
    int x, y;
    pixelOfTileIndex(blockIdx.x * blockDim.x + threadIdx.x, W, x, y);
    bool push = false;
    QueueItem it = makeItem(v3(0.0f), v3(0.0f), v3(0.0f), 0, 0);
    if (x < W && y < H) {
        Counters cn{};
        const int pix = y * W + x;
        Rng rng = seedPixel(s, x, y, spp, fc);
        const vec2 scr = v2((float)x, (float)y) / v2((float)W, (float)H);
        const Ray ray = cameraSampleRay(s, scr, sample4D<false>(s, rng, cn));
        float dist;
        const int id = bvhHit<false>(s, ray, dist, cn);
        const vec3 pos = rayPoint(ray, dist);
        if (id == -1) {
            accumulate(frame, pix, envLe<false>(s, ray.dir, cn));
        } else if (id - s.objPrimCount >= 0) {
            accumulate(frame, pix, lightLe<false>(s, id - s.objPrimCount, pos, -ray.dir, cn));
        } else {
            push = true;
            it = makeItem(pos, v3(1.0f), -ray.dir, pix, id);
        }
    }
    pushItem(q, count, push, it);
}

// extendPath (pt_global_queue_streaming.glsl:89-180) for one item; returns whether the
// path continues (out = the next vertex).
__device__ bool extendPath(const DevScene& s, const PassParams& prm, const QueueItem& item, int depth, int spp,
                           uint32_t fc, float4* frame, QueueItem& out)
{
    Counters cn{};
    const int W = s.filmW;
    vec3 pos = v3(item.posWox.x, item.posWox.y, item.posWox.z);
    vec3 throughput = v3(item.tputWoy.x, item.tputWoy.y, item.tputWoy.z);
    const vec3 wo = v3(item.posWox.w, item.tputWoy.w, __int_as_float(item.pixWozId.y));
    const int pix = item.pixWozId.x;
    const int id = item.pixWozId.z;
    Rng rng = seedPixel(s, pix % W, pix / W, spp, fc);
    rng.dim = 4 + 9 * depth;
    const bool uniform = prm.lightEnvUniform != 0;
    const float portion = prm.lightPortion;

    SurfaceInfo surf = triangleSurfaceInfo<false>(s, id, pos, cn);
    const int matTex = s.matTexIndices[id];
    const int matId = matTex & 0x0000ffff;
    const uint32_t matType = loadMaterialType<false>(s, matId, cn);
    if (matType != MatDielectric && matType != MatThinDielectric) {
        if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
    }
    const BSDFParam mp = loadMaterial<false>(s, matType, matId, matTex >> 16, surf.uv, cn);

    const float ud = sample1D<false>(s, rng, cn);
    const vec4 us = sample4D<false>(s, rng, cn);
    if (prm.sampleLight) {
        const LightLiSample ls = sampleLightAndEnv<false>(s, pos, ud, us, uniform, portion, cn);
        if (ls.pdf > 0.0f) {
            const vec3 f = materialBSDF(matType, mp, wo, ls.wi, surf.ns, ModeRadiance);
            const float fp = materialPdf(matType, mp, wo, ls.wi, surf.ns);
            const float weight = ls.pdf * ls.pdf / (ls.pdf * ls.pdf + fp * fp);
            accumulate(frame, pix, f * throughput * satDot(surf.ns, ls.wi) * ls.coef * weight);
        }
    }

    const vec3 u3 = sample3D<false>(s, rng, cn);
    const BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeRadiance, u3, rng);
    const vec3 wi = bs.wi;
    const float bsdfPdf = bs.pdf;
    const bool delta = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
    if (bsdfPdf < 1e-8f) return false;
    throughput *= bs.bsdf / bsdfPdf * (delta ? 1.0f : absDot(surf.ns, wi));

    const Ray ray = rayOffseted(pos, wi);
    float dist;
    const int nextId = bvhHit<false>(s, ray, dist, cn);
    const int lightId = nextId - s.objPrimCount;
    const vec3 nextPos = rayPoint(ray, dist);
    if (nextId == -1) {
        const vec3 radiance = envLe<false>(s, wi, cn);
        float weight = 1.0f;
        if (prm.sampleLight && !delta) {
            const float envPdf = envPdfLi<false>(s, wi, cn) * pdfSelectEnv(s, uniform, portion);
            weight = (envPdf <= 0.0f) ? 0.0f : bsdfPdf * bsdfPdf / (bsdfPdf * bsdfPdf + envPdf * envPdf);
        }
        accumulate(frame, pix, radiance * throughput * weight);
        return false;
    }
    if (lightId >= 0) {
        const vec3 radiance = lightLe<false>(s, lightId, nextPos, -wi, cn);
        float weight = 1.0f;
        if (prm.sampleLight && !delta) {
            const float lightPdf =
                lightPdfLi<false>(s, lightId, pos, nextPos, cn) * pdfSelectLight<false>(s, lightId, uniform, portion, cn);
            weight = (lightPdf <= 0.0f) ? 0.0f : bsdfPdf * bsdfPdf / (bsdfPdf * bsdfPdf + lightPdf * lightPdf);
        }
        accumulate(frame, pix, radiance * throughput * weight);
        return false;
    }
    // Russian roulette as selects: the survival division is always evaluated (the same
    // bits where it is kept) and the next item is always formed.  Written as a branch
    // (`if (u >= p) return false; throughput /= p;`), ROCm 7.2's compiler stored a stale
    // wo.z into the surviving paths' items (the register holding -wi.z was written only
    // on the path without roulette), which the oracle comparisons caught as NaNs.
    bool survive = true;
    if (prm.russianRoulette) {
        const float continueProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 0.95f);
        const float u = sample1D<false>(s, rng, cn);
        survive = !(u >= continueProb);
        const vec3 scaled = throughput / continueProb;
        throughput = select3(survive, scaled, throughput);
    }
    out = makeItem(nextPos, throughput, -wi, pix, nextId);
    return survive;
}

// Occupancy-sized grid; the block-uniform loop keeps whole waves together for the
// ballot in pushItem.
__global__ __launch_bounds__(kQBlock) void k_gq_extend(DevScene s, PassParams prm, int depth, int spp, uint32_t fc,
                                                       float4* __restrict__ frame, const QueueItem* __restrict__ qin,
                                                       const int* __restrict__ countIn, QueueItem* __restrict__ qout,
                                                       int* __restrict__ countOut)
{
    const int n = *countIn;
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        bool push = false;
        // initialised on every path: with `out` left unset where extendPath returns
        // false, the compiler dropped the item's wo on the roulette's surviving path
        QueueItem out = makeItem(v3(0.0f), v3(0.0f), v3(0.0f), 0, 0);
        if (i < n) push = extendPath(s, prm, qin[i], depth, spp, fc, frame, out);
        pushItem(qout, countOut, push, out);
    }
}

}  // namespace

size_t globalQueueBytes(int w, int h) { return (size_t)2 * w * h * sizeof(QueueItem); }

int globalQueueBlocks()
{
    int perCU = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_gq_extend, kQBlock, 0) != hipSuccess) perCU = 0;
    return residentBlocksPerCU(reinterpret_cast<const void*>(k_gq_extend), kQBlock, perCU);
}

hipError_t launchGlobalQueue(const DevScene& s, const PassParams& p, float4* frame, void* queues, int* counts,
                             int extendBlocks, hipStream_t st)
{
    QueueItem* q[2] = {static_cast<QueueItem*>(queues),
                       static_cast<QueueItem*>(queues) + (size_t)s.filmW * s.filmH};
    const int nIndex = ((s.filmW + 7) / 8) * ((s.filmH + 7) / 8) * 64;
    const int nCounts = p.maxDepth + 1;
    for (int k = 0; k < p.nPasses; k++) {
        const int spp = p.firstPass + k;
        const uint32_t fc = (uint32_t)(p.freeCounter + k);
        k_gq_reset<<<1, kQBlock, 0, st>>>(counts, nCounts);
        k_gq_primary<<<(nIndex + kQBlock - 1) / kQBlock, kQBlock, 0, st>>>(s, spp, fc, frame, q[0], counts);
        for (int d = 1; d <= p.maxDepth; d++)
            k_gq_extend<<<extendBlocks, kQBlock, 0, st>>>(s, p, d, spp, fc, frame, q[(d - 1) & 1], counts + d - 1,
                                                          q[d & 1], counts + d);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace zgl
