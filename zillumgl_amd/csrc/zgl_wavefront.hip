// zgl_wavefront.hip -- the naive MIS path tracer (path_integ_naive.glsl:35-174) on a
// wavefront schedule: the same estimator and the same frame, bit for bit, as
// k_naive_persistent, with traversal and shading in separate kernels.
//
// Why.  In the persistent kernel a lane whose MTBVH walk has ended can only get a new
// ray from its own shading step, so it idles until its wave leaves the walk loop (on the
// 1080p workload ~30 of 64 lanes walk per round).  Here the walker takes rays from a
// queue: a lane that finishes a walk takes the next ray at once, and shading runs on
// compacted lists of paths, one phase at a time.
//
// A pass is rendered as
//   k_wf_camera   one path per pixel: seeds, camera ray -> ray queue 0, path list 0;
//   for it = 0 .. maxDepth:
//     k_wf_trace  every ray of queue it (closest-hit or shadow), results into the paths;
//     k_wf_shade  every path of list it: settle the shadow ray's NEE term, the closest
//                 hit (env / light / surface, Russian roulette), shade the next vertex
//                 (NEE set-up, BSDF sample) and queue its shadow and extension rays into
//                 queue it + 1, or finish the path into the frame.
// Queues and lists are compacted with one atomic per wave; the counts stay on the device
// and every kernel reads its own, so a pass is 2 (maxDepth + 2) launches with no host
// synchronisation.  Each pixel has one path per pass and passes run in order, so the
// frame receives the persistent kernel's sums in the same order.
#include "zgl_internal.h"

#include <algorithm>

namespace zgl {

namespace {

constexpr int kWfBlock = 256;
constexpr int kWfRows = 4;      // MTBVH rows per walk round (as the persistent kernel)
constexpr int kWfChunk = 256;   // rays a wave takes from the queue per atomic
constexpr int kWfRefill = 8;    // idle lanes that make a wave take new rays

// path flags
constexpr int fBounceMask = 0xffff;
constexpr int fPendValid = 1 << 16;
constexpr int fDelta = 1 << 17;
constexpr int fTerminate = 1 << 18;
constexpr int fPrimary = 1 << 19;  // the pending closest-hit ray is the camera ray
constexpr int fShadow = 1 << 20;   // a shadow ray is pending
constexpr int fExtend = 1 << 21;   // an extension ray is pending

// Path state, one slot per pixel, structure of arrays.
struct WfState {
    float4* a;  // result.xyz, throughput.x
    float4* b;  // throughput.yz, pos.xy
    float4* c;  // pos.z, wo.xyz
    float4* d;  // pend.xyz, bsdfPdf
    float4* e;  // wi.xyz, rrProb
    int4* f;    // randSeed, sampleSeed, dim, flags
    int* hitId;
    float* hitDist;
    int* occluded;
    float4* res;  // per path: the sample's value when it ends (accumulated in pass order)
    int N;        // pixels; path = passInBatch * N + pixel
    int firstSpp; // spp of the batch's first pass
    uint32_t firstFc;
};

struct __align__(16) WfRay {
    float4 oT;  // origin, tmax
    float4 dP;  // direction, (path << 1 | shadow) bits
};

struct WfQueues {
    int* list[2];
    WfRay* rays[2];
    int* counts;  // [it] paths of list it, [L + it] rays of queue it, [2L + it] trace fetch cursor
    int L;        // maxDepth + 2
};

__device__ __forceinline__ void pixelOfTile(int idx, int W, int& x, int& y)
{
    const int tilesX = (W + 7) / 8;
    const int tile = idx >> 6, in = idx & 63;
    x = (tile % tilesX) * 8 + (in & 7);
    y = (tile / tilesX) * 8 + (in >> 3);
}

// All threads of the block call this (256 threads, 4 waves); one atomic per block.
template <class T> __device__ __forceinline__ void pushBlock(T* q, int* count, bool push, const T& it, int* sh)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long m = __ballot(push);
    if (lane == 0) sh[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = sh[0] + sh[1] + sh[2] + sh[3];
        sh[4] = tot ? atomicAdd(count, tot) : 0;
    }
    __syncthreads();
    int base = sh[4];
    for (int k = 0; k < w; k++) base += sh[k];
    if (push) q[base + __popcll(m & ((1ull << lane) - 1ull))] = it;
    __syncthreads();
}
// All lanes of the wave call these; one atomic per wave.
template <class T> __device__ __forceinline__ void pushWave(T* q, int* count, bool push, const T& it)
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

__device__ __forceinline__ WfRay makeRay(const Ray& r, float tmax, int path, bool shadow)
{
    WfRay w;
    w.oT = make_float4(r.ori.x, r.ori.y, r.ori.z, tmax);
    w.dP = make_float4(r.dir.x, r.dir.y, r.dir.z, __int_as_float((path << 1) | (shadow ? 1 : 0)));
    return w;
}

// Zero the counts of a pass; iteration 0 holds every pixel's path and camera ray.
__global__ __launch_bounds__(kWfBlock) void k_wf_reset(int* counts, int n, int L, int pixels)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) counts[i] = (i == 0 || i == L) ? pixels : 0;
}

// ---------------------------------------------------------------- camera
__global__ __launch_bounds__(kWfBlock) void k_wf_camera(DevScene s, int nIndex, WfState st, WfQueues q)
{
    const int W = s.filmW, H = s.filmH;
    const int pl = blockIdx.y;  // pass within the batch
    const int spp = st.firstSpp + pl;
    const uint32_t fc = st.firstFc + (uint32_t)pl;
    int x, y;
    pixelOfTile(blockIdx.x * blockDim.x + threadIdx.x, W, x, y);
    const bool valid = x < W && y < H;
    int path = 0;
    WfRay r;
    if (valid) {
        Counters cn;
        path = pl * st.N + y * W + x;
        const vec2 scr = v2((float)x, (float)y) / v2((float)W, (float)H);
        const PixelSeed ps = pixelSeed(s, scr);
        Rng rng;
        rng.randSeed = ps.tx * fc + ps.ty;
        rng.sampleSeed = ps.tx * ps.ty;
        rng.sampleOffset = sampleOffsetOf(spp);
        rng.dim = 0;
        const vec4 u = sample4D<false>(s, rng, cn);
        const Ray ray = cameraSampleRay(s, scr, u);
        // pos / wi hold the camera ray until its hit is shaded
        st.b[path] = make_float4(0.f, 0.f, ray.ori.x, ray.ori.y);
        st.c[path] = make_float4(ray.ori.z, 0.f, 0.f, 0.f);
        st.e[path] = make_float4(ray.dir.x, ray.dir.y, ray.dir.z, 0.f);
        st.f[path] = make_int4((int)rng.randSeed, (int)rng.sampleSeed, rng.dim, fPrimary);
        r = makeRay(ray, 1e8f, path, false);
    }
    // every pixel starts a path: list and queue slots by pixel, no atomics
    if (valid) {
        q.list[0][path] = path;
        q.rays[0][path] = r;
    }
}

// ---------------------------------------------------------------- trace
// Persistent walker: lanes take rays from the queue (a wave claims kWfChunk at a time)
// whenever kWfRefill of them are idle or none walks, and walk them with the persistent
// kernel's rounds (kWfRows rows, then one leaf test for the lanes at a leaf).
__global__ __launch_bounds__(kWfBlock) void k_wf_trace(DevScene s, const WfRay* __restrict__ rays,
                                                       const int* __restrict__ count, int* __restrict__ fetch,
                                                       WfState st)
{
    const int n = *count;
    const int lane = threadIdx.x & 63;
    Counters cn;
    Walk wk;
    wk.p.o = v3(0.f); wk.p.d = v3(0.f); wk.p.dInv = v3(0.f); wk.p.mode = 0;
    wk.p.nx = wk.p.ny = wk.p.nz = false; wk.p.special = false; wk.p.table = 0;
    wk.dist = 0.f; wk.k = 0; wk.end = 0; wk.closest = -1; wk.leaf = -1; wk.any = false;
    bool walking = false;
    bool own = false;  // the lane holds a ray whose result is still to be stored
    int tag = 0;
    // each wave walks a fixed slice of the queue (no shared cursor)
    const int waves = gridDim.x * (blockDim.x >> 6);
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int per = (n + waves - 1) / waves;
    int cur = min(n, wave * per);
    const int end = min(n, cur + per);
    (void)fetch;
    for (;;) {
        const unsigned long long idle = __ballot(!walking);
        const int nIdle = __popcll(idle);
        const bool drained = cur >= end;
        if (!drained && (nIdle >= kWfRefill || nIdle == 64)) {
            {
                const int take = min(nIdle, end - cur);
                const int rank = __popcll(idle & ((1ull << lane) - 1ull));
                if (!walking && rank < take) {
                    const WfRay r = rays[cur + rank];
                    tag = __float_as_int(r.dP.w);
                    Ray ray;
                    ray.ori = v3(r.oT.x, r.oT.y, r.oT.z);
                    ray.dir = v3(r.dP.x, r.dP.y, r.dP.z);
                    walkStart<false>(s, wk, ray, (tag & 1) != 0, r.oT.w, cn);
                    walking = true;
                    own = true;
                }
                cur += take;
            }
        }
        if (__ballot(walking) == 0ull) {
            if (drained) break;
            continue;
        }
        const bool general = __ballot(walking && wk.p.special) == 0ull;
        const WalkMem wm = walkMem(s);
#pragma unroll
        for (int u = 0; u < kWfRows; ++u) {
            if (walking && wk.leaf < 0) {
                if (walkRow<false>(wm, wk, general, cn)) walking = false;
            }
        }
        if (walking && wk.leaf >= 0) {
            if (walkLeaf<false>(wm, wk, cn)) walking = false;
        }
        // a walk that ended stores its result in its path
        if (!walking && own) {
            const int path = tag >> 1;
            if (tag & 1) {
                st.occluded[path] = wk.closest >= 0 ? 1 : 0;
            } else {
                st.hitId[path] = wk.closest;
                st.hitDist[path] = wk.dist;
            }
            own = false;
        }
    }
}

// ---------------------------------------------------------------- shade
__device__ __forceinline__ void finishPath(const WfState& st, int path, vec3 result)
{
    st.res[path] = make_float4(result.x, result.y, result.z, 0.0f);
}

// The batch's samples into the frame, pass by pass in order, as the persistent kernel
// adds them: a NaN sample is skipped, pass 0 replaces the frame.
__global__ __launch_bounds__(kWfBlock) void k_wf_accumulate(WfState st, int nPasses, float4* __restrict__ frame)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= st.N) return;
    float4 m = frame[pix];
    bool stored = false;
    for (int k = 0; k < nPasses; k++) {
        const float4 r = st.res[(size_t)k * st.N + pix];
        const vec3 result = v3(r.x, r.y, r.z);
        if (hasNan(result)) continue;
        const int spp = st.firstSpp + k;
        const vec3 last = (spp == 0) ? v3(0.0f) : v3(m.x, m.y, m.z);
        const vec3 sum = last + result;
        m = make_float4(sum.x, sum.y, sum.z, 1.0f);
        stored = true;
    }
    if (stored) frame[pix] = m;
}

__global__ __launch_bounds__(kWfBlock) void k_wf_shade(DevScene s, PassParams prm, int it, WfState st, WfQueues q)
{
    const int n = q.counts[it];
    const int* __restrict__ listIn = q.list[it & 1];
    int* listOut = q.list[(it + 1) & 1];
    WfRay* raysOut = q.rays[(it + 1) & 1];
    int* countList = q.counts + it + 1;
    int* countRays = q.counts + q.L + it + 1;
    const bool uniform = prm.lightEnvUniform != 0;
    const float portion = prm.lightPortion;
    __shared__ int sh[8];
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        bool keep = false, shadow = false, extend = false;
        int path = 0;
        WfRay shadowOut, extendOut;
        if (i < n) {
            Counters cn;
            path = listIn[i];
            const float4 A = st.a[path], Bv = st.b[path], Cv = st.c[path], D = st.d[path], E = st.e[path];
            const int4 Fv = st.f[path];
            vec3 result = v3(A.x, A.y, A.z);
            vec3 throughput = v3(A.w, Bv.x, Bv.y);
            vec3 pos = v3(Bv.z, Bv.w, Cv.x);
            vec3 wo = v3(Cv.y, Cv.z, Cv.w);
            vec3 pend = v3(D.x, D.y, D.z);
            float bsdfPdf = D.w;
            vec3 wi = v3(E.x, E.y, E.z);
            float rrProb = E.w;
            Rng rng;
            rng.randSeed = (uint32_t)Fv.x;
            rng.sampleSeed = (uint32_t)Fv.y;
            rng.dim = Fv.z;
            rng.sampleOffset = sampleOffsetOf(st.firstSpp + path / st.N);
            int flags = Fv.w;
            int bounce = flags & fBounceMask;
            bool alive = true;
            bool surface = false;
            int id = -1;

            // 1. the shadow ray of the previous vertex
            if (flags & fShadow) {
                if (st.occluded[path] == 0 && (flags & fPendValid)) result += pend;
                if (!(flags & fExtend)) alive = false;  // the BSDF sample ended the path
            }
            // 2. the closest hit of the camera or extension ray
            if (alive && (flags & fPrimary)) {
                id = st.hitId[path];
                const Ray ray = Ray{pos, wi};
                const vec3 hitPos = rayPoint(ray, st.hitDist[path]);
                if (id == -1) {
                    result = envLe<false>(s, wi, cn);
                    alive = false;
                } else if (id - s.objPrimCount >= 0) {
                    result = lightLe<false>(s, id - s.objPrimCount, hitPos, -wi, cn);
                    alive = false;
                } else {
                    wo = -wi;
                    result = v3(0.f);
                    throughput = v3(1.f);
                    bounce = 1;
                    pos = hitPos;
                    if (prm.maxDepth >= 1) surface = true;
                    else alive = false;
                }
            } else if (alive && (flags & fExtend)) {
                const int nextId = st.hitId[path];
                const int lightId = nextId - s.objPrimCount;
                const bool delta = (flags & fDelta) != 0;
                const vec3 nextPos = rayPoint(rayOffseted(pos, wi), st.hitDist[path]);
                if (nextId == -1) {
                    vec3 radiance = envLe<false>(s, wi, cn);
                    float weight = 1.0f;
                    if (prm.sampleLight && !delta) {
                        float envPdf = envPdfLi<false>(s, wi, cn) * pdfSelectEnv(s, uniform, portion);
                        weight = (envPdf <= 0.0f) ? 0.0f : bsdfPdf * bsdfPdf / (bsdfPdf * bsdfPdf + envPdf * envPdf);
                    }
                    result += radiance * throughput * weight;
                    alive = false;
                } else if (lightId >= 0) {
                    vec3 radiance = lightLe<false>(s, lightId, nextPos, -wi, cn);
                    float weight = 1.0f;
                    if (prm.sampleLight && !delta) {
                        float lightPdf = lightPdfLi<false>(s, lightId, pos, nextPos, cn) *
                                         pdfSelectLight<false>(s, lightId, uniform, portion, cn);
                        weight = (lightPdf <= 0.0f) ? 0.0f
                                                    : bsdfPdf * bsdfPdf / (bsdfPdf * bsdfPdf + lightPdf * lightPdf);
                    }
                    result += radiance * throughput * weight;
                    alive = false;
                } else {
                    // id and wo are set before the roulette: a killed path never reads them.
                    // (Set after it, inside the branch below, the surface step read a stale wo
                    // on this compiler with roulette enabled; tests/test_gpu_wavefront.py
                    // covers the roulette case.)
                    id = nextId;
                    wo = -wi;
                    if (prm.russianRoulette) {
                        if (sample1D<false>(s, rng, cn) >= rrProb) alive = false;
                        else throughput /= rrProb;
                    }
                    if (alive) {
                        pos = nextPos;
                        bounce++;
                        if (bounce > prm.maxDepth) alive = false;
                        else surface = true;
                    }
                }
            }
            // 3. a surface vertex: NEE set-up and BSDF sample
            flags = 0;
            if (alive && surface) {
                SurfaceInfo surf = triangleSurfaceInfo<false>(s, id, pos, cn);
                const int matTex = s.matTexIndices[id];
                const int matId = matTex & 0x0000ffff;
                const uint32_t matType = loadMaterialType<false>(s, matId, cn);
                if (matType != MatDielectric && matType != MatThinDielectric) {
                    if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
                }
                const BSDFParam mp = loadMaterial<false>(s, matType, matId, matTex >> 16, surf.uv, cn);
                Ray shadowRay;
                float shadowDist = 0.f;
                bool pendValid = false;
                if (prm.sampleLight) {
                    const float ud = sample1D<false>(s, rng, cn);
                    const vec4 us = sample4D<false>(s, rng, cn);
                    const float pdfSampleLight = pdfSampleLightOf(s, uniform, portion);
                    const bool pickLight = ud < pdfSampleLight;
                    const float pdfSelect = pickLight ? pdfSampleLight : 1.0f - pdfSampleLight;
                    LightLiSample ls = invalidLi();
                    bool valid = false;
                    if (pickLight) {
                        const int lid = lightSampleOne<false>(s, v2(us.x, us.y), cn);
                        const int triId = lid + s.objPrimCount;
                        vec3 a, b, c;
                        triangleVerts<false>(s, triId, a, b, c, cn);
                        const vec3 yp = sampleTriangleUniform(a, b, c, v2(us.z, us.w));
                        const vec3 wl = normalize(yp - pos);
                        const vec3 norm = triangleSurfaceInfo<false>(s, triId, yp, cn).ng;
                        const float cosTheta = dot(norm, -wl);
                        if (cosTheta >= 1e-6f) {
                            shadow = true;
                            shadowRay = rayOffseted(pos, wl);
                            const float dist = distance(pos, yp);
                            float pdf = dist * dist / (triangleArea(a, b, c) * cosTheta);
                            shadowDist = dist - 1e-4f - 1e-6f;
                            if (!(pdf < 1e-8f)) {
                                const vec3 weight = lightLe<false>(s, lid, yp, -wl, cn);
                                const float pdfSample = lightPdfSampleOne<false>(s, lid, cn);
                                pdf *= pdfSample;
                                ls = LightLiSample{wl, weight / pdf, pdf};
                                valid = true;
                            }
                        }
                    } else {
                        const vec4 sp = envSampleWi<false>(s, us, cn);
                        const vec3 we = v3(sp.x, sp.y, sp.z);
                        shadow = true;
                        shadowRay = rayOffseted(pos, we);
                        shadowDist = 1e8f;
                        if (sp.w != 0.0f) {
                            ls = LightLiSample{we, envLe<false>(s, we, cn) / sp.w, sp.w};
                            valid = true;
                        }
                    }
                    if (valid) {
                        ls.coef /= pdfSelect;
                        ls.pdf *= pdfSelect;
                        if (ls.pdf > 0.0f) {
                            vec3 f = materialBSDF(matType, mp, wo, ls.wi, surf.ns, ModeRadiance);
                            float fp = materialPdf(matType, mp, wo, ls.wi, surf.ns);
                            float weight = ls.pdf * ls.pdf / (ls.pdf * ls.pdf + fp * fp);
                            pend = f * throughput * satDot(surf.ns, ls.wi) * ls.coef * weight;
                            pendValid = true;
                        }
                    }
                }
                const vec3 u3 = sample3D<false>(s, rng, cn);
                const BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeRadiance, u3, rng);
                wi = bs.wi;
                bsdfPdf = bs.pdf;
                const bool delta = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
                const bool terminate = bsdfPdf < 1e-8f;
                if (!terminate) {
                    throughput *= bs.bsdf / bsdfPdf * (delta ? 1.0f : absDot(surf.ns, wi));
                    rrProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 0.95f);
                }
                extend = !terminate;
                flags = bounce | (pendValid ? fPendValid : 0) | (delta ? fDelta : 0) | (terminate ? fTerminate : 0) |
                        (shadow ? fShadow : 0) | (extend ? fExtend : 0);
                if (shadow) shadowOut = makeRay(shadowRay, shadowDist, path, true);
                if (extend) extendOut = makeRay(rayOffseted(pos, wi), 1e8f, path, false);
                keep = shadow || extend;
                if (!keep) alive = false;
            }
            if (!alive || !keep) {
                finishPath(st, path, result);
                keep = false;
            } else {
                st.a[path] = make_float4(result.x, result.y, result.z, throughput.x);
                st.b[path] = make_float4(throughput.y, throughput.z, pos.x, pos.y);
                st.c[path] = make_float4(pos.z, wo.x, wo.y, wo.z);
                st.d[path] = make_float4(pend.x, pend.y, pend.z, bsdfPdf);
                st.e[path] = make_float4(wi.x, wi.y, wi.z, rrProb);
                st.f[path] = make_int4((int)rng.randSeed, (int)rng.sampleSeed, rng.dim, flags);
            }
        }
        pushBlock(listOut, countList, keep, path, sh);
        pushBlock(raysOut, countRays, shadow && keep, shadowOut, sh);
        pushBlock(raysOut, countRays, extend && keep, extendOut, sh);
    }
}

}  // namespace

// Bytes per path: state 6 x 16, hit 12, result 16, two list slots 8, two queues of two
// rays 128.
constexpr size_t kWfPathBytes = 6 * 16 + 12 + 16 + 2 * 4 + 2 * 2 * sizeof(WfRay);

size_t wavefrontBytes(int w, int h, int passes)
{
    return (size_t)w * h * passes * kWfPathBytes + 64 * 1024;
}

int wavefrontBatch(int w, int h, int nPasses, size_t budget)
{
    const size_t per = (size_t)w * h * kWfPathBytes;
    int cap = (int)std::max<size_t>(1, budget / std::max<size_t>(per, 1));
    cap = std::min(cap, (int)((1u << 30) / std::max<size_t>((size_t)w * h, 1)));  // path << 1 fits an int
    cap = std::max(cap, 1);
    const int batches = (nPasses + cap - 1) / cap;
    return (nPasses + batches - 1) / batches;
}

void wavefrontBlocks(int& traceBlocksPerCU, int& shadeBlocksPerCU)
{
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_wf_trace, kWfBlock, 0) != hipSuccess) a = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wf_shade, kWfBlock, 0) != hipSuccess) b = 0;
    traceBlocksPerCU = residentBlocksPerCU(reinterpret_cast<const void*>(k_wf_trace), kWfBlock, a);
    shadeBlocksPerCU = residentBlocksPerCU(reinterpret_cast<const void*>(k_wf_shade), kWfBlock, b);
}

hipError_t launchNaiveWavefront(const DevScene& s, const PassParams& p, float4* frame, void* work, int batch,
                                int traceBlocks, int shadeBlocks, hipStream_t st)
{
    const size_t n = (size_t)s.filmW * s.filmH;
    const size_t m = n * (size_t)batch;
    char* w = static_cast<char*>(work);
    auto take = [&](size_t bytes) {
        char* r = w;
        w += (bytes + 255) & ~(size_t)255;
        return r;
    };
    WfState S;
    S.a = reinterpret_cast<float4*>(take(m * 16));
    S.b = reinterpret_cast<float4*>(take(m * 16));
    S.c = reinterpret_cast<float4*>(take(m * 16));
    S.d = reinterpret_cast<float4*>(take(m * 16));
    S.e = reinterpret_cast<float4*>(take(m * 16));
    S.f = reinterpret_cast<int4*>(take(m * 16));
    S.hitId = reinterpret_cast<int*>(take(m * 4));
    S.hitDist = reinterpret_cast<float*>(take(m * 4));
    S.occluded = reinterpret_cast<int*>(take(m * 4));
    S.res = reinterpret_cast<float4*>(take(m * 16));
    S.N = (int)n;
    WfQueues q;
    q.list[0] = reinterpret_cast<int*>(take(m * 4));
    q.list[1] = reinterpret_cast<int*>(take(m * 4));
    q.rays[0] = reinterpret_cast<WfRay*>(take(2 * m * sizeof(WfRay)));
    q.rays[1] = reinterpret_cast<WfRay*>(take(2 * m * sizeof(WfRay)));
    q.L = p.maxDepth + 2;
    q.counts = reinterpret_cast<int*>(take((size_t)3 * q.L * 4));
    const int nIndex = ((s.filmW + 7) / 8) * ((s.filmH + 7) / 8) * 64;
    for (int k0 = 0; k0 < p.nPasses; k0 += batch) {
        const int passes = std::min(batch, p.nPasses - k0);
        S.firstSpp = p.firstPass + k0;
        S.firstFc = (uint32_t)(p.freeCounter + k0);
        k_wf_reset<<<1, kWfBlock, 0, st>>>(q.counts, 3 * q.L, q.L, (int)(n * passes));
        k_wf_camera<<<dim3((nIndex + kWfBlock - 1) / kWfBlock, passes), kWfBlock, 0, st>>>(s, nIndex, S, q);
        for (int it = 0; it <= p.maxDepth; it++) {
            k_wf_trace<<<traceBlocks, kWfBlock, 0, st>>>(s, q.rays[it & 1], q.counts + q.L + it,
                                                         q.counts + 2 * q.L + it, S);
            k_wf_shade<<<shadeBlocks, kWfBlock, 0, st>>>(s, p, it, S, q);
        }
        k_wf_accumulate<<<(unsigned)((n + kWfBlock - 1) / kWfBlock), kWfBlock, 0, st>>>(S, passes, frame);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace zgl
