// zgl_persistent.hip -- persistent, regenerating schedule of the naive MIS path tracer.
//
// Same estimator as path_integ_naive.glsl:35-174 (and k_naive), scheduled for wave64:
//   * every lane owns one pixel at a time and renders that pixel's passes in order,
//     so its running sum is formed exactly as one launch per pass would form it;
//   * a lane's path is a state machine whose states end by issuing one ray (camera,
//     shadow or extension).  Walks advance one MTBVH row per step for all lanes of the
//     wave; when a ballot shows that no more than `threshold` lanes are still walking,
//     the lanes whose walk finished run their shading transition and issue their next
//     ray, so lanes with short walks do not wait for the longest walk of the wave;
//   * lanes that finish all passes of their pixel take the next pixel from a global
//     counter, one atomic per wave (ballot + prefix count over the lanes that need one).
// NEE contributions are evaluated before the shadow walk with the throughput of that
// bounce and added after it if the ray is unoccluded, which is the reference's order of
// additions; the BSDF is sampled after the NEE samples are drawn, the reference's order
// of random-number use.
#include "zgl_internal.h"

#include <algorithm>

namespace zgl {

namespace {

enum : int {
    kNewPixel = 0,
    kNewSample,
    kPrimary,   // camera ray walking / walked
    kSurface,   // at a surface vertex, nothing issued yet
    kShadow,    // NEE shadow ray walking / walked
    kExtend,    // extension ray walking / walked
    kFinish,    // sample complete
    kDone,
    kWait,      // batch: the wave's scene has no pixels left; wait for the next scene
};

// MTBVH rows each walking lane advances per round of the walking loop (one leaf test
// and one ballot per round).
constexpr int kRowsPerRound = 4;
constexpr int kRegions = kUnitRegions;  // unit regions, one per XCD

// Pixel order for the dynamic fetch: 8x8 tiles in row-major tile order.
__device__ __forceinline__ void pixelOfIndex(int idx, int W, int& x, int& y)
{
    const int tilesX = (W + 7) / 8;
    const int tile = idx >> 6, in = idx & 63;
    x = (tile % tilesX) * 8 + (in & 7);
    y = (tile / tilesX) * 8 + (in >> 3);
}

}  // namespace

__global__ void k_reset_counter(int* c) { if (threadIdx.x < kRegions && blockIdx.x == 0) c[threadIdx.x] = 0; }

template <bool C> __device__ void flushCountersP(const Counters& cn, unsigned long long* out)
{
    if (!C) return;
    atomicAdd(out + 0, (unsigned long long)cn.bytes);
    atomicAdd(out + 1, (unsigned long long)cn.rows);
    atomicAdd(out + 2, (unsigned long long)cn.tests);
    atomicAdd(out + 3, (unsigned long long)cn.traversals);
}

// The persistent schedule.  The work is one sample per (pixel, pass), handed out in
// (8x8 tile, pass) groups; each sample's value goes to res[pass * W * H + pixel] and
// k_accumulate adds the passes to the frame in order, so that no lane is left with a
// long run of one pixel's passes when the work runs out.
// BATCH = false: one scene (s0, prm0, res0, counter0).
// LOG (single scene only): every traversal query's result is also written to
// hitLog[(pass * W * H + pixel) * kHitLogQueries + query], in issue order (primary,
// then per vertex the shadow query if one is cast and the extension), up to
// kHitLogQueries per sample: {closest primitive, distance bits} for closest-hit queries,
// {1 if occluded else 0, 0} for shadow queries.  The adjoint replays read it.
// BATCH = true: the scenes of `items` in order; a wave works on one scene at a time
// (scene data stays wave-uniform), takes its samples from that scene's counter, and
// moves to the next scene once its lanes have finished the samples they hold; its
// regeneration threshold is that scene's (items[].threshold).
template <bool C, bool BATCH, bool LOG>
__device__ __forceinline__ void naiveBody(const DevScene& s0, const PassParams& prm0, float4* res0, int* counter0,
                                          const BatchScene* items, int nScenes, int* sceneCounters, int threshold,
                                          unsigned long long* counters, int2* hitLog)
{
    const int lane = threadIdx.x & 63;
    Counters cn;
    int cs = 0;
    int region = (int)(blockIdx.x % kRegions);  // the workgroup's XCD under round-robin dispatch
    int exhausted = 0;  // regions of the current scene found used up in a row (the same in every lane)

    int phase = kNewPixel;
    int pix = -1, pass = 0;
    int query = 0;  // LOG: traversal queries of the current sample so far
    vec2 scr = v2(0.f, 0.f);
    PixelSeed ps = {0u, 0u};
    Rng rng = {0u, 0u, 0, 0};
    vec3 result = v3(0.f), throughput = v3(1.f), wo = v3(0.f), pos = v3(0.f), wi = v3(0.f);
    int bounce = 0;
    float bsdfPdf = 0.f, rrProb = 0.f;
    bool delta = false, terminate = false;
    vec3 pend = v3(0.f);
    bool pendValid = false;
    uint32_t pendBytes = 0;
    bool walking = false;
    Walk wk;
    wk.p.o = v3(0.f); wk.p.d = v3(0.f); wk.p.dInv = v3(0.f); wk.p.mode = 0;
    wk.p.nx = wk.p.ny = wk.p.nz = false; wk.p.table = 0;
    wk.dist = 0.f; wk.k = 0; wk.closest = -1; wk.any = false;

    for (;;) {
        const int csu = BATCH ? __builtin_amdgcn_readfirstlane(cs) : 0;
        const DevScene& s = BATCH ? items[csu].s : sceneArg(s0);
        const PassParams& prm = BATCH ? items[csu].prm : prm0;
        float4* const res = BATCH ? items[csu].res : res0;
        int* const pixelCounter = BATCH ? sceneCounters + csu * kRegions : counter0;
        const int sceneThreshold = BATCH ? items[csu].threshold : threshold;
        const int W = s.filmW, H = s.filmH;
        const int nIndex = ((W + 7) / 8) * ((H + 7) / 8) * 64;
        // work units: (tile, pass, pixel of the tile)
        const int nUnits = nIndex * prm.nPasses;
        const bool uniform = prm.lightEnvUniform != 0;
        const float portion = prm.lightPortion;
        // LOG: the result of the walk that just ended, as the next query of the sample
        auto logQuery = [&](bool shadowQuery) {
            if constexpr (LOG) {
                if (query < kHitLogQueries) {
                    const int2 e = shadowQuery ? make_int2(wk.closest >= 0 ? 1 : 0, 0)
                                               : make_int2(wk.closest, __float_as_int(wk.dist));
                    hitLog[((size_t)pass * ((size_t)W * H) + pix) * kHitLogQueries + query] = e;
                }
                query++;
            }
        };

        // ------------------------------------------------------------ shading
        if (!walking) {
            // completion of the ray that was walked
            if (phase == kPrimary) {
                logQuery(false);
                const int id = wk.closest;
                const Ray ray = makeRay(wk.p.o, wk.p.d);
                pos = rayPoint(ray, wk.dist);
                if (id == -1) {
                    result = envLe<C>(s, ray.dir, cn);
                    phase = kFinish;
                } else if (id - s.objPrimCount >= 0) {
                    result = lightLe<C>(s, id - s.objPrimCount, pos, -ray.dir, cn);
                    phase = kFinish;
                } else {
                    wo = -ray.dir;
                    result = v3(0.f);
                    throughput = v3(1.f);
                    bounce = 1;
                    phase = (prm.maxDepth >= 1) ? kSurface : kFinish;
                }
            } else if (phase == kShadow) {
                logQuery(true);
                if (wk.closest < 0) {
                    if (C) cn.bytes += pendBytes;
                    if (pendValid) result += pend;
                }
                if (terminate) phase = kFinish;
                else {
                    walkStart<C>(s, wk, rayOffseted(pos, wi), false, 1e8f, cn);
                    walking = true;
                    phase = kExtend;
                }
            } else if (phase == kExtend) {
                logQuery(false);
                const int nextId = wk.closest;
                const int lightId = nextId - s.objPrimCount;
                const vec3 nextPos = rayPoint(makeRay(wk.p.o, wk.p.d), wk.dist);
                if (nextId == -1) {
                    vec3 radiance = envLe<C>(s, wi, cn);
                    float weight = 1.0f;
                    if (prm.sampleLight && !delta) {
                        float envPdf = envPdfLi<C>(s, wi, cn) * pdfSelectEnv(s, uniform, portion);
                        weight = (envPdf <= 0.0f) ? 0.0f : bsdfPdf * bsdfPdf / (bsdfPdf * bsdfPdf + envPdf * envPdf);
                    }
                    result += radiance * throughput * weight;
                    phase = kFinish;
                } else if (lightId >= 0) {
                    vec3 radiance = lightLe<C>(s, lightId, nextPos, -wi, cn);
                    float weight = 1.0f;
                    if (prm.sampleLight && !delta) {
                        float lightPdf = lightPdfLi<C>(s, lightId, pos, nextPos, cn) *
                                         pdfSelectLight<C>(s, lightId, uniform, portion, cn);
                        weight = (lightPdf <= 0.0f) ? 0.0f
                                                    : bsdfPdf * bsdfPdf / (bsdfPdf * bsdfPdf + lightPdf * lightPdf);
                    }
                    result += radiance * throughput * weight;
                    phase = kFinish;
                } else {
                    // the next vertex's state is set before the roulette, and the roulette
                    // is written as selects: state a surviving path set after a roulette
                    // branch has been compiled wrongly (the global-queue tracer)
                    wk.closest = nextId;
                    pos = nextPos;
                    wo = -wi;
                    bool survive = true;
                    if (prm.russianRoulette) {
                        survive = !(sample1D<C>(s, rng, cn) >= rrProb);
                        const vec3 scaled = throughput / rrProb;
                        throughput = select3(survive, scaled, throughput);
                    }
                    bounce++;
                    phase = (survive && bounce <= prm.maxDepth) ? kSurface : kFinish;
                }
            }

            // a surface vertex: NEE sample + BSDF sample, issue the next ray
            if (phase == kSurface) {
                const int id = wk.closest;
                SurfaceInfo surf = triangleSurfaceInfo<C>(s, id, pos, cn);
                const int matTex = s.matTexIndices[id];
                const int matId = matTex & 0x0000ffff;
                cnt<C>(cn, 4);
                const uint32_t matType = loadMaterialType<C>(s, matId, cn);
                if (matType != MatDielectric && matType != MatThinDielectric) {
                    if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
                }
                const BSDFParam mp = loadMaterial<C>(s, matType, matId, matTex >> 16, surf.uv, cn);
                bool shadow = false;
                Ray shadowRay;
                float shadowDist = 0.f;
                pendValid = false;
                pendBytes = 0;
                if (prm.sampleLight) {
                    const float ud = sample1D<C>(s, rng, cn);
                    const vec4 us = sample4D<C>(s, rng, cn);
                    const float pdfSampleLight = pdfSampleLightOf(s, uniform, portion);
                    const bool pickLight = ud < pdfSampleLight;
                    const float pdfSelect = pickLight ? pdfSampleLight : 1.0f - pdfSampleLight;
                    Counters deferred;  // bytes the reference reads only after an unoccluded test
                    LightLiSample ls = invalidLi();
                    bool valid = false;
                    if (pickLight) {
                        const int lid = lightSampleOne<C>(s, v2(us.x, us.y), cn);
                        const int triId = lid + s.objPrimCount;
                        vec3 a, b, c;
                        triangleVerts<C>(s, triId, a, b, c, cn);
                        const vec3 y = sampleTriangleUniform(a, b, c, v2(us.z, us.w));
                        const vec3 wl = normalize(y - pos);
                        const vec3 norm = triangleSurfaceInfo<C>(s, triId, y, cn).ng;
                        const float cosTheta = dot(norm, -wl);
                        if (cosTheta >= 1e-6f) {
                            shadow = true;
                            shadowRay = rayOffseted(pos, wl);
                            const float dist = distance(pos, y);
                            float pdf = dist * dist / (triangleArea(a, b, c) * cosTheta);
                            shadowDist = dist - 1e-4f - 1e-6f;
                            if (!(pdf < 1e-8f)) {
                                const vec3 weight = lightLe<C>(s, lid, y, -wl, deferred);
                                const float pdfSample = lightPdfSampleOne<C>(s, lid, deferred);
                                pdf *= pdfSample;
                                ls = LightLiSample{wl, weight / pdf, pdf};
                                valid = true;
                            }
                        }
                    } else {
                        const vec4 sp = envSampleWi<C>(s, us, cn);
                        const vec3 we = v3(sp.x, sp.y, sp.z);
                        shadow = true;
                        shadowRay = rayOffseted(pos, we);
                        shadowDist = 1e8f;
                        if (sp.w != 0.0f) {
                            ls = LightLiSample{we, envLe<C>(s, we, deferred) / sp.w, sp.w};
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
                    if (C) pendBytes = deferred.bytes;
                }
                const vec3 u3 = sample3D<C>(s, rng, cn);
                const BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeRadiance, u3, rng);
                wi = bs.wi;
                bsdfPdf = bs.pdf;
                delta = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
                terminate = bsdfPdf < 1e-8f;
                if (!terminate) {
                    throughput *= bs.bsdf / bsdfPdf * (delta ? 1.0f : absDot(surf.ns, wi));
                    rrProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 0.95f);
                }
                if (shadow) {
                    walkStart<C>(s, wk, shadowRay, true, shadowDist, cn);
                    walking = true;
                    phase = kShadow;
                } else if (terminate) {
                    phase = kFinish;
                } else {
                    walkStart<C>(s, wk, rayOffseted(pos, wi), false, 1e8f, cn);
                    walking = true;
                    phase = kExtend;
                }
            }

            if (phase == kFinish) {
                res[(size_t)pass * ((size_t)W * H) + pix] = make_float4(result.x, result.y, result.z, 0.0f);
                phase = kNewPixel;
            }
        }

        // pixels are handed out per wave: one atomic for all lanes that need one.  One
        // region of 8x8 tiles per XCD (workgroups are dispatched to XCDs round-robin), so
        // that the waves sharing an L2 start with neighbouring pixels; a wave whose region
        // is used up moves on to the next, and once all eight are, stops (one scene) or
        // waits to move to the next scene (batch).
        const unsigned long long need = __ballot(!walking && phase == kNewPixel);
        if (need) {
            const int n = __popcll(need);
            const int leader = __ffsll((long long)need) - 1;
            int base = 0;
            // the wave's region of the image first (one per XCD), then the others
            const int len = ((nUnits / kRegions + 63) / 64) * 64;
            const int start = min(nUnits, region * len);
            const int limit = min(nUnits, start + len);
            if (lane == leader) base = atomicAdd(pixelCounter + region, n);
            base = __shfl(base, leader) + start;
            if (base >= limit) {
                region = (region + 1) % kRegions;
                exhausted++;
            } else {
                exhausted = 0;
            }
            if (!walking && phase == kNewPixel) {
                const unsigned long long below = need & ((1ull << lane) - 1ull);
                const int idx = base + __popcll(below);
                pix = -1;
                if (exhausted >= kRegions) {
                    phase = BATCH ? kWait : kDone;
                } else if (idx >= limit) {
                    // this region is used up: fetch again from the next one
                } else {
                    const int group = idx >> 6;
                    const int unitPass = group % prm.nPasses;
                    const int pixIdx = (group / prm.nPasses) * 64 + (idx & 63);
                    int x, y;
                    pixelOfIndex(pixIdx, W, x, y);
                    if (x < W && y < H) {  // slots of partial edge tiles fetch again next round
                        pix = y * W + x;
                        scr = v2((float)x, (float)y) / v2((float)W, (float)H);
                        ps = pixelSeed(s, scr);
                        pass = unitPass;
                        phase = kNewSample;
                    }
                }
            }
        }

        if (!walking && phase == kNewSample) {
            const int spp = prm.firstPass + pass;
            const uint32_t fc = (uint32_t)(prm.freeCounter + pass);
            rng.randSeed = ps.tx * fc + ps.ty;
            rng.sampleSeed = ps.tx * ps.ty;
            rng.sampleOffset = sampleOffsetOf(spp);
            rng.dim = 0;
            query = 0;
            if (C) cn.bytes += 64 + 32;
            const vec4 u = sample4D<C>(s, rng, cn);
            walkStart<C>(s, wk, cameraSampleRay(s, scr, u), false, 1e8f, cn);
            walking = true;
            phase = kPrimary;
        }

        if (BATCH) {
            // every lane idle and some waiting: move the wave to the next scene
            if (__ballot(walking || (phase != kWait && phase != kDone)) == 0ull && __ballot(phase == kWait) != 0ull) {
                cs = csu + 1;
                region = (int)(blockIdx.x % kRegions);
                exhausted = 0;
                if (phase == kWait) phase = (cs < nScenes) ? kNewPixel : kDone;
                continue;
            }
        }
        if (__ballot(phase != kDone) == 0ull) break;

        // ------------------------------------------------------------ walking
        // Each round advances every walking lane by up to kRowsPerRound rows; a lane
        // that reaches a leaf whose box is hit waits there, and the waiting lanes test
        // their triangles together once per round, so that the triangle test (as long
        // as the box test) runs once per round instead of on nearly every row step for
        // a few lanes.  A lane whose shadow walk ends settles its NEE term and starts its
        // extension walk within the loop (the same operations, in the same order, as the
        // shading step would do), instead of idling until the wave leaves the loop.
        {
            bool general = __ballot(walking && wk.p.special) == 0ull;
            const WalkMem wm = walkMem(s);
            for (;;) {
#pragma unroll
                for (int u = 0; u < kRowsPerRound; ++u) {
                    if (walking && wk.leaf < 0) {
                        if (walkRow<C>(wm, wk, general, cn)) walking = false;
                    }
                }
                if (walking && wk.leaf >= 0) {
                    if (walkLeaf<C>(wm, wk, cn)) walking = false;
                }
                // a shadow walk that ended settles its NEE term and starts the extension walk
                if (!walking && phase == kShadow) {
                    logQuery(true);
                    if (wk.closest < 0) {
                        if (C) cn.bytes += pendBytes;
                        if (pendValid) result += pend;
                    }
                    if (terminate) {
                        phase = kFinish;
                    } else {
                        walkStart<C>(s, wk, rayOffseted(pos, wi), false, 1e8f, cn);
                        walking = true;
                        phase = kExtend;
                    }
                }
                general = __ballot(walking && wk.p.special) == 0ull;
                if (__popcll(__ballot(walking)) <= sceneThreshold) break;
            }
        }
    }
    flushCountersP<C>(cn, counters);
}

template <bool C, bool LOG>
__global__ __launch_bounds__(256, 4) void k_naive_persistent(DevScene s, PassParams prm, float4* __restrict__ res,
                                                          int* __restrict__ pixelCounter, int threshold,
                                                          unsigned long long* counters, int2* __restrict__ hitLog)
{
    naiveBody<C, false, LOG>(s, prm, res, pixelCounter, nullptr, 0, nullptr, threshold, counters, hitLog);
}

// A launch's samples into the frame, pass by pass in order: a NaN sample is skipped and
// pass 0 replaces the frame (path_integ_naive.glsl:168-173).
__device__ __forceinline__ void accumulatePixel(const float4* __restrict__ res, int pix, int n, int nPasses,
                                                int firstPass, float4* __restrict__ frame)
{
    float4 m = frame[pix];
    bool stored = false;
    for (int k = 0; k < nPasses; k++) {
        const float4 r = res[(size_t)k * n + pix];
        const vec3 result = v3(r.x, r.y, r.z);
        if (hasNan(result)) continue;
        const vec3 last = (firstPass + k == 0) ? v3(0.0f) : v3(m.x, m.y, m.z);
        const vec3 sum = last + result;
        m = make_float4(sum.x, sum.y, sum.z, 1.0f);
        stored = true;
    }
    if (stored) frame[pix] = m;
}

__global__ __launch_bounds__(256) void k_accumulate(const float4* __restrict__ res, int n, int nPasses, int firstPass,
                                                    float4* __restrict__ frame)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix < n) accumulatePixel(res, pix, n, nPasses, firstPass, frame);
}

// The same for every scene of a batch: blockIdx.y (strided by gridDim.y) picks the scene.
__global__ __launch_bounds__(256) void k_accumulate_batch(const BatchScene* __restrict__ items, int nScenes)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = blockIdx.y; k < nScenes; k += gridDim.y) {
        const BatchScene& it = items[k];
        const int n = it.s.filmW * it.s.filmH;
        if (pix < n) accumulatePixel(it.res, pix, n, it.prm.nPasses, it.prm.firstPass, it.frame);
    }
}

__global__ __launch_bounds__(256, 4) void k_naive_batch(const BatchScene* __restrict__ items, int nScenes,
                                                     int* __restrict__ sceneCounters)
{
    naiveBody<false, true, false>(items[0].s, items[0].prm, nullptr, nullptr, items, nScenes, sceneCounters, 0,
                                  nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// Triple path tracer, camera pass (triple_path_pass_pt.glsl:44-195) on the same
// persistent schedule: one sample per (8x8 tile, pass) unit, XCD regions, walks advanced
// a few rows per ballot round with lanes shading and reissuing rays once no more than
// `threshold` are still walking.  The shading steps are tripleCameraTrace's, in its order
// of operations and random-number use; the s=1 term is evaluated before its visibility
// walk and added if the walk finds no occluder.  Each sample's value goes to
// res[pass * W * H + pixel]; k_triple_accumulate adds them to the planar film.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 4) void k_triple_persistent(DevScene s0, PassParams prm, float4* __restrict__ res,
                                                           int* __restrict__ pixelCounter, int threshold)
{
    const int lane = threadIdx.x & 63;
    Counters cn;
    int region = (int)(blockIdx.x % kRegions);
    int exhausted = 0;
    const int W = s0.filmW, H = s0.filmH;
    const int nIndex = ((W + 7) / 8) * ((H + 7) / 8) * 64;
    const int nUnits = nIndex * prm.nPasses;

    int phase = kNewPixel;
    int pix = -1, pass = 0;
    vec2 scr = v2(0.f, 0.f);
    PixelSeed ps = {0u, 0u};
    Rng rng = {0u, 0u, 0, 0};
    vec3 result = v3(0.f), throughput = v3(1.f), wo = v3(0.f), pos = v3(0.f), wi = v3(0.f);
    vec3 prevNorm = v3(0.f);
    int bounce = 0, id = -1;
    float t1s0 = 0.f, t1s1 = 0.f, pdfNext = 0.f, cPrev = 1.f, rrProb = 0.f;
    bool terminate = false;
    vec3 pend = v3(0.f);
    bool walking = false;
    Walk wk;
    wk.p.o = v3(0.f); wk.p.d = v3(0.f); wk.p.dInv = v3(0.f); wk.p.mode = 0;
    wk.p.nx = wk.p.ny = wk.p.nz = false; wk.p.table = 0;
    wk.dist = 0.f; wk.k = 0; wk.closest = -1; wk.any = false;

    for (;;) {
        const DevScene& s = sceneArg(s0);
        if (!walking) {
            if (phase == kPrimary) {
                const int hit = wk.closest;
                const Ray ray = makeRay(wk.p.o, wk.p.d);
                const float primDist = wk.dist;
                pos = rayPoint(ray, primDist);
                if (hit == -1) {
                    result = envLe<false>(s, ray.dir, cn);
                    phase = kFinish;
                } else if (hit - s.objPrimCount >= 0) {
                    result = lightLe<false>(s, hit - s.objPrimCount, pos, -ray.dir, cn);
                    phase = kFinish;
                } else {
                    const SurfaceInfo surf = triangleSurfaceInfo<false>(s, hit, pos, cn);
                    wo = -ray.dir;
                    prevNorm = s.camF;
                    result = v3(0.0f);
                    throughput = v3(1.0f);
                    const CameraPdf camPdf = cameraPdfIe(s, ray);
                    const float primaryPdf =
                        remapP(camPdf.pdfPos) / remapP(camPdf.pdfDir * absDot(surf.ns, ray.dir) / square(primDist));
                    t1s0 = primaryPdf;
                    t1s1 = primaryPdf;
                    id = hit;
                    bounce = 1;
                    phase = (prm.maxDepth >= 1) ? kSurface : kFinish;
                }
            } else if (phase == kShadow) {
                if (wk.closest < 0) result += pend;
                if (terminate) {
                    phase = kFinish;
                } else {
                    walkStart<false>(s, wk, rayOffseted(pos, wi), false, 1e8f, cn);
                    walking = true;
                    phase = kExtend;
                }
            } else if (phase == kExtend) {
                const int nextId = wk.closest;
                const int lightId = nextId - s.objPrimCount;
                const vec3 nextPos = rayPoint(makeRay(wk.p.o, wk.p.d), wk.dist);
                const float dist = wk.dist;
                if (nextId == -1) {
                    phase = kFinish;
                } else if (lightId >= 0) {
                    const vec3 nLit = triangleSurfaceInfo<false>(s, nextId, nextPos, cn).ng;
                    const LightPdf pl = lightPdfLe<false>(s, lightId, makeRay(nextPos, -wi), cn);
                    const float pdfPLit = remapP(pl.pdfPos * lightPdfSampleOne<false>(s, lightId, cn));
                    const float coefToLight = remapP(pdfNext * satDot(nLit, -wi));
                    const float coefToSurf = remapP(pl.pdfDir * absDot(prevNorm, wi));
                    const float coefDist = remapP(dist * dist);
                    const float s1s0 = pdfPLit * coefDist / coefToLight;
                    const float t1s0x = t1s0 * coefToSurf * pdfPLit * cPrev / coefToLight;
                    const float weight = fisnan(t1s0) ? 0.0f : 1.0f / (1.0f + s1s0 + t1s0x);
                    result += lightLe<false>(s, lightId, nextPos, -wi, cn) * throughput * weight;
                    phase = kFinish;
                } else {
                    // set before the roulette, which is written as selects (see naiveBody)
                    pos = nextPos;
                    wo = -wi;
                    id = nextId;
                    bool survive = true;
                    if (prm.russianRoulette) {
                        survive = !(sample1D<false>(s, rng, cn) >= rrProb);
                        const vec3 scaled = throughput / rrProb;
                        throughput = select3(survive, scaled, throughput);
                    }
                    // (also on a path that ends here, which does not read them again)
                    const float coef = cPrev / remapP(pdfNext * absDot(triangleNormalShad<false>(s, nextId, nextPos, cn), wi));
                    t1s0 *= coef;
                    t1s1 *= coef;
                    bounce++;
                    phase = (survive && bounce <= prm.maxDepth) ? kSurface : kFinish;
                }
            }

            if (phase == kSurface) {
                SurfaceInfo surf = triangleSurfaceInfo<false>(s, id, pos, cn);
                const int matTex = s.matTexIndices[id];
                const int matId = matTex & 0x0000ffff;
                const uint32_t matType = loadMaterialType<false>(s, matId, cn);
                if (matType != MatDielectric && matType != MatThinDielectric) {
                    if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
                }
                const BSDFParam mp = loadMaterial<false>(s, matType, matId, matTex >> 16, surf.uv, cn);
                bool shadow = false;
                Ray shadowRay;
                float shadowDist = 0.f;
                {
                    const int light = lightSampleOne<false>(s, sample2D<false>(s, rng, cn), cn);
                    const int triId = light + s.objPrimCount;
                    const float pdfSource = lightPdfSampleOne<false>(s, light, cn);
                    const vec3 pLit = triangleSampleUniformId<false>(s, triId, sample2D<false>(s, rng, cn), cn);
                    const vec3 wl = normalize(pLit - pos);
                    const vec3 Le = lightLe<false>(s, light, pLit, -wl, cn);
                    if (!isBlack(Le)) {
                        // visible(pos, pLit): offset 1e-5 along the direction, 2e-5 short of pLit
                        shadow = true;
                        shadowDist = distance(pos, pLit) - 2e-5f;
                        const vec3 wv = normalize(pLit - pos);
                        shadowRay = makeRay(pos + wv * 1e-5f, wv);
                        const vec3 nLit = triangleSurfaceInfo<false>(s, triId, pLit, cn).ng;
                        const float pA = pdfSource / triangleAreaId<false>(s, triId, cn);
                        const float dist2 = distSquare(pos, pLit);
                        const float pS = pA * dist2 / absDot(nLit, wl);
                        const vec3 f = materialBSDF(matType, mp, wo, wl, surf.ns, ModeRadiance);
                        const float fpdf = materialPdf(matType, mp, wo, wl, surf.ns);
                        const float pdfRev = materialPdf(matType, mp, wl, wo, surf.ns);
                        const float pdfPLit = remapP(pA);
                        const float coefToSurf = remapP(0.5f * PiInv * absDot(surf.ns, wl));
                        const float coefToLight = remapP(fpdf * satDot(nLit, -wl));
                        const float coefToPrev = (bounce == 1) ? 1.0f : remapP(pdfRev * absDot(prevNorm, wo));
                        const float coefDist = remapP(dist2);
                        const float s1s0 = pdfPLit * coefDist / coefToLight;
                        const float t1s1x = t1s1 * coefToSurf * coefToPrev / coefDist;
                        const float weight = s1s0 / (1.0f + s1s0 + s1s0 * t1s1x);
                        pend = Le * f * throughput * absDot(surf.ns, wl) / pS * weight;
                    }
                }
                const BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeRadiance, sample3D<false>(s, rng, cn), rng);
                wi = bs.wi;
                const float bsdfPdf = bs.pdf;
                const bool deltaBsdf = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
                terminate = bsdfPdf < 1e-8f;
                if (!terminate) {
                    throughput *= bs.bsdf / bsdfPdf * (deltaBsdf ? 1.0f : absDot(surf.ns, wi));
                    pdfNext = materialPdf(matType, mp, wo, wi, surf.ns);
                    const float pdfPrev = materialPdf(matType, mp, wi, wo, surf.ns);
                    cPrev = (bounce == 1) ? 1.0f : remapP(pdfPrev * absDot(prevNorm, wo));
                    rrProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 0.95f);
                    prevNorm = surf.ns;
                }
                if (shadow) {
                    walkStart<false>(s, wk, shadowRay, true, shadowDist, cn);
                    walking = true;
                    phase = kShadow;
                } else if (terminate) {
                    phase = kFinish;
                } else {
                    walkStart<false>(s, wk, rayOffseted(pos, wi), false, 1e8f, cn);
                    walking = true;
                    phase = kExtend;
                }
            }

            if (phase == kFinish) {
                res[(size_t)pass * ((size_t)W * H) + pix] = make_float4(result.x, result.y, result.z, 0.0f);
                phase = kNewPixel;
            }
        }

        // samples are handed out per wave, one region of 8x8 tiles per XCD first
        const unsigned long long need = __ballot(!walking && phase == kNewPixel);
        if (need) {
            const int n = __popcll(need);
            const int leader = __ffsll((long long)need) - 1;
            int base = 0;
            const int len = ((nUnits / kRegions + 63) / 64) * 64;
            const int start = min(nUnits, region * len);
            const int limit = min(nUnits, start + len);
            if (lane == leader) base = atomicAdd(pixelCounter + region, n);
            base = __shfl(base, leader) + start;
            if (base >= limit) {
                region = (region + 1) % kRegions;
                exhausted++;
            } else {
                exhausted = 0;
            }
            if (!walking && phase == kNewPixel) {
                const unsigned long long below = need & ((1ull << lane) - 1ull);
                const int idx = base + __popcll(below);
                pix = -1;
                if (exhausted >= kRegions) {
                    phase = kDone;
                } else if (idx < limit) {
                    const int group = idx >> 6;
                    const int unitPass = group % prm.nPasses;
                    const int pixIdx = (group / prm.nPasses) * 64 + (idx & 63);
                    int x, y;
                    pixelOfIndex(pixIdx, W, x, y);
                    if (x < W && y < H) {
                        pix = y * W + x;
                        scr = v2((float)x, (float)y) / v2((float)W, (float)H);
                        ps = pixelSeed(s, scr);
                        pass = unitPass;
                        phase = kNewSample;
                    }
                }
            }
        }

        if (!walking && phase == kNewSample) {
            const int spp = prm.firstPass + pass;
            const uint32_t fc = (uint32_t)(prm.freeCounter + pass);
            rng.randSeed = ps.tx * fc + ps.ty;
            rng.sampleSeed = ps.tx * ps.ty;
            rng.sampleOffset = sampleOffsetOf(spp);
            rng.dim = 0;
            const vec4 u = sample4D<false>(s, rng, cn);
            walkStart<false>(s, wk, cameraSampleRay(s, scr, u), false, 1e8f, cn);
            walking = true;
            phase = kPrimary;
        }

        if (__ballot(phase != kDone) == 0ull) break;

        {
            bool general = __ballot(walking && wk.p.special) == 0ull;
            const WalkMem wm = walkMem(s);
            for (;;) {
#pragma unroll
                for (int u = 0; u < kRowsPerRound; ++u) {
                    if (walking && wk.leaf < 0) {
                        if (walkRow<false>(wm, wk, general, cn)) walking = false;
                    }
                }
                if (walking && wk.leaf >= 0) {
                    if (walkLeaf<false>(wm, wk, cn)) walking = false;
                }
                // a visibility walk that ended settles its s=1 term and starts the extension
                if (!walking && phase == kShadow) {
                    if (wk.closest < 0) result += pend;
                    if (terminate) {
                        phase = kFinish;
                    } else {
                        walkStart<false>(s, wk, rayOffseted(pos, wi), false, 1e8f, cn);
                        walking = true;
                        phase = kExtend;
                    }
                }
                general = __ballot(walking && wk.p.special) == 0ull;
                if (__popcll(__ballot(walking)) <= threshold) break;
            }
        }
    }
}

// The camera pass's samples into the planar RGB film (3 floats per pixel, rows of 3W),
// pass by pass in order, NaN samples skipped (triple_path_pass_pt.glsl:218-222).
__global__ __launch_bounds__(256) void k_triple_accumulate(const float4* __restrict__ res, int W, int H, int nPasses,
                                                         float* __restrict__ film)
{
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = W * H;
    if (pix >= n) return;
    float* f = film + (size_t)(pix / W) * 3 * W + 3 * (pix % W);
    vec3 acc = v3(f[0], f[1], f[2]);
    for (int k = 0; k < nPasses; k++) {
        const float4 r = res[(size_t)k * n + pix];
        const vec3 v = v3(r.x, r.y, r.z);
        if (!hasNan(v)) acc = acc + v;
    }
    f[0] = acc.x;
    f[1] = acc.y;
    f[2] = acc.z;
}

// Zeroes the per-scene pixel counters of a batch (a kernel, so that it can sit in a graph).
__global__ void k_reset_counters(int* c, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

}  // namespace zgl

namespace zgl {

hipError_t launchNaivePersistent(const DevScene& s, const PassParams& p, float4* frame, float4* res, int resPasses,
                                 int* pixelCounter, int threshold, int blocks, unsigned long long* counters,
                                 hipStream_t st)
{
    if (threshold < 0) threshold = 0;
    if (threshold > 63) threshold = 63;
    const int n = s.filmW * s.filmH;
    // passes in chunks of at most resPasses (the capacity of res)
    for (int k0 = 0; k0 < p.nPasses; k0 += resPasses) {
        PassParams q = p;
        q.firstPass = p.firstPass + k0;
        q.freeCounter = p.freeCounter + k0;
        q.nPasses = std::min(resPasses, p.nPasses - k0);
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, pixelCounter);
        if (counters)
            hipLaunchKernelGGL((k_naive_persistent<true, false>), dim3(blocks), dim3(256), 0, st, s, q, res,
                               pixelCounter, threshold, counters, nullptr);
        else
            hipLaunchKernelGGL((k_naive_persistent<false, false>), dim3(blocks), dim3(256), 0, st, s, q, res,
                               pixelCounter, threshold, counters, nullptr);
        hipLaunchKernelGGL(k_accumulate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, res, n, q.nPasses,
                           q.firstPass, frame);
    }
    return hipGetLastError();
}

hipError_t launchNaiveLogged(const DevScene& s, const PassParams& p, float4* res, int2* hitLog, int* pixelCounter,
                             int threshold, int blocks, hipStream_t st)
{
    if (threshold < 0) threshold = 0;
    if (threshold > 63) threshold = 63;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, pixelCounter);
    hipLaunchKernelGGL((k_naive_persistent<false, true>), dim3(blocks), dim3(256), 0, st, s, p, res, pixelCounter,
                       threshold, nullptr, hitLog);
    return hipGetLastError();
}

hipError_t launchNaiveBatch(const BatchScene* items, int nScenes, int maxPixels, int* sceneCounters, int blocks,
                            hipStream_t st)
{
    const int nCounters = nScenes * kRegions;
    hipLaunchKernelGGL(k_reset_counters, dim3((nCounters + 255) / 256), dim3(256), 0, st, sceneCounters, nCounters);
    hipLaunchKernelGGL(k_naive_batch, dim3(blocks), dim3(256), 0, st, items, nScenes, sceneCounters);
    hipLaunchKernelGGL(k_accumulate_batch, dim3((unsigned)((maxPixels + 255) / 256), (unsigned)std::min(nScenes, 4096)),
                       dim3(256), 0, st, items, nScenes);
    return hipGetLastError();
}

hipError_t launchTriplePersistent(const DevScene& s, const PassParams& p, float* film, float4* res, int resPasses,
                                  int* pixelCounter, int threshold, int blocks, hipStream_t st)
{
    if (threshold < 0) threshold = 0;
    if (threshold > 63) threshold = 63;
    const int n = s.filmW * s.filmH;
    for (int k0 = 0; k0 < p.nPasses; k0 += resPasses) {
        PassParams q = p;
        q.firstPass = p.firstPass + k0;
        q.freeCounter = p.freeCounter + k0;
        q.nPasses = std::min(resPasses, p.nPasses - k0);
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, pixelCounter);
        hipLaunchKernelGGL(k_triple_persistent, dim3(blocks), dim3(256), 0, st, s, q, res, pixelCounter, threshold);
        hipLaunchKernelGGL(k_triple_accumulate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, res, s.filmW,
                           s.filmH, q.nPasses, film);
    }
    return hipGetLastError();
}

int triplePersistentBlocksPerCU()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_triple_persistent, 256, 0) != hipSuccess) n = 0;
    return residentBlocksPerCU(reinterpret_cast<const void*>(k_triple_persistent), 256, n);
}

int naiveBatchBlocksPerCU()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_naive_batch, 256, 0) != hipSuccess) n = 0;
    return residentBlocksPerCU(reinterpret_cast<const void*>(k_naive_batch), 256, n);
}

int naivePersistentBlocksPerCU()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_naive_persistent<false, false>, 256, 0) != hipSuccess) n = 0;
    return residentBlocksPerCU(reinterpret_cast<const void*>(k_naive_persistent<false, false>), 256, n);
}

}  // namespace zgl
