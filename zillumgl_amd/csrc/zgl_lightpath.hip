// zgl_lightpath.hip -- the light-path kernels on the persistent, regenerating wave64
// schedule: the light tracer (TRIPLE = false, light_path_integ.glsl:45-159) and the
// triple tracer's light pass (TRIPLE = true, triple_path_pass_lpt.glsl:58-193).  k_light
// and k_triple_lpt (one thread per unit) stay as their byte-counting variants.
//
// A light path is sampled on an emissive triangle, connected to the camera at every
// vertex (and, for the light tracer, at the light point itself) by a visibility test and
// a splat into the planar film, and extended by BSDF sampling.  Here each lane runs a
// path as a state machine whose states end by issuing one walk -- a connection's
// visibility walk or the extension walk -- and the walks advance a few MTBVH rows per
// round for the whole wave, as in the naive tracer's persistent kernel:
//   * a connection's contribution is computed before its visibility walk and splatted
//     after it if the walk finds no occluder;
//   * the random numbers the reference draws after a visibility test (the emitted ray at
//     the light point; the BSDF sample and the roulette number at a vertex) are drawn
//     before the walk in the same order, since a walk draws none;
//   * a unit is (pass, thread id) of the reference's dispatch: one path for the light
//     tracer, loopsPerPass paths drawn in turn from one random stream for the triple
//     light pass; a lane whose unit is done takes the next, one atomic per wave;
//   * a lane whose connection walk ends splats and starts its extension walk inside the
//     walk loop.
// The roulette is written as selects (survival as a flag, the division always
// evaluated): written as a branch, the compiler mishandled a surviving path's state in
// the global-queue tracer.
#include "zgl_internal.h"

#include <algorithm>

namespace zgl {

namespace {

enum : int {
    kNewUnit = 0,  // the lane needs a (pass, id) unit
    kStart,        // a path of the lane's unit is to be started
    kConnect,      // a camera connection's visibility walk
    kExtend,       // the extension walk to the next vertex
    kDone,
};

constexpr int kRowsPerRound = 4;

__global__ void k_reset_light_counter(int* c) { if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0; }

// visible(x, y) (intersection.glsl:429-434) as the walk it issues.
__device__ __forceinline__ void visibilityRay(vec3 x, vec3 y, Ray& ray, float& dist)
{
    dist = distance(x, y) - 2e-5f;
    const vec3 w = normalize(y - x);
    ray = makeRay(x + w * 1e-5f, w);
}

template <bool TRIPLE>
__global__ __launch_bounds__(256, 4) void k_lightpath_persistent(DevScene s0, PassParams prm, float* __restrict__ film,
                                                              int* __restrict__ counter, int threshold)
{
    const int lane = threadIdx.x & 63;
    Counters cn;
    const int nUnits = prm.pathsPerPass * prm.nPasses;
    const int loops = TRIPLE ? prm.loopsPerPass : 1;

    int phase = kNewUnit;
    int loopsLeft = 0;
    bool walking = false;
    Walk wk;
    wk.p.o = v3(0.f); wk.p.d = v3(0.f); wk.p.dInv = v3(0.f); wk.p.mode = 0;
    wk.p.nx = wk.p.ny = wk.p.nz = false; wk.p.table = 0;
    wk.dist = 0.f; wk.k = 0; wk.closest = -1; wk.any = false;
    Rng rng = {0u, 0u, 0, 0};
    vec3 throughput = v3(0.f), wo = v3(0.f);
    Ray next = makeRay(v3(0.f), v3(0.f));  // the extension ray issued after the connection
    int bounce = 0;
    bool cont = false;                      // the path goes on after its connection
    vec3 pendRes = v3(0.f);
    vec2 pendUv = v2(0.f, 0.f);
    bool pendValid = false;
    // triple light pass: MIS state (triple_path_pass_lpt.glsl:58-80)
    vec3 prevNorm = v3(0.f);
    float prevPdfDir = 0.f, s0t1 = 0.f, s1t1 = 0.f;

    for (;;) {
        const DevScene& s = sceneArg(s0);
        auto endPath = [&]() { phase = loopsLeft > 0 ? kStart : kNewUnit; };
        // after a connection walk: the splat if unoccluded, then the extension walk or the end
        auto settle = [&]() {
            if (wk.closest < 0 && pendValid) splatFilm(s, film, pendUv, pendRes);
            if (cont) {
                walkStart<false>(s, wk, next, false, 1e8f, cn);
                walking = true;
                phase = kExtend;
            } else {
                endPath();
            }
        };
        // the walk a state ends with: the connection if there is one, else the extension
        auto issue = [&](bool connect, const Ray& vis, float visDist) {
            if (connect) {
                walkStart<false>(s, wk, vis, true, visDist, cn);
                walking = true;
                phase = kConnect;
            } else if (cont) {
                walkStart<false>(s, wk, next, false, 1e8f, cn);
                walking = true;
                phase = kExtend;
            } else {
                endPath();
            }
        };
        // a new path of the lane's unit, from its random stream
        auto startPath = [&]() {
            loopsLeft--;
            const int light = lightSampleOne<false>(s, sample2D<false>(s, rng, cn), cn);
            const float pdfSource = lightPdfSampleOne<false>(s, light, cn);
            const int triId = light + s.objPrimCount;
            bool connect = false;
            Ray vis = makeRay(v3(0.f), v3(0.f));
            float visDist = 0.f;
            pendValid = false;
            if (TRIPLE) {
                // the light point is drawn and not used (triple_path_pass_lpt.glsl:64)
                (void)triangleSampleUniformId<false>(s, triId, sample2D<false>(s, rng, cn), cn);
            } else {
                // the light point's own connection (light_path_integ.glsl:56-70)
                const vec3 pLit = triangleSampleUniformId<false>(s, triId, sample2D<false>(s, rng, cn), cn);
                const CameraIiSample ci = cameraSampleIi(s, pLit, sample2D<false>(s, rng, cn));
                if (ci.pdf > 0.0f) {
                    connect = true;
                    const vec3 pCam = pLit + ci.wi * ci.dist;
                    const float pdfPos = 1.0f / triangleAreaId<false>(s, triId, cn);
                    visibilityRay(pLit, pCam, vis, visDist);
                    const vec3 Le = lightLe<false>(s, light, pLit, ci.wi, cn);
                    pendRes = Le * ci.Ii / (ci.pdf * pdfPos * pdfSource);
                    pendUv = ci.uv;
                    pendValid = !isBlack(pendRes);
                }
            }
            const LightLeSample le = lightSampleOneLe<false>(s, light, sample4D<false>(s, rng, cn), cn);
            const vec3 nl = triangleSurfaceInfo<false>(s, triId, le.ray.ori, cn).ng;
            wo = -le.ray.dir;
            next = rayOffseted(le.ray.ori, le.ray.dir);
            throughput = le.Le * absDot(nl, -wo) / (pdfSource * le.pdfPos * le.pdfDir);
            if (TRIPLE) {
                prevNorm = nl;
                prevPdfDir = le.pdfDir;
                s0t1 = 1.0f / remapP(le.pdfPos * pdfSource);
                s1t1 = 1.0f;
            }
            bounce = 1;
            cont = 1 <= prm.maxDepth;
            issue(connect, vis, visDist);
        };

        if (!walking) {
            if (phase == kConnect) {
                settle();
            } else if (phase == kExtend) {
                // the vertex the extension walk found
                const int id = wk.closest;
                if (id == -1 || id - s.objPrimCount >= 0) {
                    endPath();
                } else {
                    const float dist = wk.dist;
                    const vec3 pos = rayPoint(makeRay(wk.p.o, wk.p.d), dist);
                    SurfaceInfo surf = triangleSurfaceInfo<false>(s, id, pos, cn);
                    const int matTex = s.matTexIndices[id];
                    const int matId = matTex & 0x0000ffff;
                    const uint32_t matType = loadMaterialType<false>(s, matId, cn);
                    if (matType != MatDielectric && matType != MatThinDielectric) {
                        if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
                    }
                    if (TRIPLE) {
                        const float coefToPos = remapP(prevPdfDir * absDot(surf.ns, wo));
                        s0t1 /= coefToPos;
                        s1t1 /= coefToPos / (bounce == 1 ? remapP(dist * dist) : 1.0f);
                    }
                    const BSDFParam mp = loadMaterial<false>(s, matType, matId, matTex >> 16, surf.uv, cn);
                    const CameraIiSample ci = cameraSampleIi(s, pos, sample2D<false>(s, rng, cn));
                    bool connect = false;
                    Ray vis = makeRay(v3(0.f), v3(0.f));
                    float visDist = 0.f;
                    pendValid = false;
                    if (ci.pdf > 0.0f) {
                        connect = true;
                        const vec3 pCam = pos + ci.wi * ci.dist;
                        visibilityRay(pos, pCam, vis, visDist);
                        const float cosWi = satDot(surf.ng, ci.wi) * fabsG(dot(surf.ns, wo) / dot(surf.ng, wo));
                        const vec3 f = materialBSDF(matType, mp, wo, ci.wi, surf.ns, ModeImportance);
                        vec3 res = ci.Ii * f * throughput * cosWi / ci.pdf;
                        if (TRIPLE) {
                            // the t=1 strategy against s=0 and s=1 (triple_path_pass_lpt.glsl:118-135)
                            const float coefToSurf =
                                remapP(cameraPdfIe(s, makeRay(pCam, -ci.wi)).pdfDir * satDot(surf.ns, ci.wi));
                            const float coefToPrev =
                                remapP(materialPdf(matType, mp, ci.wi, wo, surf.ns) * absDot(prevNorm, wo));
                            const float coefDist = remapP(ci.dist * ci.dist);
                            const float coef0 = coefToSurf * coefToPrev / coefDist;
                            const float coef1 = ((bounce == 1) ? 1.0f : coefToPrev) * coefToSurf / coefDist;
                            const float weight = 1.0f / (s0t1 * coef0 + s1t1 * coef1 + 1.0f);
                            res = res * weight;
                        }
                        pendValid = !hasNan(res) && !fisnan(ci.pdf) && ci.pdf > 1e-8f && !isBlack(res);
                        pendRes = TRIPLE ? res * prm.lptScale : res;
                        pendUv = ci.uv;
                    }
                    const BSDFSample bs =
                        materialSample(matType, mp, surf.ns, wo, ModeImportance, sample3D<false>(s, rng, cn), rng);
                    const vec3 wi = bs.wi;
                    const float bsdfPdf = bs.pdf;
                    const bool deltaBsdf = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
                    bool alive = !(bsdfPdf < 1e-8f || fisnan(bsdfPdf));
                    if (prm.russianRoulette) {
                        // the reference draws this number only after a valid BSDF sample;
                        // drawn either way here, it changes nothing for a path ending here
                        const float continueProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 1.0f);
                        const bool survive = !(sample1D<false>(s, rng, cn) >= continueProb);
                        const vec3 scaled = throughput / continueProb;
                        throughput = select3(survive, scaled, throughput);
                        alive = alive && survive;
                    }
                    if (TRIPLE) {
                        const float coefToPrev = remapP(materialPdf(matType, mp, wi, wo, surf.ns) * absDot(prevNorm, wo));
                        s0t1 *= coefToPrev;
                        s1t1 *= (bounce == 1) ? 1.0f : coefToPrev;
                        prevPdfDir = materialPdf(matType, mp, wo, wi, surf.ns);
                        prevNorm = surf.ns;
                    }
                    const float cosWi =
                        deltaBsdf ? 1.0f : fabsG(dot(surf.ng, wi) * dot(surf.ns, wo) / dot(surf.ng, wo));
                    throughput *= bs.bsdf * cosWi / bsdfPdf;
                    next = rayOffseted(pos, wi);
                    wo = -wi;
                    cont = alive && bounce + 1 <= prm.maxDepth;
                    bounce++;
                    issue(connect, vis, visDist);
                }
            }
        }
        while (!walking && phase == kStart) startPath();

        // new units, one atomic per wave for the lanes that need one
        const unsigned long long need = __ballot(!walking && phase == kNewUnit);
        if (need) {
            const int leader = __ffsll((long long)need) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(counter, __popcll(need));
            base = __shfl(base, leader);
            if (!walking && phase == kNewUnit) {
                const int idx = base + __popcll(need & ((1ull << lane) - 1ull));
                if (idx >= nUnits) {
                    phase = kDone;
                } else {
                    const int k = idx / prm.pathsPerPass;
                    const uint32_t pid = (uint32_t)(idx % prm.pathsPerPass);
                    const int spp = prm.firstPass + k;
                    rng.randSeed = (uint32_t)spp * (uint32_t)(prm.pathsPerPass * loops) + pid +
                                   (uint32_t)(prm.freeCounter + k);
                    rng.sampleSeed = 0;
                    rng.sampleOffset = sampleOffsetOf(spp);
                    rng.dim = 0;
                    loopsLeft = loops;
                    phase = kStart;
                }
            }
        }
        while (!walking && phase == kStart) startPath();
        if (__ballot(phase != kDone) == 0ull) break;

        // walking: up to kRowsPerRound rows per lane per round, leaves tested together once
        // per round; a finished connection walk splats and starts its extension walk here
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
            if (!walking && phase == kConnect) settle();
            general = __ballot(walking && wk.p.special) == 0ull;
            if (__popcll(__ballot(walking)) <= threshold) break;
        }
    }
}

}  // namespace

int lightPersistentBlocksPerCU(bool triple)
{
    int n = 0;
    const hipError_t e =
        triple ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_lightpath_persistent<true>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_lightpath_persistent<false>, 256, 0);
    if (e != hipSuccess) n = 0;
    return triple ? residentBlocksPerCU(reinterpret_cast<const void*>(k_lightpath_persistent<true>), 256, n)
                  : residentBlocksPerCU(reinterpret_cast<const void*>(k_lightpath_persistent<false>), 256, n);
}

hipError_t launchLightPersistent(const DevScene& s, const PassParams& p, bool triple, float* film, int* counter,
                                 int threshold, int blocks, hipStream_t st)
{
    if (threshold < 0) threshold = 0;
    if (threshold > 63) threshold = 63;
    // units are counted in int32 and each wave may take 64 past the end: chunks of passes
    const int chunk = std::max(1, (int)std::min<int64_t>(p.nPasses, ((int64_t)1 << 30) / p.pathsPerPass));
    for (int k0 = 0; k0 < p.nPasses; k0 += chunk) {
        PassParams q = p;
        q.firstPass = p.firstPass + k0;
        q.freeCounter = p.freeCounter + k0;
        q.nPasses = std::min(chunk, p.nPasses - k0);
        hipLaunchKernelGGL(k_reset_light_counter, dim3(1), dim3(64), 0, st, counter);
        if (triple)
            hipLaunchKernelGGL(k_lightpath_persistent<true>, dim3(blocks), dim3(256), 0, st, s, q, film, counter,
                               threshold);
        else
            hipLaunchKernelGGL(k_lightpath_persistent<false>, dim3(blocks), dim3(256), 0, st, s, q, film, counter,
                               threshold);
    }
    return hipGetLastError();
}

}  // namespace zgl
