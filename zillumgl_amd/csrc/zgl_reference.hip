// zgl_reference.hip -- the converged FP64 reference of the naive MIS path tracer.
//
// The estimator of path_integ_naive.glsl:35-174 (the one k_naive_persistent renders),
// evaluated in double by the double instance of the device library (zgl_f64.h): one
// thread per pixel renders the launch's passes in order and adds their samples in FP64,
// skipping NaN samples as the reference's accumulation does (path_integ_naive.glsl:
// 168-173).  Scene data, literals and random numbers are the FP32 ones; every other
// quantity is double, so the per-pixel seeds, every ray, box and triangle test, BSDF,
// pdf and weight are the FP64 values of the same expressions.  It renders the converged
// image that the RMSE of the FP32 renderer is measured against; it runs at a fraction of
// the FP32 kernel's rate (double arithmetic and one thread per pixel).
#include "zgl_f64.h"
#include "zgl_internal.h"

namespace zgl {
namespace f64 {
namespace {

// One sample: the operations of k_naive_persistent's state machine in its order, the
// walks done in place (bvhHit / bvhTest) instead of resumed across ballot rounds.
ZGL_D vec3 naiveSample(const DevScene& s, const PassParams& prm, vec2 scr, Rng& rng)
{
    Counters cn;
    const bool uniform = prm.lightEnvUniform != 0;
    const real portion = prm.lightPortion;
    const vec4 u = sample4D<false>(s, rng, cn);
    const Ray ray = cameraSampleRay(s, scr, u);
    real primDist;
    int id = bvhHit<false>(s, ray, primDist, cn);
    vec3 pos = rayPoint(ray, primDist);
    if (id == -1) return envLe<false>(s, ray.dir, cn);
    if (id - s.objPrimCount >= 0) return lightLe<false>(s, id - s.objPrimCount, pos, -ray.dir, cn);

    vec3 wo = -ray.dir, result = v3(0.0f), throughput = v3(1.0f);
    for (int bounce = 1; bounce <= prm.maxDepth; bounce++) {
        SurfaceInfo surf = triangleSurfaceInfo<false>(s, id, pos, cn);
        const int matTex = s.matTexIndices[id];
        const int matId = matTex & 0x0000ffff;
        const uint32_t matType = loadMaterialType<false>(s, matId, cn);
        if (matType != MatDielectric && matType != MatThinDielectric) {
            if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
        }
        const BSDFParam mp = loadMaterial<false>(s, matType, matId, matTex >> 16, surf.uv, cn);
        if (prm.sampleLight) {
            const real ud = sample1D<false>(s, rng, cn);
            const vec4 us = sample4D<false>(s, rng, cn);
            const real pdfSampleLight = pdfSampleLightOf(s, uniform, portion);
            const bool pickLight = ud < pdfSampleLight;
            const real pdfSelect = pickLight ? pdfSampleLight : 1.0f - pdfSampleLight;
            LightLiSample ls = invalidLi();
            bool valid = false, occluded = false;
            if (pickLight) {
                const int lid = lightSampleOne<false>(s, v2(us.x, us.y), cn);
                const int triId = lid + s.objPrimCount;
                vec3 a, b, c;
                triangleVerts<false>(s, triId, a, b, c, cn);
                const vec3 y = sampleTriangleUniform(a, b, c, v2(us.z, us.w));
                const vec3 wl = normalize(y - pos);
                const vec3 norm = triangleSurfaceInfo<false>(s, triId, y, cn).ng;
                const real cosTheta = dot(norm, -wl);
                if (cosTheta >= 1e-6f) {
                    const real dist = distance(pos, y);
                    real pdf = dist * dist / (triangleArea(a, b, c) * cosTheta);
                    occluded = bvhTest<false>(s, rayOffseted(pos, wl), dist - 1e-4f - 1e-6f, cn);
                    if (!(pdf < 1e-8f)) {
                        const vec3 weight = lightLe<false>(s, lid, y, -wl, cn);
                        pdf *= lightPdfSampleOne<false>(s, lid, cn);
                        ls = LightLiSample{wl, weight / pdf, pdf};
                        valid = true;
                    }
                }
            } else {
                const vec4 sp = envSampleWi<false>(s, us, cn);
                const vec3 we = v3(sp.x, sp.y, sp.z);
                occluded = bvhTest<false>(s, rayOffseted(pos, we), 1e8f, cn);
                if (sp.w != 0.0f) {
                    ls = LightLiSample{we, envLe<false>(s, we, cn) / sp.w, sp.w};
                    valid = true;
                }
            }
            if (valid) {
                ls.coef /= pdfSelect;
                ls.pdf *= pdfSelect;
                if (ls.pdf > 0.0f && !occluded) {
                    const vec3 f = materialBSDF(matType, mp, wo, ls.wi, surf.ns, ModeRadiance);
                    const real fp = materialPdf(matType, mp, wo, ls.wi, surf.ns);
                    const real weight = ls.pdf * ls.pdf / (ls.pdf * ls.pdf + fp * fp);
                    result += f * throughput * satDot(surf.ns, ls.wi) * ls.coef * weight;
                }
            }
        }
        const vec3 u3 = sample3D<false>(s, rng, cn);
        const BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeRadiance, u3, rng);
        const bool delta = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
        if (bs.pdf < 1e-8f) break;
        throughput *= bs.bsdf / bs.pdf * (delta ? real(1) : absDot(surf.ns, bs.wi));
        const Ray next = rayOffseted(pos, bs.wi);
        real dist;
        const int nextId = bvhHit<false>(s, next, dist, cn);
        const int lightId = nextId - s.objPrimCount;
        const vec3 nextPos = rayPoint(next, dist);
        if (nextId == -1) {
            const vec3 radiance = envLe<false>(s, bs.wi, cn);
            real weight = 1.0f;
            if (prm.sampleLight && !delta) {
                const real envPdf = envPdfLi<false>(s, bs.wi, cn) * pdfSelectEnv(s, uniform, portion);
                weight = (envPdf <= 0.0f) ? real(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + envPdf * envPdf);
            }
            result += radiance * throughput * weight;
            break;
        }
        if (lightId >= 0) {
            const vec3 radiance = lightLe<false>(s, lightId, nextPos, -bs.wi, cn);
            real weight = 1.0f;
            if (prm.sampleLight && !delta) {
                const real lightPdf = lightPdfLi<false>(s, lightId, pos, nextPos, cn) *
                                      pdfSelectLight<false>(s, lightId, uniform, portion, cn);
                weight = (lightPdf <= 0.0f) ? real(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + lightPdf * lightPdf);
            }
            result += radiance * throughput * weight;
            break;
        }
        // the next vertex's state is set before the roulette, and the roulette is a select:
        // state a surviving path set after a roulette branch has been compiled wrongly
        // (here, as in the global-queue tracer, the paths that survived it rendered wrong)
        id = nextId;
        pos = nextPos;
        wo = -bs.wi;
        if (prm.russianRoulette) {
            const real continueProb = fminG(maxComponent(bs.bsdf / bs.pdf), 0.95f);
            const bool survive = !(sample1D<false>(s, rng, cn) >= continueProb);
            const vec3 scaled = throughput / continueProb;
            throughput = select3(survive, scaled, throughput);
            if (!survive) break;
        }
    }
    return result;
}

// One thread per pixel of a 16x16 tile; sum holds W*H*3 doubles (RGB), added to.
__global__ __launch_bounds__(256) void k_naive_f64(DevScene s, PassParams prm, double* __restrict__ sum)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= s.filmW || y >= s.filmH) return;
    const int pix = y * s.filmW + x;
    const vec2 scr = v2((real)x, (real)y) / v2((real)s.filmW, (real)s.filmH);
    const PixelSeed ps = pixelSeed(s, scr);
    double r = 0.0, g = 0.0, b = 0.0;
    for (int k = 0; k < prm.nPasses; k++) {
        const int spp = prm.firstPass + k;
        const uint32_t fc = (uint32_t)(prm.freeCounter + k);
        Rng rng;
        rng.randSeed = ps.tx * fc + ps.ty;
        rng.sampleSeed = ps.tx * ps.ty;
        rng.sampleOffset = sampleOffsetOf(spp);
        rng.dim = 0;
        const vec3 c = naiveSample(s, prm, scr, rng);
        if (hasNan(c)) continue;
        r += c.x;
        g += c.y;
        b += c.z;
    }
    sum[3 * (size_t)pix + 0] += r;
    sum[3 * (size_t)pix + 1] += g;
    sum[3 * (size_t)pix + 2] += b;
}

}  // namespace
}  // namespace f64

hipError_t launchNaiveF64(const DevScene& s, const PassParams& p, double* sum, hipStream_t st)
{
    const dim3 grid((unsigned)((s.filmW + 15) / 16), (unsigned)((s.filmH + 15) / 16));
    hipLaunchKernelGGL(f64::k_naive_f64, grid, dim3(256), 0, st, s, p, sum);
    return hipGetLastError();
}

}  // namespace zgl
