// zgl_adjoint.hip -- path-replay adjoint of the naive MIS path tracer.
//
// The reference has no differentiable mode; this file adds one around the estimator of
// path_integ_naive.glsl:35-174.  Parameters: the five differentiable parameters of
// every principled or Lambertian material (baseColor rgb, raw roughness, metallic) and
// the RGB texels of the environment map.
//
// Detached sampling.  Every sampling decision -- directions, BSDF lobe, light or
// environment texel, the pdfs that divide, MIS weights and Russian-roulette
// probabilities -- is taken with the uploaded scene (theta0).  Only the quantities the
// estimator multiplies by are taken at the evaluation point theta: the BSDF values
// f(wo, wi; theta) of differentiable materials and the environment radiance, read
// bilinearly from FP32 texels.  At theta = theta0 the estimator is the forward
// renderer's, bit for bit; for theta != theta0 it is a smooth function of theta with
// the same random numbers, which central differences can check.
//
// Path replay.  Each sample is traced twice with the same random numbers: the first
// pass returns its contribution C, summing its terms also in FP64; the second walks
// the same path again, keeps the part L of C that is still to come (the FP64 sum minus
// the terms met so far: exact, so that L / f stays accurate when f is small), and at
// each vertex adds
//   * for the next-event term N = f(wl) * T * cos * Le / pdf * w:
//       delta . T * cos * Le / pdf * w * df/dtheta      (material of the vertex),
//       delta . f * T * cos / pdf * w * dLe/dtexel       (environment lights),
//   * for the BSDF-sampled continuation, which scales every later term by f(wi):
//       delta . L / f * df/dtheta,
//   * for an environment hit: delta . T * w * dLe/dtexel,
// where delta is dLoss/dframe at the pixel and df/dtheta comes from the dual-number
// BSDF (zgl_dual.h).  Nothing per bounce is stored.  Gradients are accumulated with
// FP64 atomics (per block in LDS for the material parameters, into interleaved copies
// for the texels): contributions of both signs from many samples cancel, which FP32
// sums would not resolve.
#include "zgl_dual.h"
#include "zgl_internal.h"

namespace zgl {

namespace {

constexpr int kLdsMaterials = 256;  // materials whose gradients a block reduces in LDS

// Environment radiance at the evaluation texels (FP32 RGBA), with the bilinear taps.
struct EnvTaps {
    int idx[4];
    float w[4];
};
ZGL_D vec3 envLeEval(const DevScene& s, const float4* env, vec3 wi, EnvTaps& tap)
{
    wi = rotateZcs(wi, s.envRotCosN, s.envRotSinN);
    const vec2 uv = sphereToPlane(wi);
    const int W = s.envW, H = s.envH;
    float u = uv.x * (float)W - 0.5f, v = uv.y * (float)H - 0.5f;
    float fu = ffloorG(u), fv = ffloorG(v);
    float al = u - fu, be = v - fv;
    int i0 = wrapRepeat((int)fu, W), j0 = wrapRepeat((int)fv, H);
    int i1 = wrapRepeat((int)fu + 1, W), j1 = wrapRepeat((int)fv + 1, H);
    tap.idx[0] = j0 * W + i0;
    tap.idx[1] = j0 * W + i1;
    tap.idx[2] = j1 * W + i0;
    tap.idx[3] = j1 * W + i1;
    const float4 a = env[tap.idx[0]], b = env[tap.idx[1]], c = env[tap.idx[2]], d = env[tap.idx[3]];
    const vec3 t00 = v3(a.x, a.y, a.z), t10 = v3(b.x, b.y, b.z), t01 = v3(c.x, c.y, c.z), t11 = v3(d.x, d.y, d.z);
    float w00 = (1.0f - al) * (1.0f - be), w10 = al * (1.0f - be), w01 = (1.0f - al) * be, w11 = al * be;
    tap.w[0] = w00;
    tap.w[1] = w10;
    tap.w[2] = w01;
    tap.w[3] = w11;
    return ((t00 * w00 + t10 * w10) + t01 * w01) + t11 * w11;
}

// Per-block material gradients (scenes with at most kLdsMaterials materials), flushed
// to global memory once per block.  Declared at namespace scope so that every access
// is known to be LDS and compiles to LDS atomics.
__shared__ double sMatGrad[kLdsMaterials * NumMatGrad];

// Environment gradients go to one of several copies of the gradient array, picked by
// the block: next-event estimation concentrates on the brightest texels (a sun), and
// atomics from the whole device on the same few addresses serialise.  The copies are
// summed afterwards (k_sum_copies).
struct Grads {
    bool lds;      // material gradients go to sMatGrad
    double* mat;   // nMaterials * NumMatGrad (global)
    double* env;   // this block's copy of the envW * envH * 3 gradients
};
ZGL_D void addEnvGrad(const Grads& g, const EnvTaps& tap, vec3 wgt)
{
    for (int t = 0; t < 4; t++) {
        if (tap.w[t] == 0.0f) continue;
        double* e = g.env + (size_t)tap.idx[t] * 3;
        atomicAdd(e + 0, (double)(wgt.x * tap.w[t]));
        atomicAdd(e + 1, (double)(wgt.y * tap.w[t]));
        atomicAdd(e + 2, (double)(wgt.z * tap.w[t]));
    }
}
// sum_c wgt_c * df_c/dtheta_k into the material's accumulators
ZGL_D void addMatGrad(const Grads& g, int matId, vec3 wgt, const DV3& f)
{
    for (int k = 0; k < NumMatGrad; k++) {
        const float v = (wgt.x * f.x.d[k] + wgt.y * f.y.d[k]) + wgt.z * f.z.d[k];
        if (v == 0.0f) continue;
        if (g.lds) atomicAdd(&sMatGrad[matId * NumMatGrad + k], (double)v);
        else atomicAdd(g.mat + (size_t)matId * NumMatGrad + k, (double)v);
    }
}

// Principled, and Lambertian (the loader's default for any other type id).
ZGL_D bool differentiable(uint32_t t)
{
    return t != MatMetalWorkflow && t != MatDielectric && t != MatThinDielectric;
}

// BSDF of a differentiable material at the evaluation record, with tangents.
ZGL_D DV3 evalDual(uint32_t t, const float4* evalMat, int matId, vec3 wo, vec3 wi, vec3 n)
{
    const DParam p = dparamOf((const float*)(evalMat + (size_t)matId * 4));
    return (t == MatPrincipled) ? dPrincipledBRDF(wo, wi, n, p) : dLambertian(p);
}

struct Replay {
    const float4* evalMat;  // material records at the evaluation point
    const float4* evalEnv;  // envW * envH FP32 texels (rgb, pad)
};

// One sample of the naive estimator with detached sampling (see the file comment).
// ADJ = false returns the contribution; ADJ = true replays it with `rest` = that
// contribution and accumulates delta-weighted gradients.
// Running sum of a sample's terms, exact: an FP64 sum (x) plus the rounding errors of its
// additions (ex), kept with Knuth's TwoSum.  The replay subtracts the terms as it meets
// them, and the remainder -- the later terms, which a small factor may have made 1e-16 of
// the earlier ones -- must survive that cancellation.
struct Sum3 {
    double x, y, z, ex, ey, ez;
};
ZGL_D void twoSumAdd(double& s, double& e, double b)
{
    const double a = s;
    s = a + b;
    const double bp = s - a;
    e += (a - (s - bp)) + (b - bp);
}
ZGL_D void acc(Sum3& a, vec3 t, double sign)
{
    twoSumAdd(a.x, a.ex, sign * (double)t.x);
    twoSumAdd(a.y, a.ey, sign * (double)t.y);
    twoSumAdd(a.z, a.ez, sign * (double)t.z);
}
ZGL_D double remX(const Sum3& a) { return a.x + a.ex; }
ZGL_D double remY(const Sum3& a) { return a.y + a.ey; }
ZGL_D double remZ(const Sum3& a) { return a.z + a.ez; }

// Traversal results of one sample.  With `pre` (the forward kernel's log of the sample,
// written by launchNaiveLogged), both passes read the results from it.  Without, the
// first pass records every closest-hit query (primitive, distance) and every shadow
// query (occluded) in order, and the replay pass reads them back.  Either way the queries
// are the same, with the same rays, as the walks that produced the log, so are the
// results.  Queries past the log's capacity (paths deeper than the log holds) are walked.
constexpr int kHitLog = kHitLogQueries;
struct HitLog {
    int n;
    const int2* pre;
    int id[kHitLog];
    float dist[kHitLog];
};
template <bool ADJ> ZGL_D int loggedHit(const DevScene& s, const Ray& r, float& dist, HitLog& log)
{
    const int i = log.n++;
    if (log.pre && i < kHitLog) {
        const int2 e = log.pre[i];
        dist = __int_as_float(e.y);
        return e.x;
    }
    if (ADJ && i < kHitLog) {
        dist = log.dist[i];
        return log.id[i];
    }
    Counters cn;
    const int id = bvhHit<false>(s, r, dist, cn);
    if (!ADJ && i < kHitLog) {
        log.id[i] = id;
        log.dist[i] = dist;
    }
    return id;
}
template <bool ADJ> ZGL_D bool loggedTest(const DevScene& s, const Ray& r, float tmax, HitLog& log)
{
    const int i = log.n++;
    if (log.pre && i < kHitLog) return log.pre[i].x != 0;
    if (ADJ && i < kHitLog) return log.id[i] != 0;
    Counters cn;
    const bool occluded = bvhTest<false>(s, r, tmax, cn);
    if (!ADJ && i < kHitLog) log.id[i] = occluded ? 1 : 0;
    return occluded;
}

template <bool ADJ>
ZGL_D vec3 replayPath(const DevScene& s, const Replay& rp, const PassParams& prm, vec2 scr, Rng& rng, Sum3& rest,
                      Sum3& restSkip, vec3 delta, const Grads& gr, HitLog& log)
{
    // first pass: rest accumulates the terms; replay: rest loses them as they are met.
    // restSkip does the same with the terms formed from tpSkip, the throughput whose
    // first vanishing factor per channel is taken as 1.  A BSDF channel vanishes at 0 (a
    // baseColor channel at 0 and no other lobe) or below 1e-20 of the vertex's largest
    // channel (a metal's zero channel lit only by the Fresnel term (1 - c)^5 at wi ~ wo):
    // rest / f cannot recover the terms after such a factor (they are 0 or underflow in
    // FP32), and the derivative through that factor, which does not vanish, needs them.
    const double sign = ADJ ? -1.0 : 1.0;
    vec3 tpSkip = v3(1.0f);
    int zeroSeen = 0;  // bit c: channel c has met its first vanishing factor
    Counters cn;
    const bool uniform = prm.lightEnvUniform != 0;
    const float portion = prm.lightPortion;
    DevScene se = s;  // the same scene with the evaluation material records
    se.materials = rp.evalMat;

    const vec4 u = sample4D<false>(s, rng, cn);
    const Ray ray = cameraSampleRay(s, scr, u);
    float primDist;
    log.n = 0;
    int id = loggedHit<ADJ>(s, ray, primDist, log);
    vec3 pos = rayPoint(ray, primDist);
    if (id == -1) {
        EnvTaps tap;
        const vec3 le = envLeEval(s, rp.evalEnv, ray.dir, tap);
        if (ADJ) addEnvGrad(gr, tap, delta);
        return le;
    }
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
        const bool diff = differentiable(matType);
        const BSDFParam mpE = diff ? loadMaterial<false>(se, matType, matId, matTex >> 16, surf.uv, cn) : mp;

        if (prm.sampleLight) {
            const float ud = sample1D<false>(s, rng, cn);
            const vec4 us = sample4D<false>(s, rng, cn);
            const float pdfSampleLight = pdfSampleLightOf(s, uniform, portion);
            const bool pickLight = ud < pdfSampleLight;
            const float pdfSelect = pickLight ? pdfSampleLight : 1.0f - pdfSampleLight;
            LightLiSample ls = invalidLi();
            bool valid = false;
            EnvTaps tap;
            float envPdf = 0.0f;
            if (pickLight) {
                const int lid = lightSampleOne<false>(s, v2(us.x, us.y), cn);
                const int triId = lid + s.objPrimCount;
                vec3 a, b, c;
                triangleVerts<false>(s, triId, a, b, c, cn);
                const vec3 y = sampleTriangleUniform(a, b, c, v2(us.z, us.w));
                const vec3 wl = normalize(y - pos);
                const vec3 norm = triangleSurfaceInfo<false>(s, triId, y, cn).ng;
                const float cosTheta = dot(norm, -wl);
                if (cosTheta >= 1e-6f) {
                    const float dist = distance(pos, y);
                    float pdf = dist * dist / (triangleArea(a, b, c) * cosTheta);
                    const float test = dist - 1e-4f - 1e-6f;
                    if (!loggedTest<ADJ>(s, rayOffseted(pos, wl), test, log) && !(pdf < 1e-8f)) {
                        const vec3 weight = lightLe<false>(s, lid, y, -wl, cn);
                        pdf *= lightPdfSampleOne<false>(s, lid, cn);
                        ls = LightLiSample{wl, weight / pdf, pdf};
                        valid = true;
                    }
                }
            } else {
                const vec4 sp = envSampleWi<false>(s, us, cn);
                const vec3 we = v3(sp.x, sp.y, sp.z);
                if (!loggedTest<ADJ>(s, rayOffseted(pos, we), 1e8f, log) && sp.w != 0.0f) {
                    envPdf = sp.w;
                    ls = LightLiSample{we, envLeEval(s, rp.evalEnv, we, tap) / sp.w, sp.w};
                    valid = true;
                }
            }
            if (valid) {
                ls.coef /= pdfSelect;
                ls.pdf *= pdfSelect;
                if (ls.pdf > 0.0f) {
                    const vec3 f = materialBSDF(matType, mpE, wo, ls.wi, surf.ns, ModeRadiance);
                    const float fp = materialPdf(matType, mp, wo, ls.wi, surf.ns);
                    const float weight = ls.pdf * ls.pdf / (ls.pdf * ls.pdf + fp * fp);
                    const vec3 term = f * throughput * satDot(surf.ns, ls.wi) * ls.coef * weight;
                    result += term;
                    acc(rest, term, sign);
                    acc(restSkip, f * tpSkip * satDot(surf.ns, ls.wi) * ls.coef * weight, sign);
                    if (ADJ) {
                        if (diff) {
                            const DV3 fd = evalDual(matType, rp.evalMat, matId, wo, ls.wi, surf.ns);
                            addMatGrad(gr, matId, delta * throughput * (satDot(surf.ns, ls.wi) * weight) * ls.coef, fd);
                        }
                        if (!pickLight)
                            addEnvGrad(gr, tap, delta * f * throughput * (satDot(surf.ns, ls.wi) * weight / envPdf / pdfSelect));
                    }
                }
            }
        }

        const vec3 u3 = sample3D<false>(s, rng, cn);
        const BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeRadiance, u3, rng);
        const bool delta_ = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
        if (bs.pdf < 1e-8f) break;
        const vec3 fE = diff ? materialBSDF(matType, mpE, wo, bs.wi, surf.ns, ModeRadiance) : bs.bsdf;
        // vanishing channels of this vertex's factor (see restSkip)
        const float fMax = fmaxG(fabsG(fE.x), fmaxG(fabsG(fE.y), fabsG(fE.z)));
        const bool vx = !(fabsG(fE.x) > 1e-20f * fMax), vy = !(fabsG(fE.y) > 1e-20f * fMax),
                   vz = !(fabsG(fE.z) > 1e-20f * fMax);
        if (ADJ && diff) {
            // every later term carries the factor fE_c; rest holds their exact sum
            // where fE_c vanishes (and no earlier factor of channel c did), the later terms
            // with this factor taken as 1 are what restSkip still holds
            const DV3 fd = evalDual(matType, rp.evalMat, matId, wo, bs.wi, surf.ns);
            const vec3 ratio =
                v3(vx ? ((zeroSeen & 1) ? 0.0f : (float)remX(restSkip)) : (float)(remX(rest) / (double)fE.x),
                   vy ? ((zeroSeen & 2) ? 0.0f : (float)remY(restSkip)) : (float)(remY(rest) / (double)fE.y),
                   vz ? ((zeroSeen & 4) ? 0.0f : (float)remZ(restSkip)) : (float)(remZ(rest) / (double)fE.z));
            addMatGrad(gr, matId, delta * ratio, fd);
        }
        const float geo = delta_ ? 1.0f : absDot(surf.ns, bs.wi);
        {
            const vec3 factor = fE / bs.pdf * geo;
            const float one = 1.0f / bs.pdf * geo;  // the factor with fE_c taken as 1
            const bool zx = vx && !(zeroSeen & 1), zy = vy && !(zeroSeen & 2), zz = vz && !(zeroSeen & 4);
            tpSkip = tpSkip * v3(zx ? one : factor.x, zy ? one : factor.y, zz ? one : factor.z);
            zeroSeen |= (zx ? 1 : 0) | (zy ? 2 : 0) | (zz ? 4 : 0);
        }
        throughput *= fE / bs.pdf * geo;
        const float rrProb = fminG(maxComponent(bs.bsdf / bs.pdf), 0.95f);

        float dist;
        const Ray next = rayOffseted(pos, bs.wi);
        const int nextId = loggedHit<ADJ>(s, next, dist, log);
        const int lightId = nextId - s.objPrimCount;
        const vec3 nextPos = rayPoint(next, dist);
        if (nextId == -1) {
            EnvTaps tap;
            const vec3 radiance = envLeEval(s, rp.evalEnv, bs.wi, tap);
            float weight = 1.0f;
            if (prm.sampleLight && !delta_) {
                float ep = envPdfLi<false>(s, bs.wi, cn) * pdfSelectEnv(s, uniform, portion);
                weight = (ep <= 0.0f) ? 0.0f : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + ep * ep);
            }
            const vec3 term = radiance * throughput * weight;
            result += term;
            acc(rest, term, sign);
            acc(restSkip, radiance * tpSkip * weight, sign);
            if (ADJ) addEnvGrad(gr, tap, delta * throughput * weight);
            break;
        } else if (lightId >= 0) {
            const vec3 radiance = lightLe<false>(s, lightId, nextPos, -bs.wi, cn);
            float weight = 1.0f;
            if (prm.sampleLight && !delta_) {
                float lp = lightPdfLi<false>(s, lightId, pos, nextPos, cn) *
                           pdfSelectLight<false>(s, lightId, uniform, portion, cn);
                weight = (lp <= 0.0f) ? 0.0f : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + lp * lp);
            }
            const vec3 term = radiance * throughput * weight;
            result += term;
            acc(rest, term, sign);
            acc(restSkip, radiance * tpSkip * weight, sign);
            break;
        }
        // the next vertex's state is set before the roulette, and the roulette is written
        // as selects: a surviving path's state assigned after a roulette branch has been
        // compiled wrongly (the global-queue tracer's extension)
        id = nextId;
        pos = nextPos;
        wo = -bs.wi;
        if (prm.russianRoulette) {
            const bool survive = !(sample1D<false>(s, rng, cn) >= rrProb);
            const vec3 tpS = throughput / rrProb, skS = tpSkip / rrProb;
            throughput = select3(survive, tpS, throughput);
            tpSkip = select3(survive, skS, tpSkip);
            if (!survive) break;
        }
    }
    return result;
}

}  // namespace

// MODE 0: the replay estimator's frame (sums added like the forward renderer);
// MODE 1: gradients of sum(dFrame . frame) for the passes of the launch.
template <int MODE>
__global__ __launch_bounds__(256) void k_naive_replay(DevScene s, PassParams prm, Replay rp, float4* __restrict__ frame,
                                                      const float4* __restrict__ dFrame, double* __restrict__ dMat,
                                                      double* __restrict__ dEnv, int envCopies, int passChunk,
                                                      const int2* __restrict__ hitLog)
{
    const bool useLds = MODE == 1 && s.nMaterials <= kLdsMaterials;
    if (useLds) {
        for (int i = threadIdx.x; i < kLdsMaterials * NumMatGrad; i += blockDim.x) sMatGrad[i] = 0.0;
        __syncthreads();
    }
    const int copy = (int)(((blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x) % (unsigned)envCopies);
    const Grads gr{useLds, dMat, dEnv + (size_t)copy * ((size_t)s.envW * s.envH * 3)};
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x < s.filmW && y < s.filmH) {
        const int pix = y * s.filmW + x;
        const vec2 scr = v2((float)x, (float)y) / v2((float)s.filmW, (float)s.filmH);
        const PixelSeed ps = pixelSeed(s, scr);
        float4 mem = frame[pix];
        bool stored = false;
        vec3 delta = v3(0.0f);
        if (MODE == 1) {
            const float4 d = dFrame[pix];
            delta = v3(d.x, d.y, d.z);
        }
        // MODE 1: blockIdx.z selects a chunk of the passes (the gradients are sums over
        // samples, in any order); MODE 0 keeps a pixel's passes in one thread, in order
        const int k0 = MODE == 1 ? (int)blockIdx.z * passChunk : 0;
        const int k1 = MODE == 1 ? min(prm.nPasses, k0 + passChunk) : prm.nPasses;
        HitLog log;
        log.pre = nullptr;
        for (int k = k0; k < k1; k++) {
            if (hitLog) log.pre = hitLog + ((size_t)k * ((size_t)s.filmW * s.filmH) + pix) * kHitLog;
            const int spp = prm.firstPass + k;
            const uint32_t fc = (uint32_t)(prm.freeCounter + k);
            Rng rng;
            rng.randSeed = ps.tx * fc + ps.ty;
            rng.sampleSeed = ps.tx * ps.ty;
            rng.sampleOffset = sampleOffsetOf(spp);
            rng.dim = 0;
            const Rng start = rng;
            Sum3 rest = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, restSkip = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            const vec3 c = replayPath<false>(s, rp, prm, scr, rng, rest, restSkip, v3(0.0f), gr, log);
            if (hasNan(c)) continue;
            if (MODE == 0) {
                const vec3 last = (spp == 0) ? v3(0.0f) : v3(mem.x, mem.y, mem.z);
                const vec3 sum = last + c;
                mem = make_float4(sum.x, sum.y, sum.z, 1.0f);
                stored = true;
            } else if (delta.x != 0.0f || delta.y != 0.0f || delta.z != 0.0f) {
                rng = start;
                replayPath<true>(s, rp, prm, scr, rng, rest, restSkip, delta, gr, log);
            }
        }
        if (MODE == 0 && stored) frame[pix] = mem;
    }
    if (useLds) {
        __syncthreads();
        for (int i = threadIdx.x; i < s.nMaterials * NumMatGrad; i += blockDim.x)
            if (sMatGrad[i] != 0.0) atomicAdd(dMat + i, sMatGrad[i]);
    }
}

// FP16 environment texels to FP32 (the conversion envTexel applies), the default
// evaluation point.
__global__ void k_env_expand(const ushort4* __restrict__ env, float4* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const ushort4 t = env[i];
    out[i] = make_float4((float)__builtin_bit_cast(_Float16, (unsigned short)t.x),
                         (float)__builtin_bit_cast(_Float16, (unsigned short)t.y),
                         (float)__builtin_bit_cast(_Float16, (unsigned short)t.z), 0.0f);
}

hipError_t launchEnvExpand(const ushort4* env, float4* out, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_env_expand, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, env, out, n);
    return hipGetLastError();
}

// dst[i] = sum over c of src[c * n + i], in FP64.
__global__ void k_sum_copies(const double* __restrict__ src, double* __restrict__ dst, int64_t n, int copies)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    for (int c = 0; c < copies; c++) v += src[(size_t)c * n + i];
    dst[i] = v;
}

constexpr int kPassChunk = 8;

hipError_t launchNaiveReplay(const DevScene& s, const PassParams& p, const float4* evalMat, const float4* evalEnv,
                             float4* frame, const float4* dFrame, double* dMat, double* dEnv, int envCopies,
                             double* envSum, const int2* hitLog, hipStream_t st)
{
    const Replay rp{evalMat, evalEnv};
    const dim3 grid((s.filmW + 15) / 16, (s.filmH + 15) / 16);
    if (!dFrame) {
        hipLaunchKernelGGL(k_naive_replay<0>, grid, dim3(256), 0, st, s, p, rp, frame, dFrame, dMat, dEnv, 1,
                           p.nPasses, hitLog);
        return hipGetLastError();
    }
    if (envCopies < 1) envCopies = 1;
    // passes in chunks of kPassChunk: more, shorter threads balance the waves
    const int chunks = (p.nPasses + kPassChunk - 1) / kPassChunk;
    const dim3 gridAdj(grid.x, grid.y, (unsigned)chunks);
    hipLaunchKernelGGL(k_naive_replay<1>, gridAdj, dim3(256), 0, st, s, p, rp, frame, dFrame, dMat, dEnv, envCopies,
                       kPassChunk, hitLog);
    const int64_t n = (int64_t)s.envW * s.envH * 3;
    hipLaunchKernelGGL(k_sum_copies, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dEnv, envSum, n,
                       envCopies);
    return hipGetLastError();
}

}  // namespace zgl
