// zgl_kernels.hip -- rendering kernels for gfx950.
//
// Kernels mirror the reference's compute shaders:
//   k_trace_closest   bvhHit over a batch of rays (intersection.glsl:395-427)
//   (the naive path tracer's persistent kernel is in zgl_persistent.hip)
//   k_light           light_path_integ.glsl:34-159
//   k_triple_pt       triple_path_pass_pt.glsl:44-224
//   k_triple_lpt      triple_path_pass_lpt.glsl:36-193
//   k_film_to_frame   util/img_copy_1x32f_4x32f.glsl
//   k_post_process    post_proc.glsl + the RGB8 read-back of captureImage
// A launch may render several consecutive passes; per-pixel sums are then carried in
// registers in pass order, which is the same sequence of float additions as one
// launch per pass.
#include "zgl_internal.h"
#include "zgl_post.h"

namespace zgl {

namespace {

constexpr int kBlock = 256;

// Block index remapped so that each XCD (blocks are dealt round-robin over the 8
// XCDs) receives a contiguous run of tiles and shares its L2 between neighbours.
__device__ __forceinline__ int xcdTile(int b, int nb)
{
    const int per = (nb + 7) / 8;
    const int xcd = b & 7, slot = b >> 3;
    int t = xcd * per + slot;
    return t;
}

// 16x16 pixel tile per 256-thread block; each wave64 covers an 8x8 sub-tile.
__device__ __forceinline__ void tilePixel(int tile, int tilesX, int& x, int& y)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    x = (tile % tilesX) * 16 + (wave & 1) * 8 + (lane & 7);
    y = (tile / tilesX) * 16 + (wave >> 1) * 8 + (lane >> 3);
}

template <bool C> __device__ void flushCounters(const Counters& cn, unsigned long long* out)
{
    if (!C) return;
    atomicAdd(out + 0, (unsigned long long)cn.bytes);
    atomicAdd(out + 1, (unsigned long long)cn.rows);
    atomicAdd(out + 2, (unsigned long long)cn.tests);
    atomicAdd(out + 3, (unsigned long long)cn.traversals);
}

}  // namespace

__global__ void k_sobol_table(const uint32_t* __restrict__ M, uint32_t* __restrict__ tab, int first, int count)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= count * 256) return;
    uint32_t i = (uint32_t)first + (uint32_t)(g / 256);
    const int d = g % 256;
    const uint32_t* m = M + d * 32;
    uint32_t r = 0;
    for (int b = 0; i != 0u && b < 32; i >>= 1, b++)
        if (i & 1u) r ^= m[b];
    tab[g] = r;
}

// sRGB8 texels to linear FP32 through the 256-entry decode table.
__global__ void k_srgb_expand(const uint8_t* __restrict__ rgb, const float* __restrict__ lut, float4* __restrict__ out,
                              int64_t n)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    out[g] = make_float4(lut[rgb[3 * g]], lut[rgb[3 * g + 1]], lut[rgb[3 * g + 2]], 0.0f);
}

__global__ void k_fill(float* p, int64_t n, float v)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; g < n; g += (int64_t)gridDim.x * blockDim.x) p[g] = v;
}

// post_proc.glsl:43-58: frame sum -> display colour (RGBA32F, alpha 1) and/or RGB8
// rows flipped top-down (captureImage: stbi_flip_vertically_on_write).
__global__ void k_post_process(const float4* __restrict__ frame, float4* __restrict__ out,
                               unsigned char* __restrict__ rgb8, int w, int h, float scale, int tone)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= w * h) return;
    const int x = g % w, y = g / w;
    const float4 f = frame[g];
    const vec3 m = displayColor(v3(f.x, f.y, f.z), scale, tone);
    if (out) out[g] = make_float4(m.x, m.y, m.z, 1.0f);
    if (rgb8) {
        unsigned char* o = rgb8 + ((size_t)(h - 1 - y) * w + x) * 3;
        o[0] = toUnorm8(m.x);
        o[1] = toUnorm8(m.y);
        o[2] = toUnorm8(m.z);
    }
}

__global__ void k_film_to_frame(const float* __restrict__ film, float4* __restrict__ frame, int w, int h)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= w * h) return;
    int x = g % w, y = g / w;
    const float* f = film + (size_t)y * 3 * w + 3 * x;
    frame[g] = make_float4(f[0], f[1], f[2], 1.0f);
}

// ---------------------------------------------------------------------------
// Closest hit over a ray batch.  Rays are SoA: ox[n], oy[n], oz[n], dx[n], dy[n], dz[n].
// Outputs the primitive (-1 on miss), t (1e8 on miss, as bvhHit's dist) and a
// barycentric bin = 16 * min(15, int(16 u/det)) + min(15, int(16 v/det)).
// ---------------------------------------------------------------------------
template <bool C>
__global__ __launch_bounds__(kBlock) void k_trace_closest(DevScene s, const float* __restrict__ rays, int64_t n,
                                                          int32_t* __restrict__ primOut, float* __restrict__ tOut,
                                                          uint32_t* __restrict__ binOut,
                                                          unsigned long long* counters)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    Counters cn;
    if (i < n) {
        Ray r;
        r.ori = v3(rays[i], rays[n + i], rays[2 * n + i]);
        r.dir = v3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        float dist;
        int prim = bvhHit<C>(s, r, dist, cn);
        uint32_t bin = 0xffffffffu;
        if (prim >= 0) {
            vec3 a, b, c;
            loadTri(s, prim, a, b, c);
            float t = 0.0f, u = 0.0f, v = 0.0f, det = 1.0f;
            intersectTriangle(a, b, c, r.ori, r.dir, t, &u, &v, &det);
            int bu = (int)(16.0f * (u / det)), bv = (int)(16.0f * (v / det));
            bu = bu > 15 ? 15 : (bu < 0 ? 0 : bu);
            bv = bv > 15 ? 15 : (bv < 0 ? 0 : bv);
            bin = (uint32_t)(bu * 16 + bv);
        }
        primOut[i] = prim;
        tOut[i] = dist;
        binOut[i] = bin;
    }
    flushCounters<C>(cn, counters);
}

__global__ __launch_bounds__(kBlock) void k_trace_any(DevScene s, const float* __restrict__ rays,
                                                      const float* __restrict__ tmax, int64_t n,
                                                      uint8_t* __restrict__ occ)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Counters cn;
    Ray r;
    r.ori = v3(rays[i], rays[n + i], rays[2 * n + i]);
    r.dir = v3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
    occ[i] = bvhTest<false>(s, r, tmax[i], cn) ? 1 : 0;
}

// light_path_integ.glsl:45-146
template <bool C>
__device__ void lightTrace(const DevScene& s, const PassParams& prm, Rng& rng, float* film, Counters& cn)
{
    int light = lightSampleOne<C>(s, sample2D<C>(s, rng, cn), cn);
    float pdfSource = lightPdfSampleOne<C>(s, light, cn);
    Ray ray;
    vec3 wo, throughput;
    {
        int triId = light + s.objPrimCount;
        vec3 pLit = triangleSampleUniformId<C>(s, triId, sample2D<C>(s, rng, cn), cn);
        CameraIiSample ci = cameraSampleIi(s, pLit, sample2D<C>(s, rng, cn));
        if (ci.pdf > 0.0f) {
            vec3 pCam = pLit + ci.wi * ci.dist;
            float pdfPos = 1.0f / triangleAreaId<C>(s, triId, cn);
            if (visible<C>(s, pLit, pCam, cn)) {
                vec3 Le = lightLe<C>(s, light, pLit, ci.wi, cn);
                vec3 contrib = Le * ci.Ii / (ci.pdf * pdfPos * pdfSource);
                if (!isBlack(contrib)) { splatFilm(s, film, ci.uv, contrib); cnt<C>(cn, 12); }
            }
        }
        LightLeSample le = lightSampleOneLe<C>(s, light, sample4D<C>(s, rng, cn), cn);
        vec3 nl = triangleSurfaceInfo<C>(s, triId, le.ray.ori, cn).ng;
        wo = -le.ray.dir;
        ray = rayOffseted(le.ray.ori, le.ray.dir);
        throughput = le.Le * absDot(nl, -wo) / (pdfSource * le.pdfPos * le.pdfDir);
    }
    for (int bounce = 1; bounce <= prm.maxDepth; bounce++) {
        float dist;
        int id = bvhHit<C>(s, ray, dist, cn);
        if (id == -1) break;
        if (id - s.objPrimCount >= 0) break;
        vec3 pos = rayPoint(ray, dist);
        SurfaceInfo surf = triangleSurfaceInfo<C>(s, id, pos, cn);
        const int matTex = s.matTexIndices[id];
        const int matId = matTex & 0x0000ffff;
        cnt<C>(cn, 4);
        uint32_t matType = loadMaterialType<C>(s, matId, cn);
        if (matType != MatDielectric && matType != MatThinDielectric) {
            if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
        }
        BSDFParam mp = loadMaterial<C>(s, matType, matId, matTex >> 16, surf.uv, cn);
        {
            CameraIiSample ci = cameraSampleIi(s, pos, sample2D<C>(s, rng, cn));
            if (ci.pdf > 0.0f) {
                vec3 pCam = pos + ci.wi * ci.dist;
                if (visible<C>(s, pos, pCam, cn)) {
                    vec3 f = materialBSDF(matType, mp, wo, ci.wi, surf.ns, ModeImportance);
                    float cosWi = satDot(surf.ng, ci.wi) * fabsG(dot(surf.ns, wo) / dot(surf.ng, wo));
                    vec3 res = ci.Ii * f * throughput * cosWi / ci.pdf;
                    if (!hasNan(res) && !fisnan(ci.pdf) && ci.pdf > 1e-8f && !isBlack(res)) {
                        splatFilm(s, film, ci.uv, res);
                        cnt<C>(cn, 12);
                    }
                }
            }
        }
        BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeImportance, sample3D<C>(s, rng, cn), rng);
        vec3 wi = bs.wi;
        float bsdfPdf = bs.pdf;
        bool deltaBsdf = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
        if (bsdfPdf < 1e-8f || fisnan(bsdfPdf)) break;
        if (prm.russianRoulette) {
            float continueProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 1.0f);
            if (sample1D<C>(s, rng, cn) >= continueProb) break;
            throughput /= continueProb;
        }
        float cosWi = deltaBsdf ? 1.0f : fabsG(dot(surf.ng, wi) * dot(surf.ns, wo) / dot(surf.ng, wo));
        throughput *= bs.bsdf * cosWi / bsdfPdf;
        ray = rayOffseted(pos, wi);
        wo = -wi;
    }
}

template <bool C>
__global__ __launch_bounds__(kBlock) void k_light(DevScene s, PassParams prm, float* __restrict__ film,
                                                  unsigned long long* counters)
{
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    Counters cn;
    if (g < (int64_t)prm.pathsPerPass * prm.nPasses) {
        const int k = (int)(g / prm.pathsPerPass);
        const uint32_t id = (uint32_t)(g % prm.pathsPerPass);
        const int spp = prm.firstPass + k;
        Rng rng;
        rng.randSeed = (uint32_t)spp * (uint32_t)prm.pathsPerPass + id + (uint32_t)(prm.freeCounter + k);
        rng.sampleSeed = 0;
        rng.sampleOffset = sampleOffsetOf(spp);
        rng.dim = 0;
        lightTrace<C>(s, prm, rng, film, cn);
    }
    flushCounters<C>(cn, counters);
}

// ---------------------------------------------------------------------------
// Triple path tracer, camera pass (triple_path_pass_pt.glsl:44-195): s=0 and s=1
// strategies weighted against t=1 with the p^2 heuristic.
// ---------------------------------------------------------------------------

template <bool C>
__device__ vec3 tripleCameraTrace(const DevScene& s, const PassParams& prm, Ray ray, Rng& rng, Counters& cn)
{
    float primDist;
    int id = bvhHit<C>(s, ray, primDist, cn);
    vec3 pos = rayPoint(ray, primDist);
    if (id == -1) return envLe<C>(s, ray.dir, cn);
    if (id - s.objPrimCount >= 0) return lightLe<C>(s, id - s.objPrimCount, pos, -ray.dir, cn);

    SurfaceInfo surf = triangleSurfaceInfo<C>(s, id, pos, cn);
    vec3 wo = -ray.dir;
    vec3 prevNorm = s.camF;
    vec3 result = v3(0.0f);
    vec3 throughput = v3(1.0f);

    CameraPdf camPdf = cameraPdfIe(s, ray);
    float primaryPdf = remapP(camPdf.pdfPos) / remapP(camPdf.pdfDir * absDot(surf.ns, ray.dir) / square(primDist));
    float t1s0 = primaryPdf, t1s1 = primaryPdf;

    for (int bounce = 1; bounce <= prm.maxDepth; bounce++) {
        if (bounce > 1) surf = triangleSurfaceInfo<C>(s, id, pos, cn);
        const int matTex = s.matTexIndices[id];
        const int matId = matTex & 0x0000ffff;
        cnt<C>(cn, 4);
        uint32_t matType = loadMaterialType<C>(s, matId, cn);
        if (matType != MatDielectric && matType != MatThinDielectric) {
            if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
        }
        BSDFParam mp = loadMaterial<C>(s, matType, matId, matTex >> 16, surf.uv, cn);
        {
            int light = lightSampleOne<C>(s, sample2D<C>(s, rng, cn), cn);
            int triId = light + s.objPrimCount;
            float pdfSource = lightPdfSampleOne<C>(s, light, cn);
            vec3 pLit = triangleSampleUniformId<C>(s, triId, sample2D<C>(s, rng, cn), cn);
            vec3 wi = normalize(pLit - pos);
            vec3 Le = lightLe<C>(s, light, pLit, -wi, cn);
            if (!isBlack(Le) && visible<C>(s, pos, pLit, cn)) {
                vec3 nLit = triangleSurfaceInfo<C>(s, triId, pLit, cn).ng;
                float pA = pdfSource / triangleAreaId<C>(s, triId, cn);
                float dist2 = distSquare(pos, pLit);
                float pS = pA * dist2 / absDot(nLit, wi);
                vec3 f = materialBSDF(matType, mp, wo, wi, surf.ns, ModeRadiance);
                float fpdf = materialPdf(matType, mp, wo, wi, surf.ns);
                float pdfRev = materialPdf(matType, mp, wi, wo, surf.ns);
                float pdfPLit = remapP(pA);
                float coefToSurf = remapP(0.5f * PiInv * absDot(surf.ns, wi));
                float coefToLight = remapP(fpdf * satDot(nLit, -wi));
                float coefToPrev = (bounce == 1) ? 1.0f : remapP(pdfRev * absDot(prevNorm, wo));
                float coefDist = remapP(dist2);
                float s1s0 = pdfPLit * coefDist / coefToLight;
                float t1s1x = t1s1 * coefToSurf * coefToPrev / coefDist;
                float weight = s1s0 / (1.0f + s1s0 + s1s0 * t1s1x);
                result += Le * f * throughput * absDot(surf.ns, wi) / pS * weight;
            }
        }
        BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeRadiance, sample3D<C>(s, rng, cn), rng);
        vec3 wi = bs.wi;
        float bsdfPdf = bs.pdf;
        bool deltaBsdf = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
        if (bsdfPdf < 1e-8f) break;
        throughput *= bs.bsdf / bsdfPdf * (deltaBsdf ? 1.0f : absDot(surf.ns, wi));

        Ray nextRay = rayOffseted(pos, wi);
        float dist;
        int nextId = bvhHit<C>(s, nextRay, dist, cn);
        int lightId = nextId - s.objPrimCount;
        vec3 nextPos = rayPoint(nextRay, dist);
        float pdfDirToNext = materialPdf(matType, mp, wo, wi, surf.ns);
        float pdfDirToPrev = materialPdf(matType, mp, wi, wo, surf.ns);

        if (nextId == -1) break;
        if (lightId >= 0) {
            vec3 nLit = triangleSurfaceInfo<C>(s, nextId, nextPos, cn).ng;
            LightPdf pl = lightPdfLe<C>(s, lightId, makeRay(nextPos, -wi), cn);
            float pdfPLit = remapP(pl.pdfPos * lightPdfSampleOne<C>(s, lightId, cn));
            float coefToLight = remapP(pdfDirToNext * satDot(nLit, -wi));
            float coefToSurf = remapP(pl.pdfDir * absDot(surf.ns, wi));
            float coefToPrev = (bounce == 1) ? 1.0f : remapP(pdfDirToPrev * absDot(prevNorm, wo));
            float coefDist = remapP(dist * dist);
            float s1s0 = pdfPLit * coefDist / coefToLight;
            float t1s0x = t1s0 * coefToSurf * pdfPLit * coefToPrev / coefToLight;
            float weight = fisnan(t1s0) ? 0.0f : 1.0f / (1.0f + s1s0 + t1s0x);
            result += lightLe<C>(s, lightId, nextPos, -wi, cn) * throughput * weight;
            break;
        }
        if (prm.russianRoulette) {
            float continueProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 0.95f);
            if (sample1D<C>(s, rng, cn) >= continueProb) break;
            throughput /= continueProb;
        }
        float coef = ((bounce == 1) ? 1.0f : remapP(pdfDirToPrev * absDot(prevNorm, wo))) /
                     remapP(pdfDirToNext * absDot(triangleNormalShad<C>(s, nextId, nextPos, cn), wi));
        t1s0 *= coef;
        t1s1 *= coef;
        prevNorm = surf.ns;
        pos = nextPos;
        wo = -wi;
        id = nextId;
    }
    return result;
}

template <bool C>
__global__ __launch_bounds__(kBlock) void k_triple_pt(DevScene s, PassParams prm, float* __restrict__ film,
                                                      unsigned long long* counters)
{
    const int tilesX = (s.filmW + 15) / 16, tilesY = (s.filmH + 15) / 16;
    const int tile = xcdTile(blockIdx.x, gridDim.x);
    Counters cn;
    if (tile < tilesX * tilesY) {
        int x, y;
        tilePixel(tile, tilesX, x, y);
        if (x < s.filmW && y < s.filmH) {
            const vec2 scr = v2((float)x, (float)y) / v2((float)s.filmW, (float)s.filmH);
            const PixelSeed ps = pixelSeed(s, scr);
            float* f = film + (size_t)y * 3 * s.filmW + 3 * x;
            vec3 acc = v3(f[0], f[1], f[2]);
            for (int k = 0; k < prm.nPasses; k++) {
                const int spp = prm.firstPass + k;
                const uint32_t fc = (uint32_t)(prm.freeCounter + k);
                Rng rng;
                rng.randSeed = ps.tx * fc + ps.ty;
                rng.sampleSeed = ps.tx * ps.ty;
                rng.sampleOffset = sampleOffsetOf(spp);
                rng.dim = 0;
                vec4 u = sample4D<C>(s, rng, cn);
                Ray ray = cameraSampleRay(s, scr, u);
                vec3 res = tripleCameraTrace<C>(s, prm, ray, rng, cn);
                if (!hasNan(res)) acc = acc + res;
            }
            f[0] = acc.x; f[1] = acc.y; f[2] = acc.z;
        }
    }
    flushCounters<C>(cn, counters);
}

// ---------------------------------------------------------------------------
// Triple path tracer, light pass (triple_path_pass_lpt.glsl:58-182): t=1 with the
// running ratios s0t1 / s1t1.
// ---------------------------------------------------------------------------
template <bool C>
__device__ void tripleLightTrace(const DevScene& s, const PassParams& prm, Rng& rng, float* film, Counters& cn)
{
    int light = lightSampleOne<C>(s, sample2D<C>(s, rng, cn), cn);
    float pdfSource = lightPdfSampleOne<C>(s, light, cn);
    Ray ray;
    vec3 wo, throughput, prevNorm;
    float prevPdfDir, s0t1, s1t1;
    {
        int triId = light + s.objPrimCount;
        (void)triangleSampleUniformId<C>(s, triId, sample2D<C>(s, rng, cn), cn);
        LightLeSample le = lightSampleOneLe<C>(s, light, sample4D<C>(s, rng, cn), cn);
        vec3 nl = triangleSurfaceInfo<C>(s, triId, le.ray.ori, cn).ng;
        wo = -le.ray.dir;
        ray = rayOffseted(le.ray.ori, le.ray.dir);
        throughput = le.Le * absDot(nl, -wo) / (pdfSource * le.pdfPos * le.pdfDir);
        prevNorm = nl;
        prevPdfDir = le.pdfDir;
        s0t1 = 1.0f / remapP(le.pdfPos * pdfSource);
        s1t1 = 1.0f;
    }
    for (int bounce = 1; bounce <= prm.maxDepth; bounce++) {
        float dist;
        int id = bvhHit<C>(s, ray, dist, cn);
        if (id == -1) break;
        if (id - s.objPrimCount >= 0) break;
        vec3 pos = rayPoint(ray, dist);
        SurfaceInfo surf = triangleSurfaceInfo<C>(s, id, pos, cn);
        const int matTex = s.matTexIndices[id];
        const int matId = matTex & 0x0000ffff;
        cnt<C>(cn, 4);
        uint32_t matType = loadMaterialType<C>(s, matId, cn);
        if (matType != MatDielectric && matType != MatThinDielectric) {
            if (dot(surf.ns, wo) < 0.0f) flipNormals(surf);
        }
        float coefToPos = remapP(prevPdfDir * absDot(surf.ns, wo));
        s0t1 /= coefToPos;
        s1t1 /= coefToPos / (bounce == 1 ? remapP(dist * dist) : 1.0f);
        BSDFParam mp = loadMaterial<C>(s, matType, matId, matTex >> 16, surf.uv, cn);
        {
            CameraIiSample ci = cameraSampleIi(s, pos, sample2D<C>(s, rng, cn));
            if (ci.pdf > 0.0f) {
                vec3 pCam = pos + ci.wi * ci.dist;
                if (visible<C>(s, pos, pCam, cn)) {
                    float cosWi = satDot(surf.ng, ci.wi) * fabsG(dot(surf.ns, wo) / dot(surf.ng, wo));
                    vec3 f = materialBSDF(matType, mp, wo, ci.wi, surf.ns, ModeImportance);
                    vec3 contrib = ci.Ii * f * throughput * cosWi / ci.pdf;
                    float coefToSurf = remapP(cameraPdfIe(s, makeRay(pCam, -ci.wi)).pdfDir * satDot(surf.ns, ci.wi));
                    float coefToPrev = remapP(materialPdf(matType, mp, ci.wi, wo, surf.ns) * absDot(prevNorm, wo));
                    float coefDist = remapP(ci.dist * ci.dist);
                    float coef0 = coefToSurf * coefToPrev / coefDist;
                    float coef1 = ((bounce == 1) ? 1.0f : coefToPrev) * coefToSurf / coefDist;
                    float weight = 1.0f / (s0t1 * coef0 + s1t1 * coef1 + 1.0f);
                    vec3 res = contrib * weight;
                    if (!hasNan(res) && !fisnan(ci.pdf) && ci.pdf > 1e-8f && !isBlack(res)) {
                        splatFilm(s, film, ci.uv, res * prm.lptScale);
                        cnt<C>(cn, 12);
                    }
                }
            }
        }
        BSDFSample bs = materialSample(matType, mp, surf.ns, wo, ModeImportance, sample3D<C>(s, rng, cn), rng);
        vec3 wi = bs.wi;
        float bsdfPdf = bs.pdf;
        bool deltaBsdf = (bs.flag == FlagSpecRefl || bs.flag == FlagSpecTrans);
        if (bsdfPdf < 1e-8f || fisnan(bsdfPdf)) break;
        if (prm.russianRoulette) {
            float continueProb = fminG(maxComponent(bs.bsdf / bsdfPdf), 1.0f);
            if (sample1D<C>(s, rng, cn) >= continueProb) break;
            throughput /= continueProb;
        }
        float coefToPrev = remapP(materialPdf(matType, mp, wi, wo, surf.ns) * absDot(prevNorm, wo));
        s0t1 *= coefToPrev;
        s1t1 *= (bounce == 1) ? 1.0f : coefToPrev;
        prevPdfDir = materialPdf(matType, mp, wo, wi, surf.ns);
        prevNorm = surf.ns;
        float cosWi = deltaBsdf ? 1.0f : fabsG(dot(surf.ng, wi) * dot(surf.ns, wo) / dot(surf.ng, wo));
        throughput *= bs.bsdf * cosWi / bsdfPdf;
        ray = rayOffseted(pos, wi);
        wo = -wi;
    }
}

template <bool C>
__global__ __launch_bounds__(kBlock) void k_triple_lpt(DevScene s, PassParams prm, float* __restrict__ film,
                                                       unsigned long long* counters)
{
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    Counters cn;
    if (g < (int64_t)prm.pathsPerPass * prm.nPasses) {
        const int k = (int)(g / prm.pathsPerPass);
        const uint32_t id = (uint32_t)(g % prm.pathsPerPass);
        const int spp = prm.firstPass + k;
        Rng rng;
        rng.randSeed = (uint32_t)spp * (uint32_t)(prm.pathsPerPass * prm.loopsPerPass) + id +
                       (uint32_t)(prm.freeCounter + k);
        rng.sampleSeed = 0;
        rng.sampleOffset = sampleOffsetOf(spp);
        rng.dim = 0;
        for (int i = 0; i < prm.loopsPerPass; i++) tripleLightTrace<C>(s, prm, rng, film, cn);
    }
    flushCounters<C>(cn, counters);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned gridFor(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launchSobolTable(const uint32_t* M, uint32_t* tab, int first, int count, hipStream_t st)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sobol_table, dim3(gridFor((int64_t)count * 256, 256)), dim3(256), 0, st, M, tab, first, count);
    return hipGetLastError();
}
hipError_t launchSrgbExpand(const uint8_t* rgb, const float* lut, float4* out, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_srgb_expand, dim3(gridFor(n, 256)), dim3(256), 0, st, rgb, lut, out, n);
    return hipGetLastError();
}
int residentBlocksPerCU(const void* kernel, int threads, int queried)
{
    int computed = 0;
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, kernel) == hipSuccess && a.numRegs > 0) {
        const int alloc = (a.numRegs + 7) / 8 * 8;
        const int wavesPerSimd = std::min(8, 512 / alloc);
        const int wavesPerBlock = (threads + 63) / 64;
        computed = std::max(1, wavesPerSimd * 4 / wavesPerBlock);
        if (a.sharedSizeBytes > 0) computed = std::min<int>(computed, (int)(((size_t)160 << 10) / a.sharedSizeBytes));
    }
    return std::max({1, queried, computed});
}

hipError_t launchFill(float* p, int64_t n, float v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    unsigned g = gridFor(n, 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
hipError_t launchPostProcess(const float4* frame, float4* out, unsigned char* rgb8, int w, int h, float scale, int tone,
                             hipStream_t st)
{
    const int n = w * h;
    hipLaunchKernelGGL(k_post_process, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, frame, out, rgb8, w, h,
                       scale, tone);
    return hipGetLastError();
}

hipError_t launchFilmToFrame(const float* film, float4* frame, int w, int h, hipStream_t st)
{
    hipLaunchKernelGGL(k_film_to_frame, dim3(gridFor((int64_t)w * h, 256)), dim3(256), 0, st, film, frame, w, h);
    return hipGetLastError();
}
hipError_t launchTraceClosest(const DevScene& s, const float* rays, int64_t n, int32_t* prim, float* t,
                              uint32_t* bin, unsigned long long* counters, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    if (counters)
        hipLaunchKernelGGL(k_trace_closest<true>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, s, rays, n, prim, t,
                           bin, counters);
    else
        hipLaunchKernelGGL(k_trace_closest<false>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, s, rays, n, prim, t,
                           bin, counters);
    return hipGetLastError();
}
hipError_t launchTraceAny(const DevScene& s, const float* rays, const float* tmax, int64_t n, uint8_t* occ,
                          hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_any, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, s, rays, tmax, n, occ);
    return hipGetLastError();
}
static inline unsigned tileGrid(const DevScene& s)
{
    const int tiles = ((s.filmW + 15) / 16) * ((s.filmH + 15) / 16);
    return (unsigned)(((tiles + 7) / 8) * 8);  // a multiple of 8 so every XCD gets the same count
}
hipError_t launchLight(const DevScene& s, const PassParams& p, float* film, unsigned long long* counters,
                       hipStream_t st)
{
    const int64_t n = (int64_t)p.pathsPerPass * p.nPasses;
    if (n <= 0) return hipSuccess;
    if (counters)
        hipLaunchKernelGGL(k_light<true>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, s, p, film, counters);
    else
        hipLaunchKernelGGL(k_light<false>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, s, p, film, counters);
    return hipGetLastError();
}
hipError_t launchTriplePT(const DevScene& s, const PassParams& p, float* film, unsigned long long* counters,
                          hipStream_t st)
{
    if (counters)
        hipLaunchKernelGGL(k_triple_pt<true>, dim3(tileGrid(s)), dim3(kBlock), 0, st, s, p, film, counters);
    else
        hipLaunchKernelGGL(k_triple_pt<false>, dim3(tileGrid(s)), dim3(kBlock), 0, st, s, p, film, counters);
    return hipGetLastError();
}
hipError_t launchTripleLPT(const DevScene& s, const PassParams& p, float* film, unsigned long long* counters,
                           hipStream_t st)
{
    const int64_t n = (int64_t)p.pathsPerPass * p.nPasses;
    if (n <= 0) return hipSuccess;
    if (counters)
        hipLaunchKernelGGL(k_triple_lpt<true>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, s, p, film, counters);
    else
        hipLaunchKernelGGL(k_triple_lpt<false>, dim3(gridFor(n, kBlock)), dim3(kBlock), 0, st, s, p, film, counters);
    return hipGetLastError();
}

}  // namespace zgl
