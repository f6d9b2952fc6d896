// zgl_device.h -- device-side rendering library for gfx950.
//
// Restates the reference's GLSL compute-shader library (src/shader/intersection.glsl,
// random.glsl, microfacet.glsl, material.glsl, material_loader.glsl, light.glsl,
// camera.glsl) over an HBM layout chosen for wave64 gathers:
//   * MTBVH rows: one 32-byte record per (table, row) = {pMin.xyz, pMax.x} {pMax.yz,
//     prim, miss}.  The reference keeps the node index in a 12-byte row and the
//     bounds in a separate 24-byte array; merging them removes one dependent gather
//     per traversal step and the "hit" link k+1 becomes the next 32 bytes.
//   * Triangles for the intersection test: one 48-byte record per triangle
//     {a.xyz, b.x} {b.yz, c.xy} {c.z, -, -, -}, so a leaf test is three 16-byte loads
//     instead of an index gather followed by three vertex gathers.
//   * Shading attributes (indices, positions, normals, uvs), materials, light and
//     environment tables keep the reference layout.
// Floating-point evaluation order follows the GLSL source expression by expression.
#pragma once

#include "zgl_box.h"
#include "zgl_math.h"

namespace zgl {

#define ZGL_D __device__ __forceinline__

// BSDF flags / types / transport modes (material.glsl:4-35)
constexpr uint32_t FlagDiffuse = 1u << 0;
constexpr uint32_t FlagGlosRefl = 1u << 1;
constexpr uint32_t FlagGlosTrans = 1u << 2;
constexpr uint32_t FlagSpecRefl = 1u << 3;
constexpr uint32_t FlagSpecTrans = 1u << 4;
constexpr uint32_t FlagInvalid = 1u << 16;
constexpr uint32_t MatLambertian = 0, MatPrincipled = 1, MatMetalWorkflow = 2, MatDielectric = 3,
                   MatThinDielectric = 4;
constexpr uint32_t ModeRadiance = 0, ModeImportance = 1;

constexpr int SobolSampleNum = 131072;
constexpr int SobolSampleDim = 256;

struct DevScene {
    const float4* rows;          // 6 * bvhSize * 2
    const float4* tris;          // nTriangles * 3
    const int32_t* indices;      // 3 per triangle
    const float* vertices;       // 3 per vertex
    const float* normals;        // 3 per vertex
    const float* texcoords;      // 2 per object vertex
    const int32_t* matTexIndices;
    const float4* materials;     // 4 texels per material
    const float* lightPower;     // 3 per light triangle
    const int32_t* lightAlias;
    const float* lightProb;
    const ushort4* env;          // RGB16F texels, padded to 8 bytes
    const int32_t* envAlias;     // (envW + 1) * envH
    const float* envProb;
    const float2* noise;         // filmW * filmH, RG32F
    const float4* tex;           // texture array: nTextures x texH x texW texels, linear RGB
                                 // (the sRGB8 layers decoded), row 0 = t 0
    const float2* texScale;      // per layer: image size / array size (texUVScale)
    const uint32_t* sobolM;      // 256 * 32 generator columns
    const uint32_t* sobolTab;    // precomputed Sobol points for a pass range
    int64_t nVertices;
    int32_t nTexcoords;
    int32_t nTriangles;
    int32_t objPrimCount;
    int32_t bvhSize;
    int32_t nMaterials;
    int32_t nLightTriangles;
    float lightSum;
    int32_t envW, envH;
    float envSum;
    float envRotCos, envRotSin;    // cos/sin of +rotation
    float envRotCosN, envRotSinN;  // cos/sin of -rotation
    int32_t sampler;               // 0: independent rand(), 1: Sobol
    int32_t sobolTabFirst, sobolTabCount;
    int32_t noiseW, noiseH;
    int32_t texW, texH, nTextures;
    // camera (camera.glsl:3-12)
    vec3 camF, camR, camU, camPos;
    mat3 camMatInv;
    float tanFOV, camAsp, lensRadius, focalDist;
    int32_t filmW, filmH;
};

// The scene of a persistent kernel whose first argument is its DevScene `arg`, read
// through the kernel-argument segment (where that first argument lies at offset 0) with
// an index the compiler cannot fold: each field is then a scalar load where it is used,
// instead of ~90 SGPRs of scene fields held across the kernel's whole loop, which
// spilled SGPRs into VGPR lanes and pushed 14 VGPRs of the naive tracer to scratch.
ZGL_D const DevScene& sceneArg(const DevScene& arg)
{
#if defined(__HIP_DEVICE_COMPILE__)
    int i = 0;
    asm volatile("" : "+s"(i));
    return ((const DevScene*)__builtin_amdgcn_kernarg_segment_ptr())[i];
#else
    return arg;
#endif
}

// Algorithmic-byte and visit counters, kept in registers and flushed once per thread.
struct Counters {
    uint32_t bytes = 0, rows = 0, tests = 0, traversals = 0;
};
template <bool C> ZGL_D void cnt(Counters& c, uint32_t b) { if (C) c.bytes += b; }

// ---------------------------------------------------------------------------
// Rays (intersection.glsl:12-42)
// ---------------------------------------------------------------------------
struct Ray { vec3 ori, dir; };
ZGL_D Ray makeRay(vec3 o, vec3 d) { return Ray{o, d}; }
ZGL_D vec3 rayPoint(const Ray& r, float t) { return r.ori + r.dir * t; }
ZGL_D Ray rayOffseted(vec3 ori, vec3 dir) { return Ray{ori + dir * 1e-4f, dir}; }

ZGL_D RayPre prepareRay(const Ray& r) { return prepareRay(r.ori, r.dir); }

// intersection.glsl:63-109.  Returns hit and writes t; u, v, det for the bin.
ZGL_D bool intersectTriangle(vec3 a, vec3 b, vec3 c, vec3 o, vec3 d, float& t, float* uo = nullptr,
                             float* vo = nullptr, float* deto = nullptr)
{
    const float eps = 1e-6f;
    vec3 ab = b - a;
    vec3 ac = c - a;
    vec3 p = cross(d, ac);
    float det = dot(ab, p);
    if (fabsG(det) < eps) return false;
    vec3 ao = o - a;
    if (det < 0.0f) { ao = -ao; det = -det; }
    float u = dot(ao, p);
    if (u < 0.0f || u > det) return false;
    vec3 q = cross(ao, ab);
    float v = dot(d, q);
    if (v < 0.0f || u + v > det) return false;
    t = dot(ac, q) / det;
    if (uo) { *uo = u; *vo = v; *deto = det; }
    return t > 0.0f;
}

ZGL_D void loadTri(const DevScene& s, int prim, vec3& a, vec3& b, vec3& c)
{
    const float4* t = s.tris + (size_t)prim * 3;
    float4 t0 = t[0], t1 = t[1], t2 = t[2];
    a = v3(t0.x, t0.y, t0.z);
    b = v3(t0.w, t1.x, t1.y);
    c = v3(t1.z, t1.w, t2.x);
}

// bvhHit / bvhTest (intersection.glsl:367-427): stackless walk over the table picked
// by cubemapFace(-dir); a missed or farther box jumps to the row's miss link.  Row
// indices and miss links are absolute (table * bvhSize + row), so a row's address is
// one shift and add.  A wave whose rays are all "general" (RayPre::special false) uses
// the three-slab test alone, which gives the same result for those rays.
ZGL_D const float4* rowPtr(const DevScene& s, int k) { return s.rows + (size_t)(uint32_t)k * 2; }

template <bool ANY, bool C>
ZGL_D int traverse(const DevScene& s, const RayPre& p, float& dist, Counters& cn)
{
    int closest = -1;
    const bool anySpecial = __ballot(p.special) != 0ull;
    if (C) cn.traversals++;
    int k = p.table * s.bvhSize;
    const int end = k + s.bvhSize;
    while (k != end) {
        // up to 4 rows, stopping at a leaf whose box is hit; then that leaf's triangle.
        // Lanes of a wave reach leaves at different rows: testing once per round keeps
        // the triangle test from running on nearly every row for a few lanes.
        int leaf = -1;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (k != end && leaf < 0) {
                const float4* row = rowPtr(s, k);
                const float4 lo = row[0];
                const float4 hi = row[1];
                if (C) { cn.rows++; cn.bytes += 36; }
                float boxDist;
                const vec3 bMin = v3(lo.x, lo.y, lo.z), bMax = v3(lo.w, hi.x, hi.y);
                const bool bHit = anySpecial ? boxHit(bMin, bMax, p, boxDist) : boxHitGeneral(bMin, bMax, p, boxDist);
                if (!bHit || boxDist > dist) {
                    k = __float_as_int(hi.w);
                } else {
                    const int prim = __float_as_int(hi.z);
                    if (prim >= 0) leaf = prim;
                    else k++;
                }
            }
        }
        if (leaf >= 0) {
            vec3 a, b, c;
            loadTri(s, leaf, a, b, c);
            if (C) { cn.tests++; cn.bytes += 48; }
            float t;
            if (intersectTriangle(a, b, c, p.o, p.d, t) && t < dist) {
                if (ANY) return leaf;
                dist = t;
                closest = leaf;
            }
            k++;
        }
    }
    return closest;
}

// Resumable traversal: the state of one bvhHit / bvhTest walk.  It advances by rows
// (walkRow); a row that is a leaf whose box is hit stops the walk there until walkLeaf
// tests its triangle.  The rows are visited in the same order with the same tests as
// traverse(), so the result is the same bit for bit however walks of different lanes
// are interleaved.
struct Walk {
    RayPre p;
    float dist;
    int k;     // absolute row index
    int end;   // one past the table's last row
    int closest;
    int leaf;  // primitive of the leaf the walk is stopped at, -1 while it moves
    bool any;
};
// The row table and the triangle positions as seen by a walk loop: loaded from the
// scene once before the loop and held in scalar registers.  Read through the scene at
// every step, the compiler reloads the pointer from the kernel arguments inside each
// row step, which puts a scalar load and its wait on every row's dependent chain; the
// empty asm keeps it from doing so.  The pointers keep the global address space, so the
// loads stay global loads.
typedef float GFloat4 __attribute__((ext_vector_type(4)));
typedef const GFloat4 __attribute__((address_space(1)))* GRows;
typedef const float __attribute__((address_space(1)))* GFloats;
struct WalkMem {
    GRows rows;
    GRows tris;
};
ZGL_D WalkMem walkMem(const DevScene& s)
{
    WalkMem m{(GRows)s.rows, (GRows)s.tris};
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(m.rows), "+s"(m.tris));
#endif
    return m;
}

template <bool C>
ZGL_D void walkStart(const DevScene& s, Walk& w, const Ray& r, bool any, float tmax, Counters& cn)
{
    if (C) cn.traversals++;
    w.p = prepareRay(r);
    w.dist = tmax;
    w.k = w.p.table * s.bvhSize;
    w.end = w.k + s.bvhSize;
    w.closest = -1;
    w.leaf = -1;
    w.any = any;
}
// One row: returns true when the walk has finished.  A leaf whose box is hit stops the
// walk (w.leaf >= 0) without testing the triangle.  `general` is true when the caller
// knows that no lane walking with it has a special ray.
template <bool C> ZGL_D bool walkRow(const WalkMem& m, Walk& w, bool general, Counters& cn)
{
    const GRows row = m.rows + (size_t)(uint32_t)w.k * 2;
    const GFloat4 lo = row[0];
    const GFloat4 hi = row[1];
    if (C) { cn.rows++; cn.bytes += 36; }
    const vec3 bMin = v3(lo.x, lo.y, lo.z), bMax = v3(lo.w, hi.x, hi.y);
    // the three-slab test straight after its loads; the other cases only when a lane of
    // the wave needs them
    float boxDist;
    bool bHit = boxHitGeneral(bMin, bMax, w.p, boxDist);
    if (!general) bHit = boxHit(bMin, bMax, w.p, boxDist);
    // next row: the miss link, the row after, or stay at a hit leaf (selects, no branch)
    const bool take = bHit && !(boxDist > w.dist);
    const int prim = __float_as_int(hi.z);
    const bool leaf = take && prim >= 0;
    w.leaf = leaf ? prim : w.leaf;
    w.k = take ? (leaf ? w.k : w.k + 1) : __float_as_int(hi.w);
    return !leaf && w.k == w.end;
}
// The triangle test of the leaf the walk is stopped at; returns true when the walk has
// finished.
template <bool C> ZGL_D bool walkLeaf(const WalkMem& m, Walk& w, Counters& cn)
{
    const int prim = w.leaf;
    w.leaf = -1;
    const GRows tri = m.tris + (size_t)prim * 3;
    const GFloat4 t0 = tri[0], t1 = tri[1];
    const float t2x = ((GFloats)(tri + 2))[0];
    const vec3 a = v3(t0.x, t0.y, t0.z), b = v3(t0.w, t1.x, t1.y), c = v3(t1.z, t1.w, t2x);
    if (C) { cn.tests++; cn.bytes += 48; }
    float t;
    if (intersectTriangle(a, b, c, w.p.o, w.p.d, t) && t < w.dist) {
        w.closest = prim;
        if (w.any) return true;
        w.dist = t;
    }
    w.k++;
    return w.k == w.end;
}

template <bool C> ZGL_D int bvhHit(const DevScene& s, const Ray& r, float& dist, Counters& cn)
{
    dist = 1e8f;
    RayPre p = prepareRay(r);
    return traverse<false, C>(s, p, dist, cn);
}
template <bool C> ZGL_D bool bvhTest(const DevScene& s, const Ray& r, float dist, Counters& cn)
{
    RayPre p = prepareRay(r);
    return traverse<true, C>(s, p, dist, cn) >= 0;
}
// intersection.glsl:429-434
template <bool C> ZGL_D bool visible(const DevScene& s, vec3 x, vec3 y, Counters& cn)
{
    float dist = distance(x, y) - 2e-5f;
    vec3 wi = normalize(y - x);
    return !bvhTest<C>(s, makeRay(x + wi * 1e-5f, wi), dist, cn);
}

// ---------------------------------------------------------------------------
// Triangle attributes (intersection.glsl:111-224), reference layout
// ---------------------------------------------------------------------------
ZGL_D vec3 ldV3(const float* p, int i) { return v3(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]); }

struct SurfaceInfo { vec3 ns, ng; vec2 uv; };
ZGL_D void flipNormals(SurfaceInfo& s) { s.ns = -s.ns; s.ng = -s.ng; }

template <bool C>
ZGL_D SurfaceInfo triangleSurfaceInfo(const DevScene& s, int id, vec3 p, Counters& cn)
{
    int ia = s.indices[3 * id + 0], ib = s.indices[3 * id + 1], ic = s.indices[3 * id + 2];
    vec3 a = ldV3(s.vertices, ia), b = ldV3(s.vertices, ib), c = ldV3(s.vertices, ic);
    vec3 na = ldV3(s.normals, ia), nb = ldV3(s.normals, ib), nc = ldV3(s.normals, ic);
    // texelFetch past the end of the texcoord buffer (light meshes have none) reads 0
    vec2 ta = (ia < s.nTexcoords) ? v2(s.texcoords[2 * ia], s.texcoords[2 * ia + 1]) : v2(0.0f, 0.0f);
    vec2 tb = (ib < s.nTexcoords) ? v2(s.texcoords[2 * ib], s.texcoords[2 * ib + 1]) : v2(0.0f, 0.0f);
    vec2 tc = (ic < s.nTexcoords) ? v2(s.texcoords[2 * ic], s.texcoords[2 * ic + 1]) : v2(0.0f, 0.0f);
    cnt<C>(cn, 108);
    vec3 pa = a - p, pb = b - p, pc = c - p;
    float areaInv = 1.0f / length(cross(b - a, c - a));
    float la = length(cross(pb, pc)) * areaInv;
    float lb = length(cross(pc, pa)) * areaInv;
    float lc = 1.0f - la - lb;
    SurfaceInfo r;
    r.ns = normalize((na * la + nb * lb) + nc * lc);
    r.ng = normalize(cross(pa, pb));
    r.uv = v2((ta.x * la + tb.x * lb) + tc.x * lc, (ta.y * la + tb.y * lb) + tc.y * lc);
    if (dot(r.ns, r.ng) < 0.0f) r.ng = -r.ng;
    return r;
}

template <bool C>
ZGL_D void triangleVerts(const DevScene& s, int id, vec3& a, vec3& b, vec3& c, Counters& cn)
{
    int ia = s.indices[3 * id + 0], ib = s.indices[3 * id + 1], ic = s.indices[3 * id + 2];
    a = ldV3(s.vertices, ia); b = ldV3(s.vertices, ib); c = ldV3(s.vertices, ic);
    cnt<C>(cn, 48);
}
// The triple tracer's pdf remap: p^2, or 1 for p < 1e-8 (triple_path_pass_pt.glsl).
ZGL_D float remapP(float p) { return p < 1e-8f ? 1.0f : p * p; }

template <bool C> ZGL_D float triangleAreaId(const DevScene& s, int id, Counters& cn)
{
    vec3 a, b, c;
    triangleVerts<C>(s, id, a, b, c, cn);
    return triangleArea(a, b, c);
}
template <bool C> ZGL_D vec3 triangleSampleUniformId(const DevScene& s, int id, vec2 u, Counters& cn)
{
    vec3 a, b, c;
    triangleVerts<C>(s, id, a, b, c, cn);
    return sampleTriangleUniform(a, b, c, u);
}
// intersection.glsl:149-173
template <bool C> ZGL_D vec3 triangleNormalShad(const DevScene& s, int id, vec3 p, Counters& cn)
{
    int ia = s.indices[3 * id + 0], ib = s.indices[3 * id + 1], ic = s.indices[3 * id + 2];
    vec3 a = ldV3(s.vertices, ia), b = ldV3(s.vertices, ib), c = ldV3(s.vertices, ic);
    vec3 na = ldV3(s.normals, ia), nb = ldV3(s.normals, ib), nc = ldV3(s.normals, ic);
    cnt<C>(cn, 84);
    vec3 pa = a - p, pb = b - p, pc = c - p;
    float areaInv = 1.0f / length(cross(b - a, c - a));
    float la = length(cross(pb, pc)) * areaInv;
    float lb = length(cross(pc, pa)) * areaInv;
    float lc = 1.0f - la - lb;
    return normalize((na * la + nb * lb) + nc * lc);
}

// ---------------------------------------------------------------------------
// Sampler (random.glsl:3-90).  The Sobol table of the reference
// (uSobolSeq[i*256 + dim], Sampler.cpp:48-64) is recomputed from the generator
// matrices; a launch precomputes the points of the passes it renders.
// ---------------------------------------------------------------------------
struct Rng {
    uint32_t randSeed;
    uint32_t sampleSeed;
    int sampleOffset;
    int dim;
};
ZGL_D float rand01(Rng& r) { r.randSeed = hashU(r.randSeed); return u32ToUnit(r.randSeed); }

ZGL_D uint32_t sobolValue(const DevScene& s, int idx)
{
    if ((uint32_t)idx >= (uint32_t)(SobolSampleNum * SobolSampleDim)) return 0u;  // texelFetch OOB
    int i = idx >> 8, d = idx & 255;
    int rel = i - s.sobolTabFirst;
    if ((uint32_t)rel < (uint32_t)s.sobolTabCount) return s.sobolTab[rel * 256 + d];
    const uint32_t* m = s.sobolM + d * 32;
    uint32_t r = 0;
    for (uint32_t u = (uint32_t)i, b = 0; u != 0u && b < 32; u >>= 1, b++)
        if (u & 1u) r ^= m[b];
    return r;
}
template <bool C> ZGL_D float sample1D(const DevScene& s, Rng& r, Counters& cn)
{
    if (s.sampler == 0) return rand01(r);
    uint32_t v = sobolValue(s, r.sampleOffset + r.dim);
    cnt<C>(cn, 4);
    v ^= r.sampleSeed;
    r.sampleSeed = hashU(r.sampleSeed);
    r.dim++;
    return u32ToUnit(v);
}
template <bool C> ZGL_D vec2 sample2D(const DevScene& s, Rng& r, Counters& cn)
{
    float a = sample1D<C>(s, r, cn);
    float b = sample1D<C>(s, r, cn);
    return v2(a, b);
}
template <bool C> ZGL_D vec3 sample3D(const DevScene& s, Rng& r, Counters& cn)
{
    float a = sample1D<C>(s, r, cn);
    float b = sample1D<C>(s, r, cn);
    float c = sample1D<C>(s, r, cn);
    return v3(a, b, c);
}
template <bool C> ZGL_D vec4 sample4D(const DevScene& s, Rng& r, Counters& cn)
{
    float a = sample1D<C>(s, r, cn);
    float b = sample1D<C>(s, r, cn);
    float c = sample1D<C>(s, r, cn);
    float d = sample1D<C>(s, r, cn);
    return v4(a, b, c, d);
}

// ---------------------------------------------------------------------------
// Textures: GL_LINEAR + GL_REPEAT bilinear lookup at level 0 in FP32.
// ---------------------------------------------------------------------------
ZGL_D int wrapRepeat(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }

ZGL_D vec2 noiseLookup(const DevScene& s, vec2 uv)
{
    const int W = s.noiseW, H = s.noiseH;
    float u = uv.x * (float)W - 0.5f, v = uv.y * (float)H - 0.5f;
    float fu = ffloorG(u), fv = ffloorG(v);
    float al = u - fu, be = v - fv;
    int i0 = wrapRepeat((int)fu, W), j0 = wrapRepeat((int)fv, H);
    int i1 = wrapRepeat((int)fu + 1, W), j1 = wrapRepeat((int)fv + 1, H);
    float2 t00 = s.noise[(size_t)j0 * W + i0], t10 = s.noise[(size_t)j0 * W + i1];
    float2 t01 = s.noise[(size_t)j1 * W + i0], t11 = s.noise[(size_t)j1 * W + i1];
    float w00 = (1.0f - al) * (1.0f - be), w10 = al * (1.0f - be), w01 = (1.0f - al) * be, w11 = al * be;
    return v2(((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
              ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y);
}

ZGL_D vec3 envTexel(const DevScene& s, int i, int j)
{
    ushort4 t = s.env[(size_t)j * s.envW + i];
    return v3((float)__builtin_bit_cast(_Float16, (unsigned short)t.x),
              (float)__builtin_bit_cast(_Float16, (unsigned short)t.y),
              (float)__builtin_bit_cast(_Float16, (unsigned short)t.z));
}
template <bool C> ZGL_D vec3 envLookup(const DevScene& s, vec2 uv, Counters& cn)
{
    const int W = s.envW, H = s.envH;
    float u = uv.x * (float)W - 0.5f, v = uv.y * (float)H - 0.5f;
    float fu = ffloorG(u), fv = ffloorG(v);
    float al = u - fu, be = v - fv;
    int i0 = wrapRepeat((int)fu, W), j0 = wrapRepeat((int)fv, H);
    int i1 = wrapRepeat((int)fu + 1, W), j1 = wrapRepeat((int)fv + 1, H);
    vec3 t00 = envTexel(s, i0, j0), t10 = envTexel(s, i1, j0), t01 = envTexel(s, i0, j1), t11 = envTexel(s, i1, j1);
    cnt<C>(cn, 24);
    float w00 = (1.0f - al) * (1.0f - be), w10 = al * (1.0f - be), w01 = (1.0f - al) * be, w11 = al * be;
    return ((t00 * w00 + t10 * w10) + t01 * w01) + t11 * w11;
}

// texture2DArray(uTextures, vec3(fract(uv) * texUVScale[texId], texId)) at level 0
// (material_loader.glsl:8-16): GL_LINEAR magnification, GL_REPEAT over the array's size,
// texels already decoded from sRGB (the GL_SRGB format decodes before filtering).
template <bool C> ZGL_D vec3 textureLookup(const DevScene& s, int texId, vec2 uv, Counters& cn)
{
    const float2 sc = s.texScale[texId];
    const int W = s.texW, H = s.texH;
    const float4* layer = s.tex + (size_t)texId * W * H;
    const float su = (uv.x - ffloorG(uv.x)) * sc.x, sv = (uv.y - ffloorG(uv.y)) * sc.y;
    float u = su * (float)W - 0.5f, v = sv * (float)H - 0.5f;
    float fu = ffloorG(u), fv = ffloorG(v);
    float al = u - fu, be = v - fv;
    int i0 = wrapRepeat((int)fu, W), j0 = wrapRepeat((int)fv, H);
    int i1 = wrapRepeat((int)fu + 1, W), j1 = wrapRepeat((int)fv + 1, H);
    const float4 a = layer[(size_t)j0 * W + i0], b = layer[(size_t)j0 * W + i1];
    const float4 c = layer[(size_t)j1 * W + i0], d = layer[(size_t)j1 * W + i1];
    cnt<C>(cn, 8 + 16);
    float w00 = (1.0f - al) * (1.0f - be), w10 = al * (1.0f - be), w01 = (1.0f - al) * be, w11 = al * be;
    return ((v3(a.x, a.y, a.z) * w00 + v3(b.x, b.y, b.z) * w10) + v3(c.x, c.y, c.z) * w01) + v3(d.x, d.y, d.z) * w11;
}

// ---------------------------------------------------------------------------
// Microfacet (microfacet.glsl:4-119)
// ---------------------------------------------------------------------------
ZGL_D float schlickW(float cosTheta) { return pow5(1.0f - cosTheta); }
ZGL_D vec3 schlickF(float cosTheta, vec3 F0) { return F0 + (v3(1.0f) - F0) * pow5(1.0f - cosTheta); }
ZGL_D vec3 schlickFr(float cosTheta, vec3 F0, float roughness)
{
    return F0 + (vmax(v3(1.0f - roughness), F0) - F0) * pow5(1.0f - cosTheta);
}
ZGL_D float schlickG(float cosTheta, float alpha)
{
    float k = alpha * 0.5f;
    return cosTheta / (cosTheta * (1.0f - k) + k);
}
ZGL_D float smithG(vec3 n, vec3 wo, vec3 wi, float alpha)
{
    return schlickG(absDot(n, wo), alpha) * schlickG(absDot(n, wi), alpha);
}
ZGL_D float ggx(float cosTheta, float alpha)
{
    if (cosTheta < 1e-6f) return 0.0f;
    float a2 = alpha * alpha;
    float nom = a2;
    float denom = cosTheta * cosTheta * (a2 - 1.0f) + 1.0f;
    denom = denom * denom * Pi;
    return nom / denom;
}
ZGL_D float ggxD(vec3 n, vec3 m, float alpha) { return ggx(dot(n, m), alpha); }
ZGL_D float ggxPdfWm(vec3 n, vec3 m, vec3 wo, float alpha) { return ggx(dot(n, m), alpha); }
ZGL_D float ggxPdfVisibleWm(vec3 n, vec3 m, vec3 wo, float alpha)
{
    return ggx(dot(n, m), alpha) * schlickG(dot(n, wo), alpha) * absDot(m, wo) / absDot(n, wo);
}
ZGL_D vec3 ggxSampleWm(vec3 n, vec3 wo, float alpha, vec2 u)
{
    vec2 xi = toConcentricDisk(u);
    vec3 h = v3(xi.x, xi.y, fsqrtG(fmaxG(0.0f, 1.0f - xi.x * xi.x - xi.y * xi.y)));
    h = normalize(h * v3(alpha, alpha, 1.0f));
    return normalToWorld(n, h);
}
ZGL_D vec3 ggxSampleVisibleWm(vec3 n, vec3 wo, float alpha, vec2 u)
{
    mat3 tbn = tbnMatrix(n);
    mat3 tbnInv = inverse(tbn);
    vec3 vh = normalize(mul(tbnInv, wo) * v3(alpha, alpha, 1.0f));
    float lensq = vh.x * vh.x + vh.y * vh.y;
    vec3 t1 = lensq > 0.0f ? v3(-vh.y, vh.x, 0.0f) / fsqrtG(lensq) : v3(1.0f, 0.0f, 0.0f);
    vec3 t2 = cross(vh, t1);
    vec2 xi = toConcentricDisk(u);
    float sv = 0.5f * (1.0f + vh.z);
    xi.y = (1.0f - sv) * fsqrtG(1.0f - xi.x * xi.x) + sv * xi.y;
    vec3 h = (t1 * xi.x + t2 * xi.y) + vh * fsqrtG(fmaxG(0.0f, 1.0f - xi.x * xi.x - xi.y * xi.y));
    h = normalize(v3(h.x * alpha, h.y * alpha, fmaxG(0.0f, h.z)));
    return normalToWorld(n, h);
}
ZGL_D float gtr1(float cosTheta, float alpha)
{
    float a2 = alpha * alpha;
    return (a2 - 1.0f) / (2.0f * Pi * logZ(alpha) * (1.0f + (a2 - 1.0f) * cosTheta * cosTheta));
}
ZGL_D float gtr1D(vec3 n, vec3 m, float alpha) { return gtr1(satDot(n, m), alpha); }
ZGL_D vec3 gtr1SampleWm(vec3 n, vec3 wo, float alpha, vec2 u)
{
    float cosTheta = fsqrtG(fmaxG(0.0f, (1.0f - powZ(alpha, 1.0f - u.x)) / (1.0f - alpha)));
    float sinTheta = fsqrtG(fmaxG(0.0f, 1.0f - cosTheta * cosTheta));
    float phi = 2.0f * u.y * Pi;
    float sp, cp;
    sincosZ(phi, &sp, &cp);
    vec3 m = normalize(v3(cp * sinTheta, sp * sinTheta, cosTheta));
    if (!sameHemisphere(n, wo, m)) m = -m;
    return normalize(normalToWorld(n, m));
}
ZGL_D float gtr1PdfWm(vec3 n, vec3 m, vec3 wo, float alpha) { return gtr1D(n, m, alpha) * absDot(n, m); }

// ---------------------------------------------------------------------------
// BSDFs (material.glsl:36-555)
// ---------------------------------------------------------------------------
struct BSDFParam {
    vec3 baseColor;
    float subsurface, metallic, roughness, specular, specularTint;
    float sheen, sheenTint, clearcoat, clearcoatGloss;
    float ior;
};
struct BSDFSample { vec3 wi; float pdf; vec3 bsdf; float eta; uint32_t flag; };
ZGL_D BSDFSample makeSample(vec3 wi, float pdf, vec3 bsdf, float eta, uint32_t flag)
{
    return BSDFSample{wi, pdf, bsdf, eta, flag};
}
ZGL_D BSDFSample invalidSample() { return makeSample(v3(0.0f), 0.0f, v3(0.0f), 0.0f, FlagInvalid); }
ZGL_D bool approximateDelta(float roughness) { return roughness < 0.02f; }

ZGL_D vec3 lambertian(const BSDFParam& p) { return p.baseColor * PiInv; }
ZGL_D float lambertianPdf(vec3 wi, vec3 n) { return satDot(wi, n) * PiInv; }
ZGL_D BSDFSample lambertianSample(vec3 n, vec3 wo, const BSDFParam& p, vec3 u)
{
    vec4 c = sampleCosineWeighted(n, v2(u.y, u.z));
    vec3 wi = v3(c.x, c.y, c.z);
    float pdf = satDot(n, wi) * PiInv;
    return makeSample(wi, pdf, p.baseColor * PiInv, 1.0f, FlagDiffuse);
}

ZGL_D vec3 metalWorkflow(vec3 wo, vec3 wi, vec3 n, const BSDFParam& p)
{
    vec3 baseColor = p.baseColor;
    float metallic = p.metallic, roughness = p.roughness;
    float alpha = square(roughness);
    vec3 h = normalize(wi + wo);
    if (!sameHemisphere(n, wo, wi)) return v3(0.0f);
    float cosWi = dot(n, wi), cosWo = dot(n, wo);
    vec3 f0 = mix(v3(0.04f), baseColor, metallic);
    vec3 f = schlickFr(satDot(h, wo), f0, roughness);
    float d = ggxD(n, h, alpha);
    float g = smithG(n, wo, wi, alpha);
    vec3 kd = v3(1.0f) - f;
    kd *= 1.0f - metallic;
    float denom = 4.0f * cosWo * cosWi;
    if (denom < 1e-7f) return v3(0.0f);
    return kd * baseColor * PiInv + f * d * g / denom;
}
ZGL_D float metalWorkflowPdf(vec3 wo, vec3 wi, vec3 n, const BSDFParam& p)
{
    float alpha = square(p.roughness);
    vec3 h = normalize(wo + wi);
    float pdfDiff = satDot(n, wi) * PiInv;
    float pdfSpec = ggxPdfVisibleWm(n, h, wo, alpha) / (4.0f * absDot(h, wo));
    float spec = 1.0f / (2.0f - p.metallic);
    return mixf(pdfDiff, pdfSpec, spec);
}
ZGL_D BSDFSample metalWorkflowSample(vec3 n, vec3 wo, const BSDFParam& p, vec3 u)
{
    float alpha = square(p.roughness);
    float spec = 1.0f / (2.0f - p.metallic);
    uint32_t type = u.x > spec ? FlagDiffuse : FlagGlosRefl;
    vec3 wi;
    if (type == FlagDiffuse) {
        vec4 c = sampleCosineWeighted(n, v2(u.y, u.z));
        wi = v3(c.x, c.y, c.z);
    } else {
        vec3 h = ggxSampleVisibleWm(n, wo, alpha, v2(u.y, u.z));
        wi = reflect(-wo, h);
    }
    if (dot(n, wi) < 0.0f) return invalidSample();
    return makeSample(wi, metalWorkflowPdf(wo, wi, n, p), metalWorkflow(wo, wi, n, p), 1.0f, type);
}

ZGL_D bool refractDir(vec3& wt, vec3 wi, vec3 n, float eta)
{
    float cosTi = dot(n, wi);
    if (cosTi < 0.0f) eta = 1.0f / eta;
    float sin2Ti = fmaxG(0.0f, 1.0f - cosTi * cosTi);
    float sin2Tt = sin2Ti / (eta * eta);
    if (sin2Tt >= 1.0f) return false;
    float cosTt = fsqrtG(1.0f - sin2Tt);
    if (cosTi < 0.0f) cosTt = -cosTt;
    wt = normalize(-wi / eta + n * (cosTi / eta - cosTt));
    return true;
}
ZGL_D float fresnelDielectric(float cosTi, float eta)
{
    cosTi = fclampG(cosTi, -1.0f, 1.0f);
    if (cosTi < 0.0f) { eta = 1.0f / eta; cosTi = -cosTi; }
    float sinTi = fsqrtG(1.0f - cosTi * cosTi);
    float sinTt = sinTi / eta;
    if (sinTt >= 1.0f) return 1.0f;
    float cosTt = fsqrtG(1.0f - sinTt * sinTt);
    float rPa = (cosTi - eta * cosTt) / (cosTi + eta * cosTt);
    float rPe = (eta * cosTi - cosTt) / (eta * cosTi + cosTt);
    return (rPa * rPa + rPe * rPe) * 0.5f;
}
ZGL_D vec3 dielectric(vec3 wo, vec3 wi, vec3 n, const BSDFParam& p, uint32_t mode)
{
    vec3 baseColor = p.baseColor;
    float roughness = p.roughness, ior = p.ior;
    if (approximateDelta(roughness)) return v3(0.0f);
    vec3 h = normalize(wo + wi);
    float hCosWo = absDot(h, wo), hCosWi = absDot(h, wi);
    float alpha = roughness * roughness;
    if (sameHemisphere(n, wo, wi)) {
        float refl = fresnelDielectric(absDot(h, wi), ior);
        return (hCosWo * hCosWi < 1e-7f) ? v3(0.0f)
               : baseColor * ggxD(n, h, alpha) * smithG(n, wo, wi, alpha) / (4.0f * hCosWo * hCosWi) * refl;
    }
    float eta = dot(n, wi) > 0.0f ? ior : 1.0f / ior;
    float sqrtDenom = dot(h, wo) + eta * dot(h, wi);
    float denom = sqrtDenom * sqrtDenom;
    denom *= absDot(n, wi) * absDot(n, wo);
    float refl = fresnelDielectric(dot(h, wi), eta);
    float factor = (mode == ModeRadiance) ? square(1.0f / eta) : 1.0f;
    return (denom < 1e-7f) ? v3(0.0f)
           : baseColor * fabsG(ggxD(n, h, alpha) * smithG(n, wo, wi, alpha) * hCosWo * hCosWi) / denom *
                 (1.0f - refl) * factor;
}
ZGL_D float dielectricPdf(vec3 wo, vec3 wi, vec3 n, const BSDFParam& p)
{
    float roughness = p.roughness, ior = p.ior;
    if (approximateDelta(roughness)) return 0.0f;
    if (sameHemisphere(n, wo, wi)) {
        vec3 h = normalize(wo + wi);
        if (dot(wo, h) < 0.0f) return 0.0f;
        float refl = fresnelDielectric(absDot(h, wi), ior);
        return ggxPdfWm(n, h, wo, roughness * roughness) / (4.0f * absDot(h, wo)) * refl;
    }
    float eta = dot(n, wo) > 0.0f ? ior : 1.0f / ior;
    vec3 h = normalize(wo + wi * eta);
    if (sameHemisphere(h, wo, wi)) return 0.0f;
    float trans = 1.0f - fresnelDielectric(absDot(h, wo), eta);
    float dHdWi = absDot(h, wi) / square(dot(h, wo) + eta * dot(h, wi));
    return ggxPdfWm(n, h, wo, roughness * roughness) * dHdWi * trans;
}
ZGL_D BSDFSample dielectricSample(vec3 n, vec3 wo, const BSDFParam& p, uint32_t mode, vec3 u)
{
    vec3 baseColor = p.baseColor;
    float roughness = p.roughness, ior = p.ior;
    if (approximateDelta(roughness)) {
        float refl = fresnelDielectric(dot(n, wo), ior);
        if (u.x < refl) return makeSample(reflect(-wo, n), 1.0f, baseColor, 1.0f, FlagSpecRefl);
        vec3 wi;
        if (!refractDir(wi, wo, n, ior)) return invalidSample();
        if (dot(n, wo) < 0.0f) ior = 1.0f / ior;
        float factor = (mode == ModeRadiance) ? square(1.0f / ior) : 1.0f;
        return makeSample(wi, 1.0f, baseColor * factor, ior, FlagSpecTrans);
    }
    float alpha = roughness * roughness;
    vec3 h = ggxSampleWm(n, wo, alpha, v2(u.y, u.z));
    if (dot(n, h) < 0.0f) h = -h;
    float refl = fresnelDielectric(dot(h, wo), ior);
    if (u.x < refl) {
        vec3 wi = -reflect(wo, h);
        if (!sameHemisphere(n, wo, wi)) return invalidSample();
        float pp = ggxPdfWm(n, h, wo, alpha) / (4.0f * absDot(h, wo));
        float hCosWo = absDot(h, wo), hCosWi = absDot(h, wi);
        vec3 r = (hCosWo * hCosWi < 1e-7f) ? v3(0.0f)
                 : baseColor * ggxD(n, h, alpha) * smithG(n, wo, wi, alpha) / (4.0f * hCosWo * hCosWi);
        if (fisnan(pp)) pp = 0.0f;
        return makeSample(wi, pp, r, 1.0f, FlagGlosRefl);
    }
    vec3 wi;
    if (!refractDir(wi, wo, h, ior)) return invalidSample();
    if (sameHemisphere(n, wo, wi)) return invalidSample();
    if (absDot(n, wi) < 1e-10f) return invalidSample();
    float hCosWo = absDot(h, wo), hCosWi = absDot(h, wi);
    if (dot(h, wo) < 0.0f) ior = 1.0f / ior;
    float sqrtDenom = dot(h, wo) + ior * dot(h, wi);
    float denom = sqrtDenom * sqrtDenom;
    float dHdWi = hCosWi / denom;
    float factor = (mode == ModeRadiance) ? square(1.0f / ior) : 1.0f;
    denom *= absDot(n, wi) * absDot(n, wo);
    vec3 t = (denom < 1e-7f) ? v3(0.0f)
             : baseColor * fabsG(ggxD(n, h, alpha) * smithG(n, wo, wi, alpha) * hCosWo * hCosWi) / denom;
    float pp = ggxPdfWm(n, h, wo, alpha) * dHdWi;
    if (fisnan(pp)) pp = 0.0f;
    return makeSample(wi, pp, t * factor, ior, FlagGlosTrans);
}
ZGL_D BSDFSample thinDielectricSample(vec3 n, vec3 wo, const BSDFParam& p, vec3 u)
{
    if (dot(n, wo) < 0.0f) n = -n;
    float refl = fresnelDielectric(dot(n, wo), p.ior);
    float trans = 1.0f - refl;
    if (refl < 1.0f) {
        refl += trans * trans * refl / (1.0f - refl * refl);
        trans = 1.0f - refl;
    }
    return (u.x < refl) ? makeSample(reflect(-wo, n), 1.0f, p.baseColor, 1.0f, FlagSpecRefl)
                        : makeSample(-wo, 1.0f, p.baseColor, 1.0f, FlagSpecTrans);
}

ZGL_D vec3 principledMetal(vec3 wo, vec3 wi, vec3 n, vec3 fm0, float alpha)
{
    float cosWo = satDot(n, wo), cosWi = satDot(n, wi);
    vec3 h = normalize(wo + wi);
    if (cosWo < 1e-10f || cosWi < 1e-10f) return v3(0.0f);
    vec3 fm = schlickF(absDot(h, wo), fm0);
    float dm = ggxD(n, h, alpha);
    float gm = smithG(n, wo, wi, alpha);
    float denom = 4.0f * cosWi * cosWo;
    if (denom < 1e-7f) return v3(0.0f);
    return fm * dm * gm / denom;
}
ZGL_D float principledMetalPdf(vec3 wo, vec3 wi, vec3 n, float alpha)
{
    vec3 h = normalize(wo + wi);
    return ggxPdfVisibleWm(n, h, wo, alpha) / (4.0f * absDot(h, wo));
}
ZGL_D vec3 principledClearcoat(vec3 wo, vec3 wi, vec3 n, vec3 baseColor, float alpha)
{
    float cosWo = satDot(n, wo), cosWi = satDot(n, wi);
    vec3 h = normalize(wo + wi);
    if (cosWo < 1e-6f || cosWi < 1e-6f) return v3(0.0f);
    vec3 fc = schlickF(absDot(h, wo), baseColor);
    float dc = gtr1D(n, h, alpha);
    float gc = smithG(n, wo, wi, 0.25f);
    float denom = 4.0f * cosWi * cosWo;
    if (denom < 1e-7f) return v3(0.0f);
    return fc * dc * gc / denom;
}
ZGL_D float principledClearcoatPdf(vec3 wo, vec3 wi, vec3 n, float alpha)
{
    vec3 h = normalize(wo + wi);
    return gtr1PdfWm(n, h, wo, alpha) / (4.0f * absDot(h, wo));
}
ZGL_D vec3 principledDiffuse(vec3 wo, vec3 wi, vec3 n, vec3 baseColor, float subsurface, float roughness)
{
    float cosWo = satDot(n, wo), cosWi = satDot(n, wi);
    if (cosWo < 1e-10f || cosWi < 1e-10f) return v3(0.0f);
    vec3 h = normalize(wo + wi);
    float hCosWi = dot(h, wi);
    float hCosWi2 = hCosWi * hCosWi;
    float fi = schlickW(cosWi), fo = schlickW(cosWo);
    vec3 fd90 = v3(0.5f + 2.0f * roughness * hCosWi2);
    vec3 fd = mix(v3(1.0f), fd90, fi) * mix(v3(1.0f), fd90, fo);
    vec3 baseDiffuse = baseColor * fd * PiInv;
    vec3 fss90 = v3(roughness * hCosWi2);
    vec3 fss = mix(v3(1.0f), fss90, fi) * mix(v3(1.0f), fss90, fo);
    vec3 ss = baseColor * PiInv * 1.25f * (fss * (1.0f / (cosWi + cosWo) - 0.5f) + 0.5f);
    return mix(baseDiffuse, ss, subsurface);
}
ZGL_D vec3 principledTint(vec3 baseColor)
{
    float lum = luminance(baseColor);
    return lum > 0.0f ? baseColor / lum : v3(1.0f);
}
ZGL_D vec3 principledBRDF(vec3 wo, vec3 wi, vec3 n, const BSDFParam& p)
{
    vec3 res = v3(0.0f);
    float alpha = square(p.roughness);
    float clearcoatAlpha = mixf(0.1f, 0.001f, p.clearcoatGloss);
    vec3 tintColor = principledTint(p.baseColor);
    vec3 fm0 = mix(v3(1.0f), tintColor, p.specularTint) * (0.08f * p.specular);
    fm0 = mix(fm0, p.baseColor, p.metallic);
    float hCosWi = dot(normalize(wo + wi), wi);
    res += principledDiffuse(wo, wi, n, p.baseColor, p.subsurface, p.roughness) * (1.0f - p.metallic);
    res += principledMetal(wo, wi, n, fm0, alpha);
    res += principledClearcoat(wo, wi, n, p.baseColor, clearcoatAlpha) * p.clearcoat * 0.25f;
    res += mix(v3(1.0f), tintColor, p.sheenTint) * schlickW(hCosWi) * p.sheen * (dot(n, wi) < 0.0f ? 0.0f : 1.0f);
    return res;
}
ZGL_D float principledBRDFPdf(vec3 wo, vec3 wi, vec3 n, const BSDFParam& p)
{
    float pdf = 0.0f;
    float alpha = square(p.roughness);
    float clearcoatAlpha = mixf(0.1f, 0.001f, p.clearcoatGloss);
    float spec = 1.0f / (2.0f - p.metallic);
    float cosinePdf = absDot(n, wi) * PiInv;
    pdf += cosinePdf * (1.0f - spec);
    pdf += principledMetalPdf(wo, wi, n, alpha) * spec;
    pdf += principledClearcoatPdf(wo, wi, n, clearcoatAlpha) * 0.25f * p.clearcoat;
    return pdf / (1.0f + 0.25f * p.clearcoat);
}
ZGL_D BSDFSample principledBRDFSample(vec3 n, vec3 wo, const BSDFParam& p, vec3 u, Rng& rng)
{
    float alpha = square(p.roughness);
    float spec = 1.0f / (2.0f - p.metallic);
    float cdf0 = 1.0f - spec, cdf1 = 1.0f, cdf2 = 1.0f + p.clearcoat * 0.25f;
    float sv = rand01(rng) * cdf2;
    vec3 wi;
    if (sv <= cdf0) {
        vec4 c = sampleCosineWeighted(n, v2(u.y, u.z));
        wi = v3(c.x, c.y, c.z);
    } else if (sv <= cdf1) {
        vec3 h = ggxSampleVisibleWm(n, wo, alpha, v2(u.y, u.z));
        wi = reflect(-wo, h);
        if (dot(n, wi) < 0.0f) wi = v3(0.0f);  // InvalidBSDFSample.wi
    } else {
        vec3 h = gtr1SampleWm(n, wo, alpha, v2(u.y, u.z));
        wi = reflect(-wo, h);
        if (dot(n, wi) < 0.0f) wi = v3(0.0f);
    }
    vec3 bsdf = principledBRDF(wo, wi, n, p);
    float pdf = principledBRDFPdf(wo, wi, n, p);
    return makeSample(wi, pdf, bsdf, 1.0f, FlagDiffuse);
}

// material_loader.glsl:3-97.  The type is the int stored in texel 3's .y slot.
template <bool C> ZGL_D uint32_t loadMaterialType(const DevScene& s, int matId, Counters& cn)
{
    cnt<C>(cn, 16);
    return (uint32_t)__float_as_int(s.materials[matId * 4 + 3].y);
}
// texId != -1 takes baseColor from the texture array (loadBaseColor, material_loader.glsl:8-16).
template <bool C>
ZGL_D BSDFParam loadMaterial(const DevScene& s, uint32_t type, int matId, int texId, vec2 uv, Counters& cn)
{
    BSDFParam r = {};
    float4 t0 = s.materials[matId * 4 + 0];
    r.baseColor = (texId == -1) ? v3(t0.x, t0.y, t0.z) : textureLookup<C>(s, texId, uv, cn);
    if (type == MatPrincipled) {
        float4 t1 = s.materials[matId * 4 + 1], t2 = s.materials[matId * 4 + 2];
        cnt<C>(cn, 48);
        r.roughness = mixf(0.0134f, 1.0f, t0.w);
        r.subsurface = t1.x; r.metallic = t1.y; r.specular = t1.z; r.specularTint = t1.w;
        r.sheen = t2.x; r.sheenTint = t2.y; r.clearcoat = t2.z; r.clearcoatGloss = t2.w;
    } else if (type == MatMetalWorkflow) {
        cnt<C>(cn, 32);
        r.roughness = mixf(0.0134f, 1.0f, t0.w);
        r.metallic = s.materials[matId * 4 + 1].y;
    } else if (type == MatDielectric) {
        cnt<C>(cn, 32);
        r.roughness = t0.w;
        r.ior = s.materials[matId * 4 + 3].x;
    } else if (type == MatThinDielectric) {
        cnt<C>(cn, 32);
        r.ior = s.materials[matId * 4 + 3].x;
    } else {
        cnt<C>(cn, 16);
    }
    return r;
}
// material_loader.glsl:99-169
ZGL_D vec3 materialBSDF(uint32_t t, const BSDFParam& p, vec3 wo, vec3 wi, vec3 n, uint32_t mode)
{
    switch (t) {
    case MatPrincipled: return principledBRDF(wo, wi, n, p);
    case MatMetalWorkflow: return metalWorkflow(wo, wi, n, p);
    case MatDielectric: return dielectric(wo, wi, n, p, mode);
    case MatThinDielectric: return v3(0.0f);
    default: return lambertian(p);
    }
}
ZGL_D float materialPdf(uint32_t t, const BSDFParam& p, vec3 wo, vec3 wi, vec3 n)
{
    switch (t) {
    case MatPrincipled: return principledBRDFPdf(wo, wi, n, p);
    case MatMetalWorkflow: return metalWorkflowPdf(wo, wi, n, p);
    case MatDielectric: return dielectricPdf(wo, wi, n, p);
    case MatThinDielectric: return 0.0f;
    default: return lambertianPdf(wi, n);
    }
}
ZGL_D BSDFSample materialSample(uint32_t t, const BSDFParam& p, vec3 n, vec3 wo, uint32_t mode, vec3 u, Rng& rng)
{
    switch (t) {
    case MatPrincipled: return principledBRDFSample(n, wo, p, u, rng);
    case MatMetalWorkflow: return metalWorkflowSample(n, wo, p, u);
    case MatDielectric: return dielectricSample(n, wo, p, mode, u);
    case MatThinDielectric: return thinDielectricSample(n, wo, p, u);
    default: return lambertianSample(n, wo, p, u);
    }
}

// ---------------------------------------------------------------------------
// Lights and environment (light.glsl:48-247)
// ---------------------------------------------------------------------------
struct LightLiSample { vec3 wi, coef; float pdf; };
ZGL_D LightLiSample invalidLi() { return LightLiSample{v3(0.0f), v3(0.0f), 0.0f}; }
ZGL_D vec3 lightPowerOf(const DevScene& s, int id) { return ldV3(s.lightPower, id); }

template <bool C> ZGL_D int lightSampleOne(const DevScene& s, vec2 u, Counters& cn)
{
    int cx = (int)((float)s.nLightTriangles * u.x);
    cnt<C>(cn, 8);
    // texelFetch past the end of the buffer reads 0 (u.x == 1.0)
    float prob = (cx >= 0 && cx < s.nLightTriangles) ? s.lightProb[cx] : 0.0f;
    int alias = (cx >= 0 && cx < s.nLightTriangles) ? s.lightAlias[cx] : 0;
    return (u.y < prob) ? cx : alias;
}
template <bool C> ZGL_D float lightPdfSampleOne(const DevScene& s, int id, Counters& cn)
{
    cnt<C>(cn, 12);
    return luminance(lightPowerOf(s, id)) / s.lightSum;
}
template <bool C> ZGL_D vec3 lightLe(const DevScene& s, int id, vec3 x, vec3 wo, Counters& cn)
{
    int triId = id + s.objPrimCount;
    SurfaceInfo si = triangleSurfaceInfo<C>(s, triId, x, cn);
    if (dot(wo, si.ng) <= 0.0f) return v3(0.0f);
    cnt<C>(cn, 12);
    return lightPowerOf(s, id) / triangleAreaId<C>(s, triId, cn) * 0.5f * PiInv;
}
template <bool C> ZGL_D float lightPdfLi(const DevScene& s, int id, vec3 x, vec3 y, Counters& cn)
{
    int triId = id + s.objPrimCount;
    vec3 norm = triangleSurfaceInfo<C>(s, triId, y, cn).ng;
    vec3 yx = normalize(x - y);
    float cosTheta = absDot(norm, yx);
    if (cosTheta < 1e-8f) return -1.0f;
    return distSquare(x, y) / (triangleAreaId<C>(s, triId, cn) * cosTheta);
}
struct LightPdf { float pdfPos, pdfDir; };
template <bool C> ZGL_D LightPdf lightPdfLe(const DevScene& s, int id, const Ray& ray, Counters& cn)
{
    int triId = id + s.objPrimCount;
    vec3 norm = triangleSurfaceInfo<C>(s, triId, ray.ori, cn).ng;
    LightPdf r;
    r.pdfPos = 1.0f / triangleAreaId<C>(s, triId, cn);
    r.pdfDir = (dot(norm, ray.dir) <= 0.0f) ? 0.0f : 0.5f * PiInv;
    return r;
}
struct LightLeSample { Ray ray; vec3 Le; float pdfPos, pdfDir; };
template <bool C> ZGL_D LightLeSample lightSampleOneLe(const DevScene& s, int id, vec4 u, Counters& cn)
{
    int triId = id + s.objPrimCount;
    vec3 ori = triangleSampleUniformId<C>(s, triId, v2(u.x, u.y), cn);
    vec3 norm = triangleSurfaceInfo<C>(s, triId, ori, cn).ng;
    vec4 samp = sampleCosineWeighted(norm, v2(u.z, u.w));
    vec3 dir = v3(samp.x, samp.y, samp.z);
    LightLeSample r;
    r.ray = rayOffseted(ori, dir);
    r.Le = lightLe<C>(s, id, ori, dir, cn);
    r.pdfPos = 1.0f / triangleAreaId<C>(s, triId, cn);
    r.pdfDir = samp.w;
    return r;
}
template <bool C> ZGL_D LightLiSample lightSampleLi(const DevScene& s, int id, vec3 x, vec2 u, Counters& cn)
{
    int triId = id + s.objPrimCount;
    vec3 a, b, c;
    triangleVerts<C>(s, triId, a, b, c, cn);
    vec3 y = sampleTriangleUniform(a, b, c, u);
    vec3 wi = normalize(y - x);
    vec3 norm = triangleSurfaceInfo<C>(s, triId, y, cn).ng;
    float cosTheta = dot(norm, -wi);
    if (cosTheta < 1e-6f) return invalidLi();
    Ray lightRay = rayOffseted(x, wi);
    float dist = distance(x, y);
    float pdf = dist * dist / (triangleArea(a, b, c) * cosTheta);
    float testDist = dist - 1e-4f - 1e-6f;
    if (bvhTest<C>(s, lightRay, testDist, cn) || pdf < 1e-8f) return invalidLi();
    vec3 weight = lightLe<C>(s, id, y, -wi, cn);
    float pdfSample = lightPdfSampleOne<C>(s, id, cn);
    pdf *= pdfSample;
    return LightLiSample{wi, weight / pdf, pdf};
}

template <bool C> ZGL_D vec3 envLe(const DevScene& s, vec3 wi, Counters& cn)
{
    wi = rotateZcs(wi, s.envRotCosN, s.envRotSinN);
    return envLookup<C>(s, sphereToPlane(wi), cn);
}
template <bool C> ZGL_D float envGetPortion(const DevScene& s, vec3 wi, Counters& cn)
{
    return luminance(envLe<C>(s, wi, cn)) / s.envSum;
}
template <bool C> ZGL_D float envPdfLi(const DevScene& s, vec3 wi, Counters& cn)
{
    if (s.envSum == 0.0f) return 0.0f;
    return envGetPortion<C>(s, wi, cn) * (float)s.envW * (float)s.envH * 0.5f * square(PiInv);
}
ZGL_D float envProbAt(const DevScene& s, int x, int y)
{
    if (x < 0 || x > s.envW || y < 0 || y >= s.envH) return 0.0f;
    return s.envProb[(size_t)y * (s.envW + 1) + x];
}
ZGL_D int envAliasAt(const DevScene& s, int x, int y)
{
    if (x < 0 || x > s.envW || y < 0 || y >= s.envH) return 0;
    return s.envAlias[(size_t)y * (s.envW + 1) + x];
}
template <bool C> ZGL_D vec4 envSampleWi(const DevScene& s, vec4 u, Counters& cn)
{
    const int w = s.envW, h = s.envH;
    int rx = (int)((float)h * u.x);
    int row = (u.y < envProbAt(s, w, rx)) ? rx : envAliasAt(s, w, rx);
    int cx = (int)((float)w * u.z);
    int col = (u.w < envProbAt(s, cx, row)) ? cx : envAliasAt(s, cx, row);
    cnt<C>(cn, 16);
    vec2 uv = v2((float)col + 0.5f, (float)row + 0.5f) / v2((float)w, (float)h);
    vec3 wi = planeToSphere(uv);
    wi = rotateZcs(wi, s.envRotCos, s.envRotSin);
    float pdf = envGetPortion<C>(s, wi, cn) * (float)w * (float)h * 0.5f * square(PiInv);
    return v4(wi.x, wi.y, wi.z, pdf);
}
template <bool C> ZGL_D LightLiSample envSampleLi(const DevScene& s, vec3 x, vec4 u, Counters& cn)
{
    vec4 sp = envSampleWi<C>(s, u, cn);
    vec3 wi = v3(sp.x, sp.y, sp.z);
    float pdf = sp.w;
    Ray ray = rayOffseted(x, wi);
    if (bvhTest<C>(s, ray, 1e8f, cn) || pdf == 0.0f) return invalidLi();
    return LightLiSample{wi, envLe<C>(s, wi, cn) / pdf, pdf};
}
ZGL_D float pdfSampleLightOf(const DevScene& s, bool uniform, float portion)
{
    if (s.nLightTriangles > 0) return uniform ? portion : s.lightSum / (s.lightSum + s.envSum);
    return 0.0f;
}
template <bool C>
ZGL_D LightLiSample sampleLightAndEnv(const DevScene& s, vec3 x, float ud, vec4 us, bool uniform, float portion,
                                      Counters& cn)
{
    float pdfSampleLight = pdfSampleLightOf(s, uniform, portion);
    bool sampleLight = ud < pdfSampleLight;
    float pdfSelect = sampleLight ? pdfSampleLight : 1.0f - pdfSampleLight;
    LightLiSample r;
    if (sampleLight) {
        int id = lightSampleOne<C>(s, v2(us.x, us.y), cn);
        r = lightSampleLi<C>(s, id, x, v2(us.z, us.w), cn);
    } else {
        r = envSampleLi<C>(s, x, us, cn);
    }
    r.coef /= pdfSelect;
    r.pdf *= pdfSelect;
    return r;
}
template <bool C>
ZGL_D float pdfSelectLight(const DevScene& s, int id, bool uniform, float portion, Counters& cn)
{
    float fst = lightPdfSampleOne<C>(s, id, cn);
    float snd = uniform ? portion : s.lightSum / (s.lightSum + s.envSum);
    return fst * snd;
}
ZGL_D float pdfSelectEnv(const DevScene& s, bool uniform, float portion)
{
    return uniform ? (1.0f - portion) : s.envSum / (s.lightSum + s.envSum);
}

// ---------------------------------------------------------------------------
// Thin-lens camera (camera.glsl:52-142)
// ---------------------------------------------------------------------------
ZGL_D Ray cameraSampleRay(const DevScene& s, vec2 uv, vec4 u)
{
    vec2 texelSize = v2(1.0f / (float)s.filmW, 1.0f / (float)s.filmH);
    vec2 biased = uv + texelSize * v2(u.x, u.y);
    vec2 ndc = biased * 2.0f - 1.0f;
    vec2 disk = toConcentricDisk(v2(u.z, u.w));
    vec3 pLens = v3(disk.x * s.lensRadius, disk.y * s.lensRadius, 0.0f);
    vec2 fp = ndc * v2(s.camAsp, 1.0f) * s.focalDist * s.tanFOV;
    vec3 pFocus = v3(fp.x, fp.y, s.focalDist);
    vec3 dir = pFocus - pLens;
    dir = normalize((s.camR * dir.x + s.camU * dir.y) + s.camF * dir.z);
    Ray r;
    r.ori = (s.camPos + s.camR * pLens.x) + s.camU * pLens.y;
    r.dir = dir;
    return r;
}

// Per-pixel seeds (path_integ_naive.glsl:152-164): a double bilinear lookup into
// the noise texture picks the scramble and the rand() seed.
struct PixelSeed { uint32_t tx, ty; };
ZGL_D PixelSeed pixelSeed(const DevScene& s, vec2 scr)
{
    vec2 nc = noiseLookup(s, scr);
    nc = noiseLookup(s, nc);
    vec2 tc = v2((float)s.filmW, (float)s.filmH) * nc;
    return PixelSeed{(uint32_t)tc.x, (uint32_t)tc.y};
}
ZGL_D int sampleOffsetOf(int spp)
{
    int off = spp * SobolSampleDim;
    if (off > SobolSampleNum * SobolSampleDim) off -= SobolSampleNum * SobolSampleDim;
    return off;
}
ZGL_D bool inFilmBound(vec2 uv) { return uv.x >= 0.0f && uv.x <= 1.0f && uv.y >= 0.0f && uv.y <= 1.0f; }
ZGL_D bool cameraDelta(const DevScene& s) { return s.lensRadius <= 1e-6f; }
ZGL_D vec2 cameraRasterPos(const DevScene& s, const Ray& ray)
{
    float cosTheta = dot(ray.dir, s.camF);
    float dFocus = s.focalDist / cosTheta;
    vec3 pFocus = mul(s.camMatInv, rayPoint(ray, dFocus) - s.camPos);
    float aspect = (float)s.filmW / (float)s.filmH;
    vec3 den = v3(aspect * s.tanFOV, 1.0f * s.tanFOV, 1.0f) * s.focalDist;
    pFocus = pFocus / den;
    return v2((pFocus.x + 1.0f) * 0.5f, (pFocus.y + 1.0f) * 0.5f);
}
ZGL_D vec3 cameraIe(const DevScene& s, const Ray& ray)
{
    float cosTheta = dot(ray.dir, s.camF);
    if (cosTheta < 1e-6f) return v3(0.0f);
    vec2 pRaster = cameraRasterPos(s, ray);
    if (!inFilmBound(pRaster)) return v3(0.0f);
    float tanFOVInv = 1.0f / s.tanFOV;
    float cos2Theta = cosTheta * cosTheta;
    float lensArea = cameraDelta(s) ? 1.0f : Pi * s.lensRadius * s.lensRadius;
    return v3(0.25f) * square(tanFOVInv / cos2Theta) / (lensArea * s.camAsp);
}
struct CameraIiSample { vec3 wi, Ii; float dist; vec2 uv; float pdf; };
ZGL_D CameraIiSample cameraSampleIi(const DevScene& s, vec3 ref, vec2 u)
{
    vec2 disk = toConcentricDisk(u);
    vec3 pLens = v3(disk.x * s.lensRadius, disk.y * s.lensRadius, 0.0f);
    vec3 y = ((s.camPos + s.camR * pLens.x) + s.camU * pLens.y) + s.camF * pLens.z;
    float dist = distance(ref, y);
    vec3 wi = normalize(y - ref);
    float cosTheta = satDot(s.camF, -wi);
    if (cosTheta < 1e-6f) return CameraIiSample{v3(0.0f), v3(0.0f), 0.0f, v2(0.0f, 0.0f), 0.0f};
    Ray ray = makeRay(y, -wi);
    vec3 Ie = cameraIe(s, ray);
    vec2 uv = cameraRasterPos(s, ray);
    float lensArea = cameraDelta(s) ? 1.0f : Pi * s.lensRadius * s.lensRadius;
    float pdf = dist * dist / (cosTheta * lensArea);
    return CameraIiSample{wi, Ie, dist, uv, pdf};
}
// Film splat of the light-path integrators (light_path_integ.glsl:34-43, addImage); the
// film is planar R32F of (3W) x H, added to with FP32 atomics.
ZGL_D void splatFilm(const DevScene& s, float* film, vec2 uv, vec3 res)
{
    if (!inFilmBound(uv)) return;
    int ix = (int)(uv.x * (float)s.filmW), iy = (int)(uv.y * (float)s.filmH);
    if (ix >= s.filmW || iy >= s.filmH) return;  // out-of-range image atomics are dropped
    float* p = film + (size_t)iy * 3 * s.filmW + 3 * ix;
    unsafeAtomicAdd(p + 0, res.x);
    unsafeAtomicAdd(p + 1, res.y);
    unsafeAtomicAdd(p + 2, res.z);
}
struct CameraPdf { float pdfPos, pdfDir; };
ZGL_D CameraPdf cameraPdfIe(const DevScene& s, const Ray& ray)
{
    float cosTheta = dot(s.camF, ray.dir);
    if (cosTheta < 1e-6f) return CameraPdf{0.0f, 0.0f};
    vec2 pRaster = cameraRasterPos(s, ray);
    if (!inFilmBound(pRaster)) return CameraPdf{0.0f, 0.0f};
    float pdfPos = cameraDelta(s) ? 1.0f : 1.0f / (Pi * s.lensRadius * s.lensRadius);
    float pdfDir = 1.0f / (cosTheta * cosTheta * cosTheta);
    return CameraPdf{pdfPos, pdfDir};
}

}  // namespace zgl
