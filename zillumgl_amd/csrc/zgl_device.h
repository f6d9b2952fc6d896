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
    float envRotation;  // uEnvRotation, the angle the cos/sin above are of
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
    // sampleOffset + dim, wrapping as GLSL int addition does
    uint32_t v = sobolValue(s, (int)((uint32_t)r.sampleOffset + (uint32_t)r.dim));
    cnt<C>(cn, 4);
    v ^= r.sampleSeed;
    r.sampleSeed = hashU(r.sampleSeed);
    r.dim++;
    return u32ToUnit(v);
}
// cos/sin of the environment rotation (negated: the lookup's) in the working precision:
// the FP32 library reads the values the host computed with the FP32 routines, the double
// instance computes them from the angle with the FP64 routines.
ZGL_D void envRotCS(const DevScene& s, float& c, float& sn) { c = s.envRotCos; sn = s.envRotSin; }
ZGL_D void envRotCSN(const DevScene& s, float& c, float& sn) { c = s.envRotCosN; sn = s.envRotSinN; }
ZGL_D void envRotCS(const DevScene& s, double& c, double& sn) { sincosF64((double)s.envRotation, &sn, &c); }
ZGL_D void envRotCSN(const DevScene& s, double& c, double& sn) { sincosF64(-(double)s.envRotation, &sn, &c); }

// the rendering path's instance of the library (made zgl's by the using-directive of
// zgl_math.h)
namespace f32 {
#include "zgl_device_body.inc"
}  // namespace f32

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

}  // namespace zgl
