// zgl_math.h -- FP32 vector math with GLSL 4.50 semantics for the rendering kernels.
//
// Every operation here is written so that the same source gives the same bits on
// gfx950 and on an x86-64 host when both are compiled with -ffp-contract=off:
//   * +, -, *, / and sqrt are IEEE-754 correctly rounded on both targets;
//   * min/max follow the GLSL specification (min(x,y) = y < x ? y : x), not the
//     IEEE minNum of the hardware instructions, so NaN operands behave the same;
//   * dot/length/normalize/mix have one fixed evaluation order;
//   * sin/cos/atan2/log/exp are computed by the routines below (Cody-Waite range
//     reduction followed by truncated Taylor series), not by libm or ocml, whose
//     last-ulp behaviour differs between the two targets.
//
// The routines follow the reference's GLSL library src/shader/math.glsl:1-219.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define ZGL_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define ZGL_HD inline
#endif

namespace zgl {

constexpr float Pi = 3.14159265358979323846f;
constexpr float PiInv = 1.0f / 3.14159265358979323846f;  // float division, as the GLSL const

ZGL_HD float fminG(float x, float y) { return (y < x) ? y : x; }
ZGL_HD float fmaxG(float x, float y) { return (x < y) ? y : x; }
ZGL_HD float fclampG(float x, float lo, float hi) { return fminG(fmaxG(x, lo), hi); }
ZGL_HD float fabsG(float x) { return __builtin_fabsf(x); }
ZGL_HD float fsqrtG(float x) { return __builtin_sqrtf(x); }
ZGL_HD float ffloorG(float x) { return __builtin_floorf(x); }
ZGL_HD bool fisnan(float x) { return __builtin_isnan(x); }
ZGL_HD float square(float x) { return x * x; }
ZGL_HD float pow5(float x) { float x2 = x * x; return x2 * x2 * x; }
ZGL_HD float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

ZGL_HD uint32_t fbits(float x) { return __builtin_bit_cast(uint32_t, x); }
ZGL_HD float bitsf(uint32_t x) { return __builtin_bit_cast(float, x); }
ZGL_HD int32_t fbitsi(float x) { return __builtin_bit_cast(int32_t, x); }
ZGL_HD float ibitsf(int32_t x) { return __builtin_bit_cast(float, x); }

// ---------------------------------------------------------------------------
// Transcendentals.  sin/cos: x = k*pi/2 + r with pi/2 split in three parts so that
// k*C1 and k*C2 are exact for |k| < 2^15; sin and cos of r in [-pi/4, pi/4] by their
// Taylor series through r^9 / r^10 (truncation error < 2e-9).
// ---------------------------------------------------------------------------
ZGL_HD void sincosZ(float x, float* s, float* c)
{
    const float TwoOverPi = 0.636619772367581343f;
    const float C1 = 1.5703125f;
    const float C2 = 4.837512969970703125e-4f;
    const float C3 = 7.54978995489188216e-8f;
    float kf = ffloorG(x * TwoOverPi + 0.5f);
    int k = (int)kf;
    float r = ((x - kf * C1) - kf * C2) - kf * C3;
    float r2 = r * r;
    float sr = r + r * r2 * (-1.66666666666666667e-1f + r2 * (8.33333333333333333e-3f +
               r2 * (-1.98412698412698413e-4f + r2 * 2.75573192239858907e-6f)));
    float cr = 1.0f - 0.5f * r2 + r2 * r2 * (4.16666666666666667e-2f + r2 * (-1.38888888888888889e-3f +
               r2 * (2.48015873015873016e-5f + r2 * -2.75573192239858907e-7f)));
    switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}
ZGL_HD float sinZ(float x) { float s, c; sincosZ(x, &s, &c); return s; }
ZGL_HD float cosZ(float x) { float s, c; sincosZ(x, &s, &c); return c; }

// atan on [0, 1]: reduce t > tan(pi/8) with atan(t) = pi/4 + atan((t-1)/(t+1)), then
// the Taylor series through t^19 on |t| <= 0.4143 (truncation error < 1e-9).
ZGL_HD float atanUnitZ(float t)
{
    float off = 0.0f;
    if (t > 0.41421356237309505f) { t = (t - 1.0f) / (t + 1.0f); off = 0.785398163397448310f; }
    float t2 = t * t;
    float p = -1.0f / 19.0f;
    p = 1.0f / 17.0f + t2 * p;
    p = -1.0f / 15.0f + t2 * p;
    p = 1.0f / 13.0f + t2 * p;
    p = -1.0f / 11.0f + t2 * p;
    p = 1.0f / 9.0f + t2 * p;
    p = -1.0f / 7.0f + t2 * p;
    p = 1.0f / 5.0f + t2 * p;
    p = -1.0f / 3.0f + t2 * p;
    return off + (t + t * t2 * p);
}

// GLSL atan(y, x); (0, 0) returns 0.
ZGL_HD float atan2Z(float y, float x)
{
    float ax = fabsG(x), ay = fabsG(y);
    float mx = fmaxG(ax, ay), mn = fminG(ax, ay);
    if (mx == 0.0f) return 0.0f;
    float a = atanUnitZ(mn / mx);
    if (ay > ax) a = 1.57079632679489662f - a;
    if (x < 0.0f) a = 3.14159265358979324f - a;
    if (y < 0.0f) a = -a;
    return a;
}

// Natural log for finite x > 0: x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
// log(m) = 2 atanh(f), f = (m-1)/(m+1), series through f^11.
ZGL_HD float logZ(float x)
{
    if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    uint32_t b = fbits(x);
    int e = (int)((b >> 23) & 0xff);
    if (e == 0) {  // subnormal: scale up by 2^23
        x = x * 8388608.0f;
        b = fbits(x);
        e = (int)((b >> 23) & 0xff) - 23;
    }
    e -= 127;
    float m = bitsf((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356237309505f) { m = m * 0.5f; e += 1; }
    float f = (m - 1.0f) / (m + 1.0f);
    float f2 = f * f;
    float p = 1.0f / 11.0f;
    p = 1.0f / 9.0f + f2 * p;
    p = 1.0f / 7.0f + f2 * p;
    p = 1.0f / 5.0f + f2 * p;
    p = 1.0f / 3.0f + f2 * p;
    float lm = 2.0f * f + 2.0f * f * f2 * p;
    float ef = (float)e;
    return ef * 0.693145751953125f + (lm + ef * 1.42860682030941723e-6f);
}

// exp: x = k ln2 + r, |r| <= ln2/2, Taylor through r^8, then scale by 2^k.
ZGL_HD float expZ(float x)
{
    if (fisnan(x)) return x;
    if (x > 88.7228f) return __builtin_inff();
    if (x < -103.97f) return 0.0f;
    float kf = ffloorG(x * 1.44269504088896341f + 0.5f);
    int k = (int)kf;
    float r = (x - kf * 0.693145751953125f) - kf * 1.42860682030941723e-6f;
    float p = 1.0f / 40320.0f;
    p = 1.0f / 5040.0f + r * p;
    p = 1.0f / 720.0f + r * p;
    p = 1.0f / 120.0f + r * p;
    p = 1.0f / 24.0f + r * p;
    p = 1.0f / 6.0f + r * p;
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    // 2^k in two steps so that k down to -149 stays representable
    int k1 = k / 2, k2 = k - k1;
    return p * bitsf((uint32_t)(k1 + 127) << 23) * bitsf((uint32_t)(k2 + 127) << 23);
}

// pow for x > 0 (the only use is the GTR1 sample with alpha in (0, 1]).
ZGL_HD float powZ(float x, float y)
{
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : __builtin_inff();
    return expZ(y * logZ(x));
}

// ---------------------------------------------------------------------------
// Small vectors.
// ---------------------------------------------------------------------------
struct vec2 { float x, y; };
struct vec3 { float x, y, z; };
struct vec4 { float x, y, z, w; };

ZGL_HD vec2 v2(float x, float y) { return vec2{x, y}; }
ZGL_HD vec3 v3(float x, float y, float z) { return vec3{x, y, z}; }
ZGL_HD vec3 v3(float s) { return vec3{s, s, s}; }
ZGL_HD vec4 v4(float x, float y, float z, float w) { return vec4{x, y, z, w}; }

ZGL_HD vec3 operator+(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
ZGL_HD vec3 operator-(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
ZGL_HD vec3 operator*(vec3 a, vec3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
ZGL_HD vec3 operator/(vec3 a, vec3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
ZGL_HD vec3 operator*(vec3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
ZGL_HD vec3 operator*(float s, vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
ZGL_HD vec3 operator/(vec3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
ZGL_HD vec3 operator+(vec3 a, float s) { return v3(a.x + s, a.y + s, a.z + s); }
ZGL_HD vec3 operator-(vec3 a, float s) { return v3(a.x - s, a.y - s, a.z - s); }
ZGL_HD vec3 operator-(vec3 a) { return v3(-a.x, -a.y, -a.z); }
// Per-component select.  A `c ? a : b` on the struct itself can be lowered to a pick
// between two stack copies (scratch memory); three float selects stay in registers.
ZGL_HD vec3 select3(bool c, vec3 a, vec3 b)
{
    return v3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}
ZGL_HD vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
ZGL_HD vec3& operator*=(vec3& a, vec3 b) { a = a * b; return a; }
ZGL_HD vec3& operator*=(vec3& a, float s) { a = a * s; return a; }
ZGL_HD vec3& operator/=(vec3& a, float s) { a = a / s; return a; }

ZGL_HD vec2 operator+(vec2 a, vec2 b) { return v2(a.x + b.x, a.y + b.y); }
ZGL_HD vec2 operator-(vec2 a, vec2 b) { return v2(a.x - b.x, a.y - b.y); }
ZGL_HD vec2 operator*(vec2 a, vec2 b) { return v2(a.x * b.x, a.y * b.y); }
ZGL_HD vec2 operator/(vec2 a, vec2 b) { return v2(a.x / b.x, a.y / b.y); }
ZGL_HD vec2 operator*(vec2 a, float s) { return v2(a.x * s, a.y * s); }
ZGL_HD vec2 operator-(vec2 a, float s) { return v2(a.x - s, a.y - s); }

ZGL_HD float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ZGL_HD float dot(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
ZGL_HD vec3 cross(vec3 a, vec3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
ZGL_HD float length(vec3 a) { return fsqrtG(dot(a, a)); }
ZGL_HD float length(vec2 a) { return fsqrtG(dot(a, a)); }
ZGL_HD float distance(vec3 a, vec3 b) { return length(a - b); }
// normalize(v) = v * (1 / |v|)
ZGL_HD vec3 normalize(vec3 a) { float inv = 1.0f / fsqrtG(dot(a, a)); return a * inv; }
ZGL_HD vec3 vmin(vec3 a, vec3 b) { return v3(fminG(a.x, b.x), fminG(a.y, b.y), fminG(a.z, b.z)); }
ZGL_HD vec3 vmax(vec3 a, vec3 b) { return v3(fmaxG(a.x, b.x), fmaxG(a.y, b.y), fmaxG(a.z, b.z)); }
ZGL_HD vec3 vabs(vec3 a) { return v3(fabsG(a.x), fabsG(a.y), fabsG(a.z)); }
ZGL_HD vec3 mix(vec3 x, vec3 y, float a) { return x * (1.0f - a) + y * a; }
// GLSL reflect(I, N) = I - 2 dot(N, I) N
ZGL_HD vec3 reflect(vec3 i, vec3 n) { return i - n * (2.0f * dot(n, i)); }
ZGL_HD float satDot(vec3 a, vec3 b) { return fmaxG(dot(a, b), 0.0f); }
ZGL_HD float absDot(vec3 a, vec3 b) { return fabsG(dot(a, b)); }
ZGL_HD float distSquare(vec3 x, vec3 y) { return dot(x - y, x - y); }
ZGL_HD float maxComponent(vec3 v) { return fmaxG(v.x, fmaxG(v.y, v.z)); }
// Rec. 601 weights, used for light selection and path logic (math.glsl:193-196)
ZGL_HD float luminance(vec3 c) { return dot(c, v3(0.299f, 0.587f, 0.114f)); }
ZGL_HD bool isBlack(vec3 c) { return luminance(c) < 1e-5f; }
ZGL_HD bool hasNan(vec3 c) { return fisnan(c.x) || fisnan(c.y) || fisnan(c.z); }
ZGL_HD bool sameHemisphere(vec3 n, vec3 a, vec3 b) { return dot(n, a) * dot(n, b) > 0.0f; }

// Column-major 3x3 matrix, M * v = c0 v.x + c1 v.y + c2 v.z.
struct mat3 { vec3 c0, c1, c2; };
ZGL_HD vec3 mul(const mat3& m, vec3 v) { return (m.c0 * v.x + m.c1 * v.y) + m.c2 * v.z; }
// General inverse by cofactors over the determinant (the GLSL builtin inverse()).
ZGL_HD mat3 inverse(const mat3& m)
{
    float a00 = m.c0.x, a01 = m.c0.y, a02 = m.c0.z;
    float a10 = m.c1.x, a11 = m.c1.y, a12 = m.c1.z;
    float a20 = m.c2.x, a21 = m.c2.y, a22 = m.c2.z;
    float b01 = a22 * a11 - a12 * a21;
    float b11 = -a22 * a10 + a12 * a20;
    float b21 = a21 * a10 - a11 * a20;
    float det = a00 * b01 + a01 * b11 + a02 * b21;
    float id = 1.0f / det;
    mat3 r;
    r.c0 = v3(b01 * id, (-a22 * a01 + a02 * a21) * id, (a12 * a01 - a02 * a11) * id);
    r.c1 = v3(b11 * id, (a22 * a00 - a02 * a20) * id, (-a12 * a00 + a02 * a10) * id);
    r.c2 = v3(b21 * id, (-a21 * a00 + a01 * a20) * id, (a11 * a00 - a01 * a10) * id);
    return r;
}

// math.glsl:73-89
ZGL_HD mat3 tbnMatrix(vec3 n)
{
    vec3 t = (fabsG(n.z) > 0.999f) ? v3(0.0f, 1.0f, 0.0f) : v3(0.0f, 0.0f, 1.0f);
    vec3 b = normalize(cross(n, t));
    t = cross(b, n);
    return mat3{t, b, n};
}
ZGL_HD vec3 normalToWorld(vec3 n, vec3 v) { return normalize(mul(tbnMatrix(n), v)); }

// math.glsl:25-41 (Shirley-Chiu concentric map)
ZGL_HD vec2 toConcentricDisk(vec2 v)
{
    if (v.x == 0.0f && v.y == 0.0f) return v2(0.0f, 0.0f);
    v = v * 2.0f - 1.0f;
    float phi, r;
    if (v.x * v.x > v.y * v.y) {
        r = v.x;
        phi = Pi * v.y / v.x * 0.25f;
    } else {
        r = v.y;
        phi = Pi * 0.5f - Pi * v.x / v.y * 0.25f;
    }
    float s, c;
    sincosZ(phi, &s, &c);
    return v2(r * c, r * s);
}

// math.glsl:99-105
ZGL_HD vec4 sampleCosineWeighted(vec3 n, vec2 u)
{
    vec2 uv = toConcentricDisk(u);
    float z = fsqrtG(1.0f - dot(uv, uv));
    vec3 v = normalToWorld(n, v3(uv.x, uv.y, z));
    return v4(v.x, v.y, v.z, PiInv * z);
}

// math.glsl:58-71 (z-up equirect)
ZGL_HD vec2 sphereToPlane(vec3 uv)
{
    float theta = atan2Z(uv.y, uv.x);
    if (theta < 0.0f) theta += Pi * 2.0f;
    float phi = atan2Z(length(v2(uv.x, uv.y)), uv.z);
    return v2(theta * PiInv * 0.5f, phi * PiInv);
}
ZGL_HD vec3 planeToSphere(vec2 uv)
{
    float theta = uv.x * Pi * 2.0f;
    float phi = uv.y * Pi;
    float st, ct, sp, cp;
    sincosZ(theta, &st, &ct);
    sincosZ(phi, &sp, &cp);
    return v3(ct * sp, st * sp, cp);
}
// rotateZ with precomputed cos/sin of the angle (math.glsl:180-185)
ZGL_HD vec3 rotateZcs(vec3 v, float cost, float sint)
{
    return v3(v.x * cost - v.y * sint, v.x * sint + v.y * cost, v.z);
}

// math.glsl:112-131; table index for the MTBVH traversal of direction dir
ZGL_HD int cubemapFace(vec3 dir)
{
    vec3 v = vabs(dir);
    int maxDim = (v.x > v.y) ? (v.x > v.z ? 0 : 2) : (v.y > v.z ? 1 : 2);
    if (maxDim == 0) return dir.x > 0.0f ? 0 : 1;
    if (maxDim == 1) return dir.y > 0.0f ? 2 : 3;
    return dir.z > 0.0f ? 4 : 5;
}

// math.glsl:139-145
ZGL_HD vec3 sampleTriangleUniform(vec3 va, vec3 vb, vec3 vc, vec2 uv)
{
    float r = fsqrtG(uv.y);
    float u = 1.0f - r;
    float v = uv.x * r;
    return (va * (1.0f - u - v) + vb * u) + vc * v;
}
ZGL_HD float triangleArea(vec3 va, vec3 vb, vec3 vc) { return 0.5f * length(cross(vc - va, vb - va)); }

// random.glsl:3-13 -- 32-bit integer hash
ZGL_HD uint32_t hashU(uint32_t seed)
{
    seed = (seed ^ 61u) ^ (seed >> 16u);
    seed *= 9u;
    seed = seed ^ (seed >> 4u);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15u);
    return seed;
}
ZGL_HD float u32ToUnit(uint32_t r) { return (float)r * (1.0f / 4294967296.0f); }

}  // namespace zgl
