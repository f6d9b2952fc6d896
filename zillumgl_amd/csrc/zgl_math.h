// zgl_math.h -- vector math with GLSL 4.50 semantics for the rendering kernels.
//
// The math itself is in zgl_math_body.inc, written over a working precision `real`:
// namespace zgl instantiates it with real = float (the rendering path, described
// below); zgl_f64.h instantiates it again with real = double in zgl::f64 for the
// converged FP64 reference, whose sin/cos/atan2/log/exp/pow are double routines by the
// same methods.
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

ZGL_HD uint32_t fbits(float x) { return __builtin_bit_cast(uint32_t, x); }
ZGL_HD float bitsf(uint32_t x) { return __builtin_bit_cast(float, x); }
ZGL_HD int32_t fbitsi(float x) { return __builtin_bit_cast(int32_t, x); }
ZGL_HD float ibitsf(int32_t x) { return __builtin_bit_cast(float, x); }

// ---------------------------------------------------------------------------
// Transcendentals.  sin/cos: x = k*pi/2 + r with pi/2 split in three parts so that
// k*C1 and k*C2 are exact for |k| < 2^15; sin and cos of r in [-pi/4, pi/4] by their
// Taylor series through r^9 / r^10 (truncation error < 2e-9).
// ---------------------------------------------------------------------------
ZGL_HD void sincosF32(float x, float* s, float* c)
{
    const float TwoOverPi = 0.636619772367581343f;
    const float C1 = 1.5703125f;
    const float C2 = 4.837512969970703125e-4f;
    const float C3 = 7.54978995489188216e-8f;
    float kf = __builtin_floorf(x * TwoOverPi + 0.5f);
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
ZGL_HD float sinF32(float x) { float s, c; sincosF32(x, &s, &c); return s; }
ZGL_HD float cosF32(float x) { float s, c; sincosF32(x, &s, &c); return c; }

// atan on [0, 1]: reduce t > tan(pi/8) with atan(t) = pi/4 + atan((t-1)/(t+1)), then
// the Taylor series through t^19 on |t| <= 0.4143 (truncation error < 1e-9).
ZGL_HD float atanUnitF32(float t)
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
ZGL_HD float atan2F32(float y, float x)
{
    float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    float mx = (ax < ay) ? ay : ax, mn = (ay < ax) ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    float a = atanUnitF32(mn / mx);
    if (ay > ax) a = 1.57079632679489662f - a;
    if (x < 0.0f) a = 3.14159265358979324f - a;
    if (y < 0.0f) a = -a;
    return a;
}

// Natural log for finite x > 0: x = m * 2^e with m in [sqrt(1/2), sqrt(2)),
// log(m) = 2 atanh(f), f = (m-1)/(m+1), series through f^11.
ZGL_HD float logF32(float x)
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
ZGL_HD float expF32(float x)
{
    if (__builtin_isnan(x)) return x;
    if (x > 88.7228f) return __builtin_inff();
    if (x < -103.97f) return 0.0f;
    float kf = __builtin_floorf(x * 1.44269504088896341f + 0.5f);
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
ZGL_HD float powF32(float x, float y)
{
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : __builtin_inff();
    return expF32(y * logF32(x));
}

// ---------------------------------------------------------------------------
// FP64 transcendentals of the double instance, by the same methods as the FP32 ones
// with double constants and longer series, so that the GPU and a host build give the
// same bits (the libraries' double routines differ in the last ulp, and sampling steps
// such as sqrt(1 - x^2) near x = 1 amplify that).  Accuracy: a few ulp of double.
// sin/cos: pi/2 split into parts of 33, 33 and 53 bits (k*C1 and k*C2 exact for
// |k| < 2^20), Taylor series through r^21 / r^22 on |r| <= pi/4 (truncation < 1e-24).
// ---------------------------------------------------------------------------
ZGL_HD void sincosF64(double x, double* s, double* c)
{
    const double TwoOverPi = 0.63661977236758134;
    const double C1 = 1.5707963267341256;
    const double C2 = 6.077100506303966e-11;
    const double C3 = 2.0222662487959506e-21;
    const double kf = __builtin_floor(x * TwoOverPi + 0.5);
    const long long k = (long long)kf;
    const double r = ((x - kf * C1) - kf * C2) - kf * C3;
    const double r2 = r * r;
    double ps = 1.0 / 51090942171709440000.0;  // 1/21!
    ps = -1.0 / 121645100408832000.0 + r2 * ps;
    ps = 1.0 / 355687428096000.0 + r2 * ps;
    ps = -1.0 / 1307674368000.0 + r2 * ps;
    ps = 1.0 / 6227020800.0 + r2 * ps;
    ps = -1.0 / 39916800.0 + r2 * ps;
    ps = 1.0 / 362880.0 + r2 * ps;
    ps = -1.0 / 5040.0 + r2 * ps;
    ps = 1.0 / 120.0 + r2 * ps;
    ps = -1.0 / 6.0 + r2 * ps;
    double pc = -1.0 / 1124000727777607680000.0;  // -1/22!
    pc = 1.0 / 2432902008176640000.0 + r2 * pc;
    pc = -1.0 / 6402373705728000.0 + r2 * pc;
    pc = 1.0 / 20922789888000.0 + r2 * pc;
    pc = -1.0 / 87178291200.0 + r2 * pc;
    pc = 1.0 / 479001600.0 + r2 * pc;
    pc = -1.0 / 3628800.0 + r2 * pc;
    pc = 1.0 / 40320.0 + r2 * pc;
    pc = -1.0 / 720.0 + r2 * pc;
    pc = 1.0 / 24.0 + r2 * pc;
    const double sr = r + r * r2 * ps;
    const double cr = (1.0 - 0.5 * r2) + r2 * r2 * pc;
    switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

// atan on [0, 1]: t > tan(pi/8) reduced by atan(t) = pi/4 + atan((t-1)/(t+1)), then
// the series through t^43 on |t| <= tan(pi/8) (truncation < 1e-17).
ZGL_HD double atanUnitF64(double t)
{
    double off = 0.0;
    if (t > 0.41421356237309505) { t = (t - 1.0) / (t + 1.0); off = 0.78539816339744831; }
    const double t2 = t * t;
    double p = -1.0 / 43.0;
    for (int n = 41; n >= 3; n -= 2) p = ((n & 2) ? -1.0 : 1.0) / (double)n + t2 * p;
    return off + (t + t * t2 * p);
}

// GLSL atan(y, x); (0, 0) returns 0.
ZGL_HD double atan2F64(double y, double x)
{
    const double ax = __builtin_fabs(x), ay = __builtin_fabs(y);
    const double mx = (ax < ay) ? ay : ax, mn = (ay < ax) ? ay : ax;
    if (mx == 0.0) return 0.0;
    double a = atanUnitF64(mn / mx);
    if (ay > ax) a = 1.5707963267948966 - a;
    if (x < 0.0) a = 3.1415926535897932 - a;
    if (y < 0.0) a = -a;
    return a;
}

// Natural log for finite x > 0: x = m * 2^e, m in [sqrt(1/2), sqrt(2)), log(m) =
// 2 atanh(f), f = (m-1)/(m+1), series through f^23; ln 2 in parts of 32 and 53 bits.
ZGL_HD double logF64(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? -__builtin_inf() : __builtin_nan("");
    if (x == __builtin_inf()) return x;
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {  // subnormal: scale up by 2^54
        x = x * 18014398509481984.0;
        b = __builtin_bit_cast(uint64_t, x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    const double f = (m - 1.0) / (m + 1.0);
    const double f2 = f * f;
    double p = 1.0 / 23.0;
    for (int n = 21; n >= 3; n -= 2) p = 1.0 / (double)n + f2 * p;
    const double lm = 2.0 * f + 2.0 * f * f2 * p;
    const double ef = (double)e;
    return ef * 0.69314718036912382 + (lm + ef * 1.9082149292705877e-10);
}

// exp: x = k ln2 + r, |r| <= ln2/2, Taylor through r^17, then scale by 2^k.
ZGL_HD double expF64(double x)
{
    if (__builtin_isnan(x)) return x;
    if (x > 709.78) return __builtin_inf();
    if (x < -745.2) return 0.0;
    const double kf = __builtin_floor(x * 1.4426950408889634 + 0.5);
    const int k = (int)kf;
    const double r = (x - kf * 0.69314718036912382) - kf * 1.9082149292705877e-10;
    double p = 1.0 / 355687428096000.0;  // 1/17!
    p = 1.0 / 20922789888000.0 + r * p;
    p = 1.0 / 1307674368000.0 + r * p;
    p = 1.0 / 87178291200.0 + r * p;
    p = 1.0 / 6227020800.0 + r * p;
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 1.0 / 2.0 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    // 2^k in two steps so that k down to -1074 stays representable
    const int k1 = k / 2, k2 = k - k1;
    return p * __builtin_bit_cast(double, (uint64_t)(k1 + 1023) << 52) *
           __builtin_bit_cast(double, (uint64_t)(k2 + 1023) << 52);
}

// pow for x > 0 (the only use is the GTR1 sample with alpha in (0, 1]).
ZGL_HD double powF64(double x, double y)
{
    if (x == 1.0 || y == 0.0) return 1.0;
    if (x == 0.0) return (y > 0.0) ? 0.0 : __builtin_inf();
    return expF64(y * logF64(x));
}

// Overloads the working-precision math of zgl_math_body.inc dispatches on.
ZGL_HD float absImpl(float x) { return __builtin_fabsf(x); }
ZGL_HD double absImpl(double x) { return __builtin_fabs(x); }
ZGL_HD float sqrtImpl(float x) { return __builtin_sqrtf(x); }
ZGL_HD double sqrtImpl(double x) { return __builtin_sqrt(x); }
ZGL_HD float floorImpl(float x) { return __builtin_floorf(x); }
ZGL_HD double floorImpl(double x) { return __builtin_floor(x); }
// IEEE minNum / maxNum (one instruction each), for operands known to hold no NaN
ZGL_HD float ieeeMin(float x, float y) { return __builtin_fminf(x, y); }
ZGL_HD double ieeeMin(double x, double y) { return __builtin_fmin(x, y); }
ZGL_HD float ieeeMax(float x, float y) { return __builtin_fmaxf(x, y); }
ZGL_HD double ieeeMax(double x, double y) { return __builtin_fmax(x, y); }
ZGL_HD void sincosImpl(float x, float* s, float* c) { sincosF32(x, s, c); }
ZGL_HD void sincosImpl(double x, double* s, double* c) { sincosF64(x, s, c); }
ZGL_HD float atan2Impl(float y, float x) { return atan2F32(y, x); }
ZGL_HD double atan2Impl(double y, double x) { return atan2F64(y, x); }
ZGL_HD float logImpl(float x) { return logF32(x); }
ZGL_HD double logImpl(double x) { return logF64(x); }
ZGL_HD float expImpl(float x) { return expF32(x); }
ZGL_HD double expImpl(double x) { return expF64(x); }
ZGL_HD float powImpl(float x, float y) { return powF32(x, y); }
ZGL_HD double powImpl(double x, double y) { return powF64(x, y); }

// Small vectors and the 3x3 matrix, over the element type.
template <class T> struct TVec2 { T x, y; };
template <class T> struct TVec3 { T x, y, z; };
template <class T> struct TVec4 { T x, y, z, w; };
// column-major: M * v = c0 v.x + c1 v.y + c2 v.z
template <class T> struct TMat3 { TVec3<T> c0, c1, c2; };

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

// The rendering path's instance of the vector math, in its own namespace so that the
// double instance (zgl::f64) never finds these functions by argument-dependent lookup on
// the shared types; the using-directive makes them zgl's for the kernels.
namespace f32 {
typedef float real;
#include "zgl_math_body.inc"
}  // namespace f32
using namespace f32;

}  // namespace zgl
