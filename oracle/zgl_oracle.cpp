// zgl_oracle.cpp -- CPU restatement of ZillumGL's rendering path.  TEST INFRASTRUCTURE.
//
// This file is the oracle: only the tests, the smoke check and the CPU-baseline leg
// of the benchmark load it, as the checker.  The product (libzgl.so) never links
// or calls it.
//
// It restates, function by function, the reference's GLSL kernels and host
// producers, each with the reference file:line it follows, and runs them on the
// reference's own buffer layout (12-byte hit-table rows + separate bounds, index
// then vertex gathers, RGB16F env texels), so it shares no data structure with
// the GPU code.  It is templated on the arithmetic type:
//   R = float   matched-sample parity: every expression is evaluated in the
//               GLSL source order, compiled with -ffp-contract=off;
//   R = double  the converged reference used for RMSE.
// Conventions for GLSL builtins whose bit-level definition is driver dependent
// (shared with the GPU code, which documents the same choices):
//   min(x,y) = y < x ? y : x;  max(x,y) = x < y ? y : x;  normalize(v) = v * (1/|v|);
//   mix(x,y,a) = x*(1-a) + y*a;  dot = ((x*x') + y*y') + z*z';  mat3 * v = ((c0 v.x +
//   c1 v.y) + c2 v.z);  inverse(mat3) by cofactors over the determinant;  bilinear
//   GL_LINEAR/GL_REPEAT filtering in FP32 with weights (1-a)(1-b), a(1-b), (1-a)b, ab
//   summed in that order;  sin/cos/atan2/log/exp for R = float by Cody-Waite
//   reduction + truncated Taylor series (identical steps to the GPU code; pinned
//   against libm double in the tests), and for R = double by the same methods with
//   longer series (identical steps to the GPU code's FP64 instance).
#include <algorithm>
#include <array>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// FP32 transcendentals (Cody-Waite + Taylor), and the same methods in FP64.
// ---------------------------------------------------------------------------
static inline float bitsToF(uint32_t b) { float f; std::memcpy(&f, &b, 4); return f; }
static inline uint32_t fToBits(float f) { uint32_t b; std::memcpy(&b, &f, 4); return b; }

static void sincosF(float x, float* s, float* c)
{
    const float TwoOverPi = 0.636619772367581343f;
    const float C1 = 1.5703125f, C2 = 4.837512969970703125e-4f, C3 = 7.54978995489188216e-8f;
    float kf = std::floor(x * TwoOverPi + 0.5f);
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
static float atanUnitF(float t)
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
static float atan2F(float y, float x)
{
    float ax = std::fabs(x), ay = std::fabs(y);
    float mx = (ax < ay) ? ay : ax, mn = (ay < ax) ? ay : ax;
    if (mx == 0.0f) return 0.0f;
    float a = atanUnitF(mn / mx);
    if (ay > ax) a = 1.57079632679489662f - a;
    if (x < 0.0f) a = 3.14159265358979324f - a;
    if (y < 0.0f) a = -a;
    return a;
}
static float logF(float x)
{
    if (!(x > 0.0f)) return (x == 0.0f) ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    uint32_t b = fToBits(x);
    int e = (int)((b >> 23) & 0xff);
    if (e == 0) { x = x * 8388608.0f; b = fToBits(x); e = (int)((b >> 23) & 0xff) - 23; }
    e -= 127;
    float m = bitsToF((b & 0x007fffffu) | 0x3f800000u);
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
static float expF(float x)
{
    if (std::isnan(x)) return x;
    if (x > 88.7228f) return INFINITY;
    if (x < -103.97f) return 0.0f;
    float kf = std::floor(x * 1.44269504088896341f + 0.5f);
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
    int k1 = k / 2, k2 = k - k1;
    return p * bitsToF((uint32_t)(k1 + 127) << 23) * bitsToF((uint32_t)(k2 + 127) << 23);
}
static float powF(float x, float y)
{
    if (x == 1.0f || y == 0.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : INFINITY;
    return expF(y * logF(x));
}

// The same methods in double (the FP64 mode): pi/2 in parts of 33, 33 and 53 bits,
// sin/cos through r^21 / r^22, atan through t^43, log's atanh series through f^23 with
// ln 2 in parts of 32 and 53 bits, exp through r^17 (identical steps to the GPU code's
// FP64 instance; pinned against libm double in the tests).
static void sincosD(double x, double* s, double* c)
{
    const double TwoOverPi = 0.63661977236758134;
    const double C1 = 1.5707963267341256, C2 = 6.077100506303966e-11, C3 = 2.0222662487959506e-21;
    double kf = std::floor(x * TwoOverPi + 0.5);
    long long k = (long long)kf;
    double r = ((x - kf * C1) - kf * C2) - kf * C3;
    double r2 = r * r;
    double ps = 1.0 / 51090942171709440000.0;
    ps = -1.0 / 121645100408832000.0 + r2 * ps;
    ps = 1.0 / 355687428096000.0 + r2 * ps;
    ps = -1.0 / 1307674368000.0 + r2 * ps;
    ps = 1.0 / 6227020800.0 + r2 * ps;
    ps = -1.0 / 39916800.0 + r2 * ps;
    ps = 1.0 / 362880.0 + r2 * ps;
    ps = -1.0 / 5040.0 + r2 * ps;
    ps = 1.0 / 120.0 + r2 * ps;
    ps = -1.0 / 6.0 + r2 * ps;
    double pc = -1.0 / 1124000727777607680000.0;
    pc = 1.0 / 2432902008176640000.0 + r2 * pc;
    pc = -1.0 / 6402373705728000.0 + r2 * pc;
    pc = 1.0 / 20922789888000.0 + r2 * pc;
    pc = -1.0 / 87178291200.0 + r2 * pc;
    pc = 1.0 / 479001600.0 + r2 * pc;
    pc = -1.0 / 3628800.0 + r2 * pc;
    pc = 1.0 / 40320.0 + r2 * pc;
    pc = -1.0 / 720.0 + r2 * pc;
    pc = 1.0 / 24.0 + r2 * pc;
    double sr = r + r * r2 * ps;
    double cr = (1.0 - 0.5 * r2) + r2 * r2 * pc;
    switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}
static double atanUnitD(double t)
{
    double off = 0.0;
    if (t > 0.41421356237309505) { t = (t - 1.0) / (t + 1.0); off = 0.78539816339744831; }
    double t2 = t * t;
    double p = -1.0 / 43.0;
    p = 1.0 / 41.0 + t2 * p;
    p = -1.0 / 39.0 + t2 * p;
    p = 1.0 / 37.0 + t2 * p;
    p = -1.0 / 35.0 + t2 * p;
    p = 1.0 / 33.0 + t2 * p;
    p = -1.0 / 31.0 + t2 * p;
    p = 1.0 / 29.0 + t2 * p;
    p = -1.0 / 27.0 + t2 * p;
    p = 1.0 / 25.0 + t2 * p;
    p = -1.0 / 23.0 + t2 * p;
    p = 1.0 / 21.0 + t2 * p;
    p = -1.0 / 19.0 + t2 * p;
    p = 1.0 / 17.0 + t2 * p;
    p = -1.0 / 15.0 + t2 * p;
    p = 1.0 / 13.0 + t2 * p;
    p = -1.0 / 11.0 + t2 * p;
    p = 1.0 / 9.0 + t2 * p;
    p = -1.0 / 7.0 + t2 * p;
    p = 1.0 / 5.0 + t2 * p;
    p = -1.0 / 3.0 + t2 * p;
    return off + (t + t * t2 * p);
}
static double atan2D(double y, double x)
{
    double ax = std::fabs(x), ay = std::fabs(y);
    double mx = (ax < ay) ? ay : ax, mn = (ay < ax) ? ay : ax;
    if (mx == 0.0) return 0.0;
    double a = atanUnitD(mn / mx);
    if (ay > ax) a = 1.5707963267948966 - a;
    if (x < 0.0) a = 3.1415926535897932 - a;
    if (y < 0.0) a = -a;
    return a;
}
static double logD(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    uint64_t b; std::memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) { x = x * 18014398509481984.0; std::memcpy(&b, &x, 8); e = (int)((b >> 52) & 0x7ff) - 54; }
    e -= 1023;
    uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m; std::memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double f = (m - 1.0) / (m + 1.0);
    double f2 = f * f;
    double p = 1.0 / 23.0;
    p = 1.0 / 21.0 + f2 * p;
    p = 1.0 / 19.0 + f2 * p;
    p = 1.0 / 17.0 + f2 * p;
    p = 1.0 / 15.0 + f2 * p;
    p = 1.0 / 13.0 + f2 * p;
    p = 1.0 / 11.0 + f2 * p;
    p = 1.0 / 9.0 + f2 * p;
    p = 1.0 / 7.0 + f2 * p;
    p = 1.0 / 5.0 + f2 * p;
    p = 1.0 / 3.0 + f2 * p;
    double lm = 2.0 * f + 2.0 * f * f2 * p;
    double ef = (double)e;
    return ef * 0.69314718036912382 + (lm + ef * 1.9082149292705877e-10);
}
static double expD(double x)
{
    if (std::isnan(x)) return x;
    if (x > 709.78) return INFINITY;
    if (x < -745.2) return 0.0;
    double kf = std::floor(x * 1.4426950408889634 + 0.5);
    int k = (int)kf;
    double r = (x - kf * 0.69314718036912382) - kf * 1.9082149292705877e-10;
    double p = 1.0 / 355687428096000.0;
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
    int k1 = k / 2, k2 = k - k1;
    uint64_t b1 = (uint64_t)(k1 + 1023) << 52, b2 = (uint64_t)(k2 + 1023) << 52;
    double s1, s2; std::memcpy(&s1, &b1, 8); std::memcpy(&s2, &b2, 8);
    return p * s1 * s2;
}
static double powD(double x, double y)
{
    if (x == 1.0 || y == 0.0) return 1.0;
    if (x == 0.0) return (y > 0.0) ? 0.0 : INFINITY;
    return expD(y * logD(x));
}

template <class R> struct M;
template <> struct M<float> {
    static void sincos(float x, float* s, float* c) { sincosF(x, s, c); }
    static float atan2(float y, float x) { return atan2F(y, x); }
    static float log(float x) { return logF(x); }
    static float pow(float x, float y) { return powF(x, y); }
    static constexpr float Pi = 3.14159265358979323846f;
    static constexpr float PiInv = 1.0f / 3.14159265358979323846f;
};
template <> struct M<double> {
    static void sincos(double x, double* s, double* c) { sincosD(x, s, c); }
    static double atan2(double y, double x) { return atan2D(y, x); }
    static double log(double x) { return logD(x); }
    static double pow(double x, double y) { return powD(x, y); }
    static constexpr double Pi = 3.14159265358979323846;
    static constexpr double PiInv = 1.0 / 3.14159265358979323846;
};

// ---------------------------------------------------------------------------
// GLSL-like vectors
// ---------------------------------------------------------------------------
template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R s) : x(s), y(s), z(s) {}
};
template <class R> struct V2 {
    R x, y;
    V2() : x(0), y(0) {}
    V2(R a, R b) : x(a), y(b) {}
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return V3<R>(a.x / s, a.y / s, a.z / s); }
template <class R> inline V3<R> operator-(V3<R> a) { return V3<R>(-a.x, -a.y, -a.z); }
template <class R> inline R gmin(R x, R y) { return (y < x) ? y : x; }
template <class R> inline R gmax(R x, R y) { return (x < y) ? y : x; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b)
{
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> normalize(V3<R> a) { R inv = R(1) / std::sqrt(dot(a, a)); return a * inv; }
template <class R> inline V3<R> mix(V3<R> x, V3<R> y, R a) { return x * (R(1) - a) + y * a; }
template <class R> inline R mixs(R x, R y, R a) { return x * (R(1) - a) + y * a; }
template <class R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - n * (R(2) * dot(n, i)); }
template <class R> inline R satDot(V3<R> a, V3<R> b) { return gmax(dot(a, b), R(0)); }
template <class R> inline R absDot(V3<R> a, V3<R> b) { return std::fabs(dot(a, b)); }
template <class R> inline R luminance(V3<R> c) { return dot(c, V3<R>(R(0.299f), R(0.587f), R(0.114f))); }
template <class R> inline bool isBlack(V3<R> c) { return luminance(c) < R(1e-5f); }
template <class R> inline bool hasNan(V3<R> c) { return std::isnan(c.x) || std::isnan(c.y) || std::isnan(c.z); }
template <class R> inline bool sameHemisphere(V3<R> n, V3<R> a, V3<R> b) { return dot(n, a) * dot(n, b) > R(0); }
template <class R> inline R sq(R x) { return x * x; }
template <class R> inline R pow5(R x) { R x2 = x * x; return x2 * x2 * x; }
template <class R> inline R maxComp(V3<R> v) { return gmax(v.x, gmax(v.y, v.z)); }

template <class R> struct M3 { V3<R> c0, c1, c2; };
template <class R> inline V3<R> mul(const M3<R>& m, V3<R> v) { return (m.c0 * v.x + m.c1 * v.y) + m.c2 * v.z; }
template <class R> M3<R> inverse3(const M3<R>& m)
{
    R a00 = m.c0.x, a01 = m.c0.y, a02 = m.c0.z, a10 = m.c1.x, a11 = m.c1.y, a12 = m.c1.z;
    R a20 = m.c2.x, a21 = m.c2.y, a22 = m.c2.z;
    R b01 = a22 * a11 - a12 * a21, b11 = -a22 * a10 + a12 * a20, b21 = a21 * a10 - a11 * a20;
    R det = a00 * b01 + a01 * b11 + a02 * b21;
    R id = R(1) / det;
    M3<R> r;
    r.c0 = V3<R>(b01 * id, (-a22 * a01 + a02 * a21) * id, (a12 * a01 - a02 * a11) * id);
    r.c1 = V3<R>(b11 * id, (a22 * a00 - a02 * a20) * id, (-a12 * a00 + a02 * a10) * id);
    r.c2 = V3<R>(b21 * id, (-a21 * a00 + a01 * a20) * id, (a11 * a00 - a01 * a10) * id);
    return r;
}

// math.glsl:73-89
template <class R> M3<R> tbn(V3<R> n)
{
    V3<R> t = (std::fabs(n.z) > R(0.999f)) ? V3<R>(0, 1, 0) : V3<R>(0, 0, 1);
    V3<R> b = normalize(cross(n, t));
    t = cross(b, n);
    return M3<R>{t, b, n};
}
template <class R> V3<R> normalToWorld(V3<R> n, V3<R> v) { return normalize(mul(tbn(n), v)); }
// math.glsl:25-41
template <class R> V2<R> concentric(V2<R> v)
{
    if (v.x == R(0) && v.y == R(0)) return V2<R>(0, 0);
    v = V2<R>(v.x * R(2) - R(1), v.y * R(2) - R(1));
    const R Pi = M<R>::Pi;
    R phi, r;
    if (v.x * v.x > v.y * v.y) { r = v.x; phi = Pi * v.y / v.x * R(0.25f); }
    else { r = v.y; phi = Pi * R(0.5f) - Pi * v.x / v.y * R(0.25f); }
    R s, c;
    M<R>::sincos(phi, &s, &c);
    return V2<R>(r * c, r * s);
}
// math.glsl:99-105; returns dir and pdf
template <class R> V3<R> cosineSample(V3<R> n, V2<R> u, R* pdf)
{
    V2<R> d = concentric(u);
    R z = std::sqrt(R(1) - (d.x * d.x + d.y * d.y));
    *pdf = M<R>::PiInv * z;
    return normalToWorld(n, V3<R>(d.x, d.y, z));
}
// math.glsl:58-71
template <class R> V2<R> sphereToPlane(V3<R> w)
{
    const R Pi = M<R>::Pi, PiInv = M<R>::PiInv;
    R theta = M<R>::atan2(w.y, w.x);
    if (theta < R(0)) theta += Pi * R(2);
    R phi = M<R>::atan2(std::sqrt(w.x * w.x + w.y * w.y), w.z);
    return V2<R>(theta * PiInv * R(0.5f), phi * PiInv);
}
template <class R> V3<R> planeToSphere(V2<R> uv)
{
    const R Pi = M<R>::Pi;
    R theta = uv.x * Pi * R(2), phi = uv.y * Pi;
    R st, ct, sp, cp;
    M<R>::sincos(theta, &st, &ct);
    M<R>::sincos(phi, &sp, &cp);
    return V3<R>(ct * sp, st * sp, cp);
}
template <class R> V3<R> rotZ(V3<R> v, R c, R s) { return V3<R>(v.x * c - v.y * s, v.x * s + v.y * c, v.z); }
// math.glsl:112-131
template <class R> int cubemapFace(V3<R> d)
{
    R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int m = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    if (m == 0) return d.x > R(0) ? 0 : 1;
    if (m == 1) return d.y > R(0) ? 2 : 3;
    return d.z > R(0) ? 4 : 5;
}
// math.glsl:139-150
template <class R> V3<R> triUniform(V3<R> a, V3<R> b, V3<R> c, V2<R> uv)
{
    R r = std::sqrt(uv.y);
    R u = R(1) - r, v = uv.x * r;
    return (a * (R(1) - u - v) + b * u) + c * v;
}
template <class R> R triArea(V3<R> a, V3<R> b, V3<R> c) { return R(0.5f) * length(cross(c - a, b - a)); }

// random.glsl:5-13
static inline uint32_t hashU(uint32_t s)
{
    s = (s ^ 61u) ^ (s >> 16u);
    s *= 9u;
    s = s ^ (s >> 4u);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15u);
    return s;
}

// ---------------------------------------------------------------------------
// Sobol matrices from the Joe-Kuo parameters (src/math/SobolMatrices256x32.h data,
// src/core/Sampler.cpp:19-28 for the XOR construction).
// ---------------------------------------------------------------------------
struct JK { int s; int a; int m[16]; };
static const JK kJK[255] = {
#include "../zillumgl_amd/csrc/sobol_joe_kuo.inc"
};
struct Sobol {
    uint32_t col[256][32];
    Sobol()
    {
        for (int k = 0; k < 32; k++) col[0][k] = 1u << (31 - k);
        for (int d = 1; d < 256; d++) {
            const JK& p = kJK[d - 1];
            for (int k = 0; k < 32; k++) {
                if (k < p.s) { col[d][k] = (uint32_t)p.m[k] << (31 - k); continue; }
                uint32_t v = col[d][k - p.s] ^ (col[d][k - p.s] >> p.s);
                for (int j = 1; j < p.s; j++)
                    if ((p.a >> (p.s - 1 - j)) & 1) v ^= col[d][k - j];
                col[d][k] = v;
            }
        }
    }
    uint32_t sample(uint32_t index, int dim) const
    {
        uint32_t r = 0;
        for (int i = 0; index != 0; index >>= 1, i++)
            if (index & 1) r ^= col[dim][i];
        return r;
    }
};
static const Sobol& sobol() { static Sobol s; return s; }

// ---------------------------------------------------------------------------
// Host producers (restated independently of the library's builder)
// ---------------------------------------------------------------------------
struct AB { float lo[3], hi[3]; };
static AB abEmpty() { AB b; for (int i = 0; i < 3; i++) { b.lo[i] = 1e8f; b.hi[i] = -1e8f; } return b; }
static AB abUnion(const AB& a, const AB& b)
{
    AB r;
    for (int i = 0; i < 3; i++) { r.lo[i] = gmin(a.lo[i], b.lo[i]); r.hi[i] = gmax(a.hi[i], b.hi[i]); }
    return r;
}
static AB abPoint(const float* p) { AB b; for (int i = 0; i < 3; i++) b.lo[i] = b.hi[i] = p[i]; return b; }
static float abSA(const AB& b)
{
    float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return 2.0f * (x * y + y * z + z * x);
}
static int abMaxExt(const AB& b)
{
    float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    if (x > y) return x > z ? 0 : 2;
    return y > z ? 1 : 2;
}
static void abCent(const AB& b, float* c) { for (int i = 0; i < 3; i++) c[i] = (b.lo[i] + b.hi[i]) * 0.5f; }
// (int) of a float as x86-64 cvttss2si: NaN / out of range -> INT_MIN
static int cvtt(float f) { return (f > -2147483904.0f && f < 2147483648.0f) ? (int)f : INT_MIN; }

struct PI { AB b; float c[3]; int idx; };

// BVH::quickBuild (src/accelerator/BVH.cpp:217-296) + partition (:97-114) +
// buildHitTable (:298-345)
static void buildBvh(const float* V, const uint32_t* I, int64_t n, std::vector<float>& boundsOut,
                     std::vector<int32_t>& table)
{
    const int64_t size = 2 * n - 1;
    std::vector<PI> p(n);
    AB root = abEmpty();
    for (int64_t i = 0; i < n; i++) {
        const float *a = V + 3 * I[3 * i], *b = V + 3 * I[3 * i + 1], *c = V + 3 * I[3 * i + 2];
        for (int k = 0; k < 3; k++) {
            p[i].b.lo[k] = gmin(gmin(a[k], b[k]), c[k]);
            p[i].b.hi[k] = gmax(gmax(a[k], b[k]), c[k]);
        }
        abCent(p[i].b, p[i].c);
        p[i].idx = (int)i;
        root = abUnion(root, abPoint(p[i].c));
    }
    std::vector<AB> bounds(size);
    std::vector<int> sizes(size);
    struct Rec { int off; AB ext; int dim, l, r; };
    std::vector<Rec> st;
    st.push_back({0, root, abMaxExt(root), 0, (int)n - 1});
    while (!st.empty()) {
        Rec t = st.back();
        st.pop_back();
        int nodeSize = (t.r - t.l) * 2 + 1;
        sizes[t.off] = (nodeSize == 1) ? (int)((uint32_t)p[t.l].idx | 0x80000000u) : nodeSize;
        if (t.l == t.r) { bounds[t.off] = p[t.l].b; continue; }
        int cnt = t.r - t.l + 1;
        if (cnt == 2) {
            bounds[t.off] = abUnion(p[t.l].b, p[t.r].b);
            if (p[t.l].c[t.dim] > p[t.r].c[t.dim]) std::swap(p[t.l], p[t.r]);
            st.push_back({t.off + 2, abPoint(p[t.r].c), 0, t.r, t.r});
            st.push_back({t.off + 1, abPoint(p[t.l].c), 0, t.l, t.l});
            continue;
        }
        float mn = t.ext.lo[t.dim], mx = t.ext.hi[t.dim];
        auto bucket = [&](float c) {
            int b = cvtt(16.0f * (c - mn) / (mx - mn));
            return std::max(std::min(b, 15), 0);
        };
        int cntB[16] = {0};
        AB boxB[16];
        for (int i = 0; i < 16; i++) boxB[i] = abEmpty();
        for (int i = t.l; i <= t.r; i++) { int b = bucket(p[i].c[t.dim]); cntB[b]++; boxB[b] = abUnion(boxB[b], p[i].b); }
        int pc[16], sc[16];
        AB pb[16], sb[16];
        pc[0] = cntB[0]; pb[0] = boxB[0]; sc[15] = cntB[15]; sb[15] = boxB[15];
        for (int i = 1; i < 16; i++) {
            pc[i] = pc[i - 1] + cntB[i]; pb[i] = abUnion(pb[i - 1], boxB[i]);
            sc[15 - i] = sc[16 - i] + cntB[15 - i]; sb[15 - i] = abUnion(sb[16 - i], boxB[15 - i]);
        }
        bounds[t.off] = pb[15];
        int split = 0;
        float best = (float)pc[0] * abSA(pb[0]) + (float)sc[1] * abSA(sb[1]);
        for (int i = 1; i < 15; i++) {
            float c = (float)pc[i] * abSA(pb[i]) + (float)sc[i + 1] * abSA(sb[i + 1]);
            if (c < best) { best = c; split = i; }
        }
        std::vector<PI> tmp(p.begin() + t.l, p.begin() + t.r + 1);
        int lo = 0, hi = cnt;
        for (int i = 0; i < cnt; i++) {
            if (bucket(tmp[i].c[t.dim]) <= split) p[t.l + lo++] = tmp[i];
            else p[t.l + --hi] = tmp[i];
        }
        if (hi == cnt) hi--;
        int mid = t.l + hi - 1;
        AB lc = abEmpty(), rc = abEmpty();
        for (int i = t.l; i <= mid; i++) lc = abUnion(lc, abPoint(p[i].c));
        for (int i = mid + 1; i <= t.r; i++) rc = abUnion(rc, abPoint(p[i].c));
        st.push_back({t.off + 2 * (mid - t.l) + 2, rc, abMaxExt(rc), mid + 1, t.r});
        st.push_back({t.off + 1, lc, abMaxExt(lc), t.l, mid});
    }
    boundsOut.resize(size * 6);
    for (int64_t i = 0; i < size; i++)
        for (int k = 0; k < 3; k++) { boundsOut[i * 6 + k] = bounds[i].lo[k]; boundsOut[i * 6 + 3 + k] = bounds[i].hi[k]; }
    table.assign(size * 18, 0);
    for (int o = 0; o < 6; o++) {
        int axis = o / 2;
        std::vector<int64_t> stack;
        stack.push_back(0);
        int64_t index = 0;
        int32_t* T = table.data() + size * 3 * o;
        while (!stack.empty()) {
            int64_t k = stack.back();
            stack.pop_back();
            bool leaf = (sizes[k] & 0x80000000) != 0;
            int ns = leaf ? 1 : sizes[k];
            T[index * 3] = (int32_t)k;
            T[index * 3 + 1] = leaf ? (int32_t)((uint32_t)sizes[k] ^ 0x80000000u) : -1;
            T[index * 3 + 2] = (int32_t)(index + ns);
            index++;
            if (leaf) continue;
            int ls = sizes[k + 1];
            if (ls & 0x80000000) ls = 1;
            int64_t L = k + 1, Rr = k + 1 + ls;
            float cl[3], cr[3];
            abCent(bounds[L], cl);
            abCent(bounds[Rr], cr);
            bool keep = (o % 2 == 0) ? (cl[axis] > cr[axis]) : (cl[axis] < cr[axis]);
            if (!keep) std::swap(L, Rr);
            stack.push_back(Rr);
            stack.push_back(L);
        }
    }
}

// AliasTable::build (src/math/AliasTable.h:11-57) with leftover probabilities kept
// (keepLeftover) or set to 1 (EnvironmentMap::setupAliasTable, :66-114).
static float aliasCore(float* pdf, int32_t* alias, int n, int stride, bool keepLeftover)
{
    float sum = 0.0f;
    for (int i = 0; i < n; i++) sum += pdf[(size_t)i * stride];
    float inv = (float)n / sum;
    std::vector<std::pair<int, float>> G, L;
    for (int i = 0; i < n; i++) {
        float& v = pdf[(size_t)i * stride];
        v *= inv;
        (v < 1.0f ? L : G).push_back({i, v});
    }
    while (!G.empty() && !L.empty()) {
        auto l = L.back(); L.pop_back();
        auto g = G.back(); G.pop_back();
        alias[(size_t)l.first * stride] = g.first;
        pdf[(size_t)l.first * stride] = l.second;
        g.second += l.second - 1.0f;
        (g.second < 1.0f ? L : G).push_back(g);
    }
    while (!G.empty()) { auto g = G.back(); G.pop_back(); alias[(size_t)g.first * stride] = g.first; pdf[(size_t)g.first * stride] = keepLeftover ? g.second : 1.0f; }
    while (!L.empty()) { auto l = L.back(); L.pop_back(); alias[(size_t)l.first * stride] = l.first; pdf[(size_t)l.first * stride] = keepLeftover ? l.second : 1.0f; }
    return sum;
}

// ---------------------------------------------------------------------------
// Scene view in the reference layout
// ---------------------------------------------------------------------------
}  // namespace orc

extern "C" {
typedef struct orc_scene {
    const float* vertices; const float* normals; int64_t n_vertices;
    const float* texcoords; int64_t n_texcoords;
    const int32_t* indices; int64_t n_triangles;
    const float* bounds; const int32_t* hit_table; int32_t bvh_size;
    const int32_t* mat_tex_indices; int32_t obj_prim_count;
    const float* materials; int32_t n_materials;
    const float* light_power; const int32_t* light_alias; const float* light_prob;
    int32_t n_light_triangles; float light_sum;
    const uint16_t* env_rgb16f; int32_t env_width, env_height;
    const int32_t* env_alias; const float* env_prob; float env_sum; float env_rotation;
    int32_t sampler;
    float cam_front[3], cam_right[3], cam_up[3], cam_pos[3], cam_mat_inv[9];
    float tan_fov, aspect, lens_radius, focal_dist;
    int32_t film_width, film_height;
    const float* noise; int32_t noise_width, noise_height;
    // albedo textures: tex_layers sRGB8 RGB layers of tex_width x tex_height, uv scales
    const uint8_t* tex_rgb8; int32_t tex_layers, tex_width, tex_height; const float* tex_scale;
} orc_scene;

typedef struct orc_params {
    int32_t integrator, first_pass, n_passes, free_counter, max_depth, russian_roulette, sample_light;
    int32_t light_env_uniform; float light_portion; int32_t blocks_one_pass, loops_per_pass;
} orc_params;
}

namespace orc {

static float halfToF(uint16_t h)
{
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1f, m = h & 0x3ffu, x;
    if (e == 0) {
        if (m == 0) x = sign;
        else { int ee = 113; while (!(m & 0x400u)) { m <<= 1; ee--; } x = sign | ((uint32_t)ee << 23) | ((m & 0x3ffu) << 13); }
    } else if (e == 31) x = sign | 0x7f800000u | (m << 13);
    else x = sign | ((e + 112) << 23) | (m << 13);
    return bitsToF(x);
}

struct Cnt { uint64_t bytes = 0, rows = 0, tests = 0, trav = 0; };

// Estimator variants, for the tests that pin the restatement (0 = the reference's
// estimators as written; set by orc_set_variant).
//   kVarEmitCos    the triple tracer's camera pass weighs its strategies with the light's
//                  actual emission pdf, cos/pi (lightSampleOneLe samples cosine-weighted
//                  directions, light.glsl:111-120), instead of lightPdfLe's uniform
//                  1/(2 pi) (light.glsl:101-109; triple_path_pass_pt.glsl:119, :162);
//   kVarPlantPrev  a planted mistake: the s=1 weight drops its reverse-pdf factor
//                  coefToPrev (triple_path_pass_pt.glsl:121);
//   kVarPlantEnvSin a planted mistake: envSampleWi's pdf gets an extra sin(theta).
constexpr int kVarEmitCos = 1, kVarPlantPrev = 2, kVarPlantEnvSin = 4;
static int gVariant = 0;

// Records of one sample of the differentiable naive estimator (replay with detached
// sampling), kept by the oracle's FP64 adjoint.
struct RepTap {
    int idx[4];
    double w[4];
};
struct RepVertex {
    int mat;
    uint32_t type;
    bool diff;
    float wo[3], wi[3], ns[3];
    float pdf, cosf;  // BSDF sampling pdf (theta0) and the cosine factor (1 for delta)
    float fConst[3];  // bs.f, the factor of a non-differentiable material
    float fE[3];      // the FP32 BSDF value the estimator multiplied by
    float rr;         // Russian-roulette division applied after this vertex (1 if none)
};
struct RepTerm {
    int kind;  // 0 camera ray hits env, 1 NEE to a light, 2 NEE to env, 3 BSDF ray hits env, 4 ... hits a light
    int v;     // NEE: the vertex; hits: the vertex whose BSDF ray hit
    // NEE: the vertex's material and frame, the light direction and the BSDF value there
    int mat;
    uint32_t type;
    bool diff;
    float wo[3], ns[3], wl[3], f[3];
    float satDot, weight, envPdf, sel;
    float coef[3];  // NEE to a light: Le / pdf / sel (independent of the parameters)
    float le[3];    // light hit: Le
    RepTap tap;
};
struct RepSample {
    std::vector<RepVertex> vs;
    std::vector<RepTerm> ts;
};

template <class R> struct Ctx {
    using V = V3<R>;
    using W = V2<R>;
    const orc_scene& s;
    const orc_params& p;
    Cnt& cn;
    // rng
    uint32_t randSeed = 0, sampleSeed = 0;
    int sampleOffset = 0, dim = 0;
    int sampler;
    R rotC, rotS, rotCN, rotSN;

    Ctx(const orc_scene& sc, const orc_params& pp, Cnt& c, int samplerType) : s(sc), p(pp), cn(c), sampler(samplerType)
    {
        if constexpr (std::is_same<R, float>::value) {
            sincosF(sc.env_rotation, &rotS, &rotC);
            sincosF(-sc.env_rotation, &rotSN, &rotCN);
        } else {
            double a, b, c, d;
            sincosD((double)sc.env_rotation, &a, &b);
            sincosD(-(double)sc.env_rotation, &c, &d);
            rotS = a; rotC = b; rotSN = c; rotCN = d;
        }
    }

    static V f3(const float* a) { return V(R(a[0]), R(a[1]), R(a[2])); }
    V vert(int i) const { return f3(s.vertices + 3 * (size_t)i); }
    V nrm(int i) const { return f3(s.normals + 3 * (size_t)i); }
    W tex(int i) const
    {
        if (i < s.n_texcoords) return W(R(s.texcoords[2 * (size_t)i]), R(s.texcoords[2 * (size_t)i + 1]));
        return W(0, 0);
    }

    // random.glsl:15-19, 61-85
    R rand()
    {
        randSeed = hashU(randSeed);
        return R(float(randSeed) * (1.0f / 4294967296.0f));
    }
    R sample1D()
    {
        if (sampler == 0) return rand();
        int idx = (int)((uint32_t)sampleOffset + (uint32_t)dim);  // GLSL int addition wraps
        uint32_t r = 0;
        if (idx >= 0 && idx < 131072 * 256) r = sobol().sample((uint32_t)(idx >> 8), idx & 255);
        cn.bytes += 4;
        r ^= sampleSeed;
        sampleSeed = hashU(sampleSeed);
        dim++;
        return R(float(r) * (1.0f / 4294967296.0f));
    }

    // intersection.glsl:63-109
    static bool triHit(V a, V b, V c, V o, V d, R& t, R* uo = nullptr, R* vo = nullptr, R* deto = nullptr)
    {
        const R eps = R(1e-6f);
        V ab = b - a, ac = c - a;
        V pv = cross(d, ac);
        R det = dot(ab, pv);
        if (std::fabs(det) < eps) return false;
        V ao = o - a;
        if (det < R(0)) { ao = -ao; det = -det; }
        R u = dot(ao, pv);
        if (u < R(0) || u > det) return false;
        V q = cross(ao, ab);
        R v = dot(d, q);
        if (v < R(0) || u + v > det) return false;
        t = dot(ac, q) / det;
        if (uo) { *uo = u; *vo = v; *deto = det; }
        return t > R(0);
    }
    bool triHitId(int id, V o, V d, R& t)
    {
        cn.bytes += 48;
        cn.tests++;
        int ia = s.indices[3 * id], ib = s.indices[3 * id + 1], ic = s.indices[3 * id + 2];
        return triHit(vert(ia), vert(ib), vert(ic), o, d, t);
    }
    // intersection.glsl:226-329
    bool boxHit(int node, V o, V d, R& tMin)
    {
        const R eps = R(1e-6f);
        V pMin = f3(s.bounds + 6 * (size_t)node), pMax = f3(s.bounds + 6 * (size_t)node + 3);
        R tMax;
        const R one = R(1) - eps;
        if (std::fabs(d.x) > one) {
            if (o.y > pMin.y && o.y < pMax.y && o.z > pMin.z && o.z < pMax.z) {
                R inv = R(1) / d.x, ta = (pMin.x - o.x) * inv, tb = (pMax.x - o.x) * inv;
                tMin = gmin(ta, tb); tMax = gmax(ta, tb);
                return tMax >= R(0) && tMax >= tMin;
            }
            return false;
        }
        if (std::fabs(d.y) > one) {
            if (o.x > pMin.x && o.x < pMax.x && o.z > pMin.z && o.z < pMax.z) {
                R inv = R(1) / d.y, ta = (pMin.y - o.y) * inv, tb = (pMax.y - o.y) * inv;
                tMin = gmin(ta, tb); tMax = gmax(ta, tb);
                return tMax >= R(0) && tMax >= tMin;
            }
            return false;
        }
        if (std::fabs(d.z) > one) {
            if (o.x > pMin.x && o.x < pMax.x && o.y > pMin.y && o.y < pMax.y) {
                R inv = R(1) / d.z, ta = (pMin.z - o.z) * inv, tb = (pMax.z - o.z) * inv;
                tMin = gmin(ta, tb); tMax = gmax(ta, tb);
                return tMax >= R(0) && tMax >= tMin;
            }
            return false;
        }
        V dInv(R(1) / d.x, R(1) / d.y, R(1) / d.z);
        V ta = (pMin - o) * dInv, tb = (pMax - o) * dInv;
        V mn(gmin(ta.x, tb.x), gmin(ta.y, tb.y), gmin(ta.z, tb.z));
        V mx(gmax(ta.x, tb.x), gmax(ta.y, tb.y), gmax(ta.z, tb.z));
        V dt = mx - mn;
        R tyz = mx.z - mn.y, tzx = mx.x - mn.z, txy = mx.y - mn.x;
        if (std::fabs(d.x) < eps) {
            if (dt.y + dt.z > tyz) { tMin = gmax(mn.y, mn.z); tMax = gmin(mx.y, mx.z); return tMax >= R(0) && tMax >= tMin; }
        }
        if (std::fabs(d.y) < eps) {
            if (dt.z + dt.x > tzx) { tMin = gmax(mn.z, mn.x); tMax = gmin(mx.z, mx.x); return tMax >= R(0) && tMax >= tMin; }
        }
        if (std::fabs(d.z) < eps) {
            if (dt.x + dt.y > txy) { tMin = gmax(mn.x, mn.y); tMax = gmin(mx.x, mx.y); return tMax >= R(0) && tMax >= tMin; }
        }
        if (dt.y + dt.z > tyz && dt.z + dt.x > tzx && dt.x + dt.y > txy) {
            tMin = gmax(gmax(mn.x, mn.y), mn.z);
            tMax = gmin(gmin(mx.x, mx.y), mx.z);
            return tMax >= R(0) && tMax >= tMin;
        }
        return false;
    }
    // intersection.glsl:367-427 (any = bvhTest)
    int bvh(V o, V d, R& dist, bool any)
    {
        int table = cubemapFace(-d) * s.bvh_size;
        int closest = -1, k = 0;
        cn.trav++;
        while (k != s.bvh_size) {
            const int32_t* row = s.hit_table + 3 * (size_t)(table + k);
            cn.bytes += 36;
            cn.rows++;
            R boxDist;
            bool hit = boxHit(row[0], o, d, boxDist);
            if (!hit || boxDist > dist) { k = row[2]; continue; }
            if (row[1] >= 0) {
                R t;
                if (triHitId(row[1], o, d, t) && t < dist) {
                    if (any) return row[1];
                    dist = t;
                    closest = row[1];
                }
            }
            k++;
        }
        return closest;
    }
    int bvhHit(V o, V d, R& dist) { dist = R(1e8f); return bvh(o, d, dist, false); }
    bool bvhTest(V o, V d, R dist) { return bvh(o, d, dist, true) >= 0; }
    bool visible(V x, V y)
    {
        R dist = length(x - y) - R(2e-5f);
        V wi = normalize(y - x);
        return !bvhTest(x + wi * R(1e-5f), wi, dist);
    }

    struct Surf { V ns, ng; W uv; };
    // intersection.glsl:188-224
    Surf surface(int id, V pnt)
    {
        cn.bytes += 108;
        int ia = s.indices[3 * id], ib = s.indices[3 * id + 1], ic = s.indices[3 * id + 2];
        V a = vert(ia), b = vert(ib), c = vert(ic);
        V na = nrm(ia), nb = nrm(ib), nc = nrm(ic);
        W ta = tex(ia), tb = tex(ib), tc = tex(ic);
        V pa = a - pnt, pb = b - pnt, pc = c - pnt;
        R areaInv = R(1) / length(cross(b - a, c - a));
        R la = length(cross(pb, pc)) * areaInv, lb = length(cross(pc, pa)) * areaInv;
        R lc = R(1) - la - lb;
        Surf r;
        r.ns = normalize((na * la + nb * lb) + nc * lc);
        r.ng = normalize(cross(pa, pb));
        r.uv = W((ta.x * la + tb.x * lb) + tc.x * lc, (ta.y * la + tb.y * lb) + tc.y * lc);
        if (dot(r.ns, r.ng) < R(0)) r.ng = -r.ng;
        return r;
    }
    void triVerts(int id, V& a, V& b, V& c)
    {
        cn.bytes += 48;
        a = vert(s.indices[3 * id]); b = vert(s.indices[3 * id + 1]); c = vert(s.indices[3 * id + 2]);
    }
    R areaOf(int id) { V a, b, c; triVerts(id, a, b, c); return triArea(a, b, c); }
    V shadingNormal(int id, V pnt)
    {
        cn.bytes += 84;
        int ia = s.indices[3 * id], ib = s.indices[3 * id + 1], ic = s.indices[3 * id + 2];
        V a = vert(ia), b = vert(ib), c = vert(ic);
        V pa = a - pnt, pb = b - pnt, pc = c - pnt;
        R areaInv = R(1) / length(cross(b - a, c - a));
        R la = length(cross(pb, pc)) * areaInv, lb = length(cross(pc, pa)) * areaInv;
        R lc = R(1) - la - lb;
        return normalize((nrm(ia) * la + nrm(ib) * lb) + nrm(ic) * lc);
    }

    // texture(...) with GL_LINEAR / GL_REPEAT
    static int wrap(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }
    W noiseAt(W uv)
    {
        int Wd = s.noise_width, H = s.noise_height;
        R u = uv.x * R(Wd) - R(0.5f), v = uv.y * R(H) - R(0.5f);
        R fu = std::floor(u), fv = std::floor(v);
        R al = u - fu, be = v - fv;
        int i0 = wrap((int)fu, Wd), j0 = wrap((int)fv, H), i1 = wrap((int)fu + 1, Wd), j1 = wrap((int)fv + 1, H);
        auto T = [&](int i, int j, int c) { return R(s.noise[((size_t)j * Wd + i) * 2 + c]); };
        R w00 = (R(1) - al) * (R(1) - be), w10 = al * (R(1) - be), w01 = (R(1) - al) * be, w11 = al * be;
        return W(((w00 * T(i0, j0, 0) + w10 * T(i1, j0, 0)) + w01 * T(i0, j1, 0)) + w11 * T(i1, j1, 0),
                 ((w00 * T(i0, j0, 1) + w10 * T(i1, j0, 1)) + w01 * T(i0, j1, 1)) + w11 * T(i1, j1, 1));
    }
    V envTex(int i, int j)
    {
        const uint16_t* t = s.env_rgb16f + ((size_t)j * s.env_width + i) * 3;
        return V(R(halfToF(t[0])), R(halfToF(t[1])), R(halfToF(t[2])));
    }
    V envAt(W uv)
    {
        cn.bytes += 24;
        int Wd = s.env_width, H = s.env_height;
        R u = uv.x * R(Wd) - R(0.5f), v = uv.y * R(H) - R(0.5f);
        R fu = std::floor(u), fv = std::floor(v);
        R al = u - fu, be = v - fv;
        int i0 = wrap((int)fu, Wd), j0 = wrap((int)fv, H), i1 = wrap((int)fu + 1, Wd), j1 = wrap((int)fv + 1, H);
        R w00 = (R(1) - al) * (R(1) - be), w10 = al * (R(1) - be), w01 = (R(1) - al) * be, w11 = al * be;
        return ((envTex(i0, j0) * w00 + envTex(i1, j0) * w10) + envTex(i0, j1) * w01) + envTex(i1, j1) * w11;
    }

    // texture2DArray(uTextures, vec3(fract(uv) * uvScale, texId)) (material_loader.glsl:8-16):
    // sRGB8 texels decoded by the GL_SRGB formula, GL_LINEAR at level 0, GL_REPEAT.
    static float srgbDecode(uint8_t b)
    {
        const double c = b / 255.0;
        return (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    V texel(int layer, int i, int j)
    {
        const uint8_t* t = s.tex_rgb8 + (((size_t)layer * s.tex_height + j) * s.tex_width + i) * 3;
        return V(R(srgbDecode(t[0])), R(srgbDecode(t[1])), R(srgbDecode(t[2])));
    }
    V texAt(int layer, W uv)
    {
        cn.bytes += 8 + 16;
        const int Wd = s.tex_width, H = s.tex_height;
        const R su = (uv.x - std::floor(uv.x)) * R(s.tex_scale[2 * layer]);
        const R sv = (uv.y - std::floor(uv.y)) * R(s.tex_scale[2 * layer + 1]);
        R u = su * R(Wd) - R(0.5f), v = sv * R(H) - R(0.5f);
        R fu = std::floor(u), fv = std::floor(v);
        R al = u - fu, be = v - fv;
        int i0 = wrap((int)fu, Wd), j0 = wrap((int)fv, H), i1 = wrap((int)fu + 1, Wd), j1 = wrap((int)fv + 1, H);
        R w00 = (R(1) - al) * (R(1) - be), w10 = al * (R(1) - be), w01 = (R(1) - al) * be, w11 = al * be;
        return ((texel(layer, i0, j0) * w00 + texel(layer, i1, j0) * w10) + texel(layer, i0, j1) * w01) +
               texel(layer, i1, j1) * w11;
    }

    // --------------------------------------------------------------- BSDFs
    struct P {
        V base;
        R ss = 0, metal = 0, rough = 0, spec = 0, specTint = 0, sheen = 0, sheenTint = 0, cc = 0, ccGloss = 0, ior = 0;
    };
    struct S { V wi; R pdf; V f; R eta; uint32_t flag; };
    static S invalid() { return S{V(0), R(0), V(0), R(0), 1u << 16}; }

    // microfacet.glsl
    static R schlickW(R c) { return pow5(R(1) - c); }
    static V schlickF(R c, V f0) { return f0 + (V(1) - f0) * pow5(R(1) - c); }
    static V schlickFr(R c, V f0, R r)
    {
        V m(gmax(R(1) - r, f0.x), gmax(R(1) - r, f0.y), gmax(R(1) - r, f0.z));
        return f0 + (m - f0) * pow5(R(1) - c);
    }
    static R schlickG(R c, R a) { R k = a * R(0.5f); return c / (c * (R(1) - k) + k); }
    static R smithG(V n, V wo, V wi, R a) { return schlickG(absDot(n, wo), a) * schlickG(absDot(n, wi), a); }
    static R ggx(R c, R a)
    {
        if (c < R(1e-6f)) return R(0);
        R a2 = a * a;
        R den = c * c * (a2 - R(1)) + R(1);
        den = den * den * M<R>::Pi;
        return a2 / den;
    }
    static R ggxPdfVis(V n, V m, V wo, R a) { return ggx(dot(n, m), a) * schlickG(dot(n, wo), a) * absDot(m, wo) / absDot(n, wo); }
    static V ggxSample(V n, R a, W u)
    {
        W xi = concentric(u);
        V h(xi.x, xi.y, std::sqrt(gmax(R(0), R(1) - xi.x * xi.x - xi.y * xi.y)));
        h = normalize(h * V(a, a, R(1)));
        return normalToWorld(n, h);
    }
    static V ggxSampleVis(V n, V wo, R a, W u)
    {
        M3<R> inv = inverse3(tbn(n));
        V vh = normalize(mul(inv, wo) * V(a, a, R(1)));
        R lensq = vh.x * vh.x + vh.y * vh.y;
        V t1 = lensq > R(0) ? V(-vh.y, vh.x, R(0)) / std::sqrt(lensq) : V(1, 0, 0);
        V t2 = cross(vh, t1);
        W xi = concentric(u);
        R sv = R(0.5f) * (R(1) + vh.z);
        xi.y = (R(1) - sv) * std::sqrt(R(1) - xi.x * xi.x) + sv * xi.y;
        V h = (t1 * xi.x + t2 * xi.y) + vh * std::sqrt(gmax(R(0), R(1) - xi.x * xi.x - xi.y * xi.y));
        h = normalize(V(h.x * a, h.y * a, gmax(R(0), h.z)));
        return normalToWorld(n, h);
    }
    static R gtr1(R c, R a)
    {
        R a2 = a * a;
        return (a2 - R(1)) / (R(2) * M<R>::Pi * M<R>::log(a) * (R(1) + (a2 - R(1)) * c * c));
    }
    static V gtr1Sample(V n, V wo, R a, W u)
    {
        R ct = std::sqrt(gmax(R(0), (R(1) - M<R>::pow(a, R(1) - u.x)) / (R(1) - a)));
        R st = std::sqrt(gmax(R(0), R(1) - ct * ct));
        R phi = R(2) * u.y * M<R>::Pi;
        R sp, cp;
        M<R>::sincos(phi, &sp, &cp);
        V m = normalize(V(cp * st, sp * st, ct));
        if (!sameHemisphere(n, wo, m)) m = -m;
        return normalize(normalToWorld(n, m));
    }

    // material.glsl:68-83
    static V lambert(const P& p) { return p.base * M<R>::PiInv; }
    static S lambertSample(V n, const P& p, V u)
    {
        R pd;
        V wi = cosineSample(n, W(u.y, u.z), &pd);
        return S{wi, satDot(n, wi) * M<R>::PiInv, p.base * M<R>::PiInv, R(1), 1u};
    }
    // material.glsl:85-149
    static V metalWF(V wo, V wi, V n, const P& p)
    {
        R a = sq(p.rough);
        V h = normalize(wi + wo);
        if (!sameHemisphere(n, wo, wi)) return V(0);
        R cwi = dot(n, wi), cwo = dot(n, wo);
        V f0 = mix(V(R(0.04f)), p.base, p.metal);
        V f = schlickFr(satDot(h, wo), f0, p.rough);
        R d = ggx(dot(n, h), a), g = smithG(n, wo, wi, a);
        V kd = V(1) - f;
        kd = kd * (R(1) - p.metal);
        R den = R(4) * cwo * cwi;
        if (den < R(1e-7f)) return V(0);
        return kd * p.base * M<R>::PiInv + f * d * g / den;
    }
    static R metalWFPdf(V wo, V wi, V n, const P& p)
    {
        R a = sq(p.rough);
        V h = normalize(wo + wi);
        R pdfDiff = satDot(n, wi) * M<R>::PiInv;
        R pdfSpec = ggxPdfVis(n, h, wo, a) / (R(4) * absDot(h, wo));
        R spec = R(1) / (R(2) - p.metal);
        return mixs(pdfDiff, pdfSpec, spec);
    }
    static S metalWFSample(V n, V wo, const P& p, V u)
    {
        R a = sq(p.rough), spec = R(1) / (R(2) - p.metal);
        uint32_t type = u.x > spec ? 1u : 2u;
        V wi;
        if (type == 1u) { R pd; wi = cosineSample(n, W(u.y, u.z), &pd); }
        else wi = reflect(-wo, ggxSampleVis(n, wo, a, W(u.y, u.z)));
        if (dot(n, wi) < R(0)) return invalid();
        return S{wi, metalWFPdf(wo, wi, n, p), metalWF(wo, wi, n, p), R(1), type};
    }
    // material.glsl:151-186
    static bool refr(V& wt, V wi, V n, R eta)
    {
        R ci = dot(n, wi);
        if (ci < R(0)) eta = R(1) / eta;
        R s2i = gmax(R(0), R(1) - ci * ci), s2t = s2i / (eta * eta);
        if (s2t >= R(1)) return false;
        R ct = std::sqrt(R(1) - s2t);
        if (ci < R(0)) ct = -ct;
        wt = normalize(-wi / eta + n * (ci / eta - ct));
        return true;
    }
    static R fresnel(R ci, R eta)
    {
        ci = gmin(gmax(ci, R(-1)), R(1));
        if (ci < R(0)) { eta = R(1) / eta; ci = -ci; }
        R si = std::sqrt(R(1) - ci * ci), stt = si / eta;
        if (stt >= R(1)) return R(1);
        R ct = std::sqrt(R(1) - stt * stt);
        R pa = (ci - eta * ct) / (ci + eta * ct), pe = (eta * ci - ct) / (eta * ci + ct);
        return (pa * pa + pe * pe) * R(0.5f);
    }
    // material.glsl:188-338
    static V diel(V wo, V wi, V n, const P& p, uint32_t mode)
    {
        if (p.rough < R(0.02f)) return V(0);
        V h = normalize(wo + wi);
        R hwo = absDot(h, wo), hwi = absDot(h, wi), a = p.rough * p.rough;
        if (sameHemisphere(n, wo, wi)) {
            R refl = fresnel(absDot(h, wi), p.ior);
            return (hwo * hwi < R(1e-7f)) ? V(0) : p.base * ggx(dot(n, h), a) * smithG(n, wo, wi, a) / (R(4) * hwo * hwi) * refl;
        }
        R eta = dot(n, wi) > R(0) ? p.ior : R(1) / p.ior;
        R sd = dot(h, wo) + eta * dot(h, wi);
        R den = sd * sd;
        den = den * (absDot(n, wi) * absDot(n, wo));
        R refl = fresnel(dot(h, wi), eta);
        R factor = (mode == 0) ? sq(R(1) / eta) : R(1);
        return (den < R(1e-7f)) ? V(0)
               : p.base * std::fabs(ggx(dot(n, h), a) * smithG(n, wo, wi, a) * hwo * hwi) / den * (R(1) - refl) * factor;
    }
    static R dielPdf(V wo, V wi, V n, const P& p)
    {
        if (p.rough < R(0.02f)) return R(0);
        if (sameHemisphere(n, wo, wi)) {
            V h = normalize(wo + wi);
            if (dot(wo, h) < R(0)) return R(0);
            R refl = fresnel(absDot(h, wi), p.ior);
            return ggx(dot(n, h), p.rough * p.rough) / (R(4) * absDot(h, wo)) * refl;
        }
        R eta = dot(n, wo) > R(0) ? p.ior : R(1) / p.ior;
        V h = normalize(wo + wi * eta);
        if (sameHemisphere(h, wo, wi)) return R(0);
        R trans = R(1) - fresnel(absDot(h, wo), eta);
        R dHdWi = absDot(h, wi) / sq(dot(h, wo) + eta * dot(h, wi));
        return ggx(dot(n, h), p.rough * p.rough) * dHdWi * trans;
    }
    static S dielSample(V n, V wo, const P& p, uint32_t mode, V u)
    {
        R ior = p.ior;
        if (p.rough < R(0.02f)) {
            R refl = fresnel(dot(n, wo), ior);
            if (u.x < refl) return S{reflect(-wo, n), R(1), p.base, R(1), 1u << 3};
            V wi;
            if (!refr(wi, wo, n, ior)) return invalid();
            if (dot(n, wo) < R(0)) ior = R(1) / ior;
            R factor = (mode == 0) ? sq(R(1) / ior) : R(1);
            return S{wi, R(1), p.base * factor, ior, 1u << 4};
        }
        R a = p.rough * p.rough;
        V h = ggxSample(n, a, W(u.y, u.z));
        if (dot(n, h) < R(0)) h = -h;
        R refl = fresnel(dot(h, wo), ior);
        if (u.x < refl) {
            V wi = -reflect(wo, h);
            if (!sameHemisphere(n, wo, wi)) return invalid();
            R pp = ggx(dot(n, h), a) / (R(4) * absDot(h, wo));
            R hwo = absDot(h, wo), hwi = absDot(h, wi);
            V r = (hwo * hwi < R(1e-7f)) ? V(0) : p.base * ggx(dot(n, h), a) * smithG(n, wo, wi, a) / (R(4) * hwo * hwi);
            if (std::isnan(pp)) pp = R(0);
            return S{wi, pp, r, R(1), 1u << 1};
        }
        V wi;
        if (!refr(wi, wo, h, ior)) return invalid();
        if (sameHemisphere(n, wo, wi)) return invalid();
        if (absDot(n, wi) < R(1e-10f)) return invalid();
        R hwo = absDot(h, wo), hwi = absDot(h, wi);
        if (dot(h, wo) < R(0)) ior = R(1) / ior;
        R sd = dot(h, wo) + ior * dot(h, wi);
        R den = sd * sd;
        R dHdWi = hwi / den;
        R factor = (mode == 0) ? sq(R(1) / ior) : R(1);
        den = den * (absDot(n, wi) * absDot(n, wo));
        V t = (den < R(1e-7f)) ? V(0) : p.base * std::fabs(ggx(dot(n, h), a) * smithG(n, wo, wi, a) * hwo * hwi) / den;
        R pp = ggx(dot(n, h), a) * dHdWi;
        if (std::isnan(pp)) pp = R(0);
        return S{wi, pp, t * factor, ior, 1u << 2};
    }
    // material.glsl:340-355
    static S thinSample(V n, V wo, const P& p, V u)
    {
        if (dot(n, wo) < R(0)) n = -n;
        R refl = fresnel(dot(n, wo), p.ior), trans = R(1) - refl;
        if (refl < R(1)) { refl += trans * trans * refl / (R(1) - refl * refl); trans = R(1) - refl; }
        return (u.x < refl) ? S{reflect(-wo, n), R(1), p.base, R(1), 1u << 3} : S{-wo, R(1), p.base, R(1), 1u << 4};
    }
    // material.glsl:357-555 (Disney)
    static V pMetal(V wo, V wi, V n, V fm0, R a)
    {
        R cwo = satDot(n, wo), cwi = satDot(n, wi);
        V h = normalize(wo + wi);
        if (cwo < R(1e-10f) || cwi < R(1e-10f)) return V(0);
        V fm = schlickF(absDot(h, wo), fm0);
        R dm = ggx(dot(n, h), a), gm = smithG(n, wo, wi, a);
        R den = R(4) * cwi * cwo;
        if (den < R(1e-7f)) return V(0);
        return fm * dm * gm / den;
    }
    static R pMetalPdf(V wo, V wi, V n, R a)
    {
        V h = normalize(wo + wi);
        return ggxPdfVis(n, h, wo, a) / (R(4) * absDot(h, wo));
    }
    static V pCoat(V wo, V wi, V n, V base, R a)
    {
        R cwo = satDot(n, wo), cwi = satDot(n, wi);
        V h = normalize(wo + wi);
        if (cwo < R(1e-6f) || cwi < R(1e-6f)) return V(0);
        V fc = schlickF(absDot(h, wo), base);
        R dc = gtr1(satDot(n, h), a), gc = smithG(n, wo, wi, R(0.25f));
        R den = R(4) * cwi * cwo;
        if (den < R(1e-7f)) return V(0);
        return fc * dc * gc / den;
    }
    static R pCoatPdf(V wo, V wi, V n, R a)
    {
        V h = normalize(wo + wi);
        return gtr1(satDot(n, h), a) * absDot(n, h) / (R(4) * absDot(h, wo));
    }
    static V pDiffuse(V wo, V wi, V n, V base, R ss, R rough)
    {
        R cwo = satDot(n, wo), cwi = satDot(n, wi);
        if (cwo < R(1e-10f) || cwi < R(1e-10f)) return V(0);
        V h = normalize(wo + wi);
        R hwi = dot(h, wi), hwi2 = hwi * hwi;
        R fi = schlickW(cwi), fo = schlickW(cwo);
        V fd90(R(0.5f) + R(2) * rough * hwi2);
        V fd = mix(V(1), fd90, fi) * mix(V(1), fd90, fo);
        V baseDiffuse = base * fd * M<R>::PiInv;
        V fss90(rough * hwi2);
        V fss = mix(V(1), fss90, fi) * mix(V(1), fss90, fo);
        V ssv = base * M<R>::PiInv * R(1.25f) * (fss * (R(1) / (cwi + cwo) - R(0.5f)) + V(R(0.5f)));
        return mix(baseDiffuse, ssv, ss);
    }
    static V tint(V base)
    {
        R lum = luminance(base);
        return lum > R(0) ? base / lum : V(1);
    }
    static V disney(V wo, V wi, V n, const P& p)
    {
        R a = sq(p.rough), ca = mixs(R(0.1f), R(0.001f), p.ccGloss);
        V tc = tint(p.base);
        V fm0 = mix(mix(V(1), tc, p.specTint) * (R(0.08f) * p.spec), p.base, p.metal);
        R hwi = dot(normalize(wo + wi), wi);
        V res(0);
        res = res + pDiffuse(wo, wi, n, p.base, p.ss, p.rough) * (R(1) - p.metal);
        res = res + pMetal(wo, wi, n, fm0, a);
        res = res + pCoat(wo, wi, n, p.base, ca) * p.cc * R(0.25f);
        res = res + mix(V(1), tc, p.sheenTint) * schlickW(hwi) * p.sheen * (dot(n, wi) < R(0) ? R(0) : R(1));
        return res;
    }
    static R disneyPdf(V wo, V wi, V n, const P& p)
    {
        R a = sq(p.rough), ca = mixs(R(0.1f), R(0.001f), p.ccGloss);
        R spec = R(1) / (R(2) - p.metal);
        R pdf = R(0);
        pdf += absDot(n, wi) * M<R>::PiInv * (R(1) - spec);
        pdf += pMetalPdf(wo, wi, n, a) * spec;
        pdf += pCoatPdf(wo, wi, n, ca) * R(0.25f) * p.cc;
        return pdf / (R(1) + R(0.25f) * p.cc);
    }
    S disneySample(V n, V wo, const P& p, V u)
    {
        R a = sq(p.rough), spec = R(1) / (R(2) - p.metal);
        R c0 = R(1) - spec, c1 = R(1), c2 = R(1) + p.cc * R(0.25f);
        R sv = rand() * c2;
        V wi;
        if (sv <= c0) { R pd; wi = cosineSample(n, W(u.y, u.z), &pd); }
        else if (sv <= c1) { wi = reflect(-wo, ggxSampleVis(n, wo, a, W(u.y, u.z))); if (dot(n, wi) < R(0)) wi = V(0); }
        else { wi = reflect(-wo, gtr1Sample(n, wo, a, W(u.y, u.z))); if (dot(n, wi) < R(0)) wi = V(0); }
        return S{wi, disneyPdf(wo, wi, n, p), disney(wo, wi, n, p), R(1), 1u};
    }
    // material_loader.glsl:3-169
    uint32_t matType(int m)
    {
        cn.bytes += 16;
        int32_t t;
        std::memcpy(&t, s.materials + 16 * (size_t)m + 13, 4);
        return (uint32_t)t;
    }
    // material_loader.glsl:8-97; texId != -1 takes baseColor from the texture array
    P loadMat(uint32_t t, int m, int texId = -1, W uv = W(0, 0))
    {
        const float* q = s.materials + 16 * (size_t)m;
        P p;
        p.base = (texId == -1) ? V(R(q[0]), R(q[1]), R(q[2])) : texAt(texId, uv);
        if (t == 1) {
            cn.bytes += 48;
            p.rough = R(mixs(0.0134f, 1.0f, q[3]));
            p.ss = R(q[4]); p.metal = R(q[5]); p.spec = R(q[6]); p.specTint = R(q[7]);
            p.sheen = R(q[8]); p.sheenTint = R(q[9]); p.cc = R(q[10]); p.ccGloss = R(q[11]);
        } else if (t == 2) {
            cn.bytes += 32;
            p.rough = R(mixs(0.0134f, 1.0f, q[3]));
            p.metal = R(q[5]);
        } else if (t == 3) {
            cn.bytes += 32;
            p.rough = R(q[3]);
            p.ior = R(q[12]);
        } else if (t == 4) {
            cn.bytes += 32;
            p.ior = R(q[12]);
        } else {
            cn.bytes += 16;
        }
        return p;
    }
    static V evalF(uint32_t t, const P& p, V wo, V wi, V n, uint32_t mode)
    {
        switch (t) {
        case 1: return disney(wo, wi, n, p);
        case 2: return metalWF(wo, wi, n, p);
        case 3: return diel(wo, wi, n, p, mode);
        case 4: return V(0);
        default: return lambert(p);
        }
    }
    static R evalPdf(uint32_t t, const P& p, V wo, V wi, V n)
    {
        switch (t) {
        case 1: return disneyPdf(wo, wi, n, p);
        case 2: return metalWFPdf(wo, wi, n, p);
        case 3: return dielPdf(wo, wi, n, p);
        case 4: return R(0);
        default: return satDot(wi, n) * M<R>::PiInv;
        }
    }
    S sampleF(uint32_t t, const P& p, V n, V wo, uint32_t mode, V u)
    {
        switch (t) {
        case 1: return disneySample(n, wo, p, u);
        case 2: return metalWFSample(n, wo, p, u);
        case 3: return dielSample(n, wo, p, mode, u);
        case 4: return thinSample(n, wo, p, u);
        default: return lambertSample(n, p, u);
        }
    }

    // --------------------------------------------------------------- lights (light.glsl)
    V power(int id) { return f3(s.light_power + 3 * (size_t)id); }
    int lightPick(W u)
    {
        cn.bytes += 8;
        int cx = (int)(R(s.n_light_triangles) * u.x);
        bool ok = cx >= 0 && cx < s.n_light_triangles;
        R prob = ok ? R(s.light_prob[cx]) : R(0);
        int al = ok ? s.light_alias[cx] : 0;
        return (u.y < prob) ? cx : al;
    }
    R lightPdfPick(int id) { cn.bytes += 12; return luminance(power(id)) / R(s.light_sum); }
    V lightLe(int id, V x, V wo)
    {
        int tri = id + s.obj_prim_count;
        Surf si = surface(tri, x);
        if (dot(wo, si.ng) <= R(0)) return V(0);
        cn.bytes += 12;
        return power(id) / areaOf(tri) * R(0.5f) * M<R>::PiInv;
    }
    R lightPdfLi(int id, V x, V y)
    {
        int tri = id + s.obj_prim_count;
        V nn = surface(tri, y).ng;
        R c = absDot(nn, normalize(x - y));
        if (c < R(1e-8f)) return R(-1);
        V d = x - y;
        return dot(d, d) / (areaOf(tri) * c);
    }
    struct Li { V wi, coef; R pdf; };
    Li lightSampleLi(int id, V x, W u)
    {
        int tri = id + s.obj_prim_count;
        V a, b, c;
        triVerts(tri, a, b, c);
        V y = triUniform(a, b, c, u);
        V wi = normalize(y - x);
        V nn = surface(tri, y).ng;
        R ct = dot(nn, -wi);
        if (ct < R(1e-6f)) return Li{V(0), V(0), R(0)};
        R dist = length(x - y);
        R pdf = dist * dist / (triArea(a, b, c) * ct);
        R test = dist - R(1e-4f) - R(1e-6f);
        if (bvhTest(x + wi * R(1e-4f), wi, test) || pdf < R(1e-8f)) return Li{V(0), V(0), R(0)};
        V w = lightLe(id, y, -wi);
        pdf = pdf * lightPdfPick(id);
        return Li{wi, w / pdf, pdf};
    }
    V envLe(V wi) { return envAt(sphereToPlane(rotZ(wi, rotCN, rotSN))); }
    R envPortion(V wi) { return luminance(envLe(wi)) / R(s.env_sum); }
    // the planted mistake kVarPlantEnvSin: a 1/sin(theta) Jacobian on top of the
    // sin(theta)-weighted table (the reference has none, light.glsl:175-179)
    R plantEnv(V wi, R pdf)
    {
        if (!(gVariant & kVarPlantEnvSin)) return pdf;
        V l = rotZ(wi, rotCN, rotSN);
        return pdf / std::sqrt(gmax(R(1e-12f), l.x * l.x + l.y * l.y));
    }
    R envPdfLi(V wi)
    {
        if (s.env_sum == 0.0f) return R(0);
        return plantEnv(wi, envPortion(wi) * R(s.env_width) * R(s.env_height) * R(0.5f) * sq(M<R>::PiInv));
    }
    R envProbAt(int x, int y)
    {
        if (x < 0 || x > s.env_width || y < 0 || y >= s.env_height) return R(0);
        return R(s.env_prob[(size_t)y * (s.env_width + 1) + x]);
    }
    int envAliasAt(int x, int y)
    {
        if (x < 0 || x > s.env_width || y < 0 || y >= s.env_height) return 0;
        return s.env_alias[(size_t)y * (s.env_width + 1) + x];
    }
    // envSampleWi (light.glsl:181-204): the row by the alias over the marginal kept in
    // column w, then the column by the row's alias; the texel centre, rotated.
    V envSampleWi(R u0, R u1, R u2, R u3, R& pdf, int* rowOut = nullptr, int* colOut = nullptr)
    {
        const int w = s.env_width, h = s.env_height;
        cn.bytes += 16;
        int rx = (int)(R(h) * u0);
        int row = (u1 < envProbAt(w, rx)) ? rx : envAliasAt(w, rx);
        int cx = (int)(R(w) * u2);
        int col = (u3 < envProbAt(cx, row)) ? cx : envAliasAt(cx, row);
        if (rowOut) { *rowOut = row; *colOut = col; }
        W uv((R(col) + R(0.5f)) / R(w), (R(row) + R(0.5f)) / R(h));
        V wi = rotZ(planeToSphere(uv), rotC, rotS);
        pdf = plantEnv(wi, envPortion(wi) * R(w) * R(h) * R(0.5f) * sq(M<R>::PiInv));
        return wi;
    }
    Li envSampleLi(V x, R u0, R u1, R u2, R u3)
    {
        R pdf;
        V wi = envSampleWi(u0, u1, u2, u3, pdf);
        if (bvhTest(x + wi * R(1e-4f), wi, R(1e8f)) || pdf == R(0)) return Li{V(0), V(0), R(0)};
        return Li{wi, envLe(wi) / pdf, pdf};
    }
    R pdfLightSel() const
    {
        if (s.n_light_triangles > 0)
            return p.light_env_uniform ? R(p.light_portion) : R(s.light_sum) / (R(s.light_sum) + R(s.env_sum));
        return R(0);
    }
    Li sampleLightEnv(V x, R ud, R u0, R u1, R u2, R u3)
    {
        R pl = pdfLightSel();
        bool sl = ud < pl;
        R sel = sl ? pl : R(1) - pl;
        Li r = sl ? lightSampleLi(lightPick(W(u0, u1)), x, W(u2, u3)) : envSampleLi(x, u0, u1, u2, u3);
        r.coef = r.coef / sel;
        r.pdf = r.pdf * sel;
        return r;
    }
    R pdfSelLight(int id)
    {
        R fst = lightPdfPick(id);
        R snd = p.light_env_uniform ? R(p.light_portion) : R(s.light_sum) / (R(s.light_sum) + R(s.env_sum));
        return fst * snd;
    }
    R pdfSelEnv() const
    {
        return p.light_env_uniform ? R(1) - R(p.light_portion) : R(s.env_sum) / (R(s.light_sum) + R(s.env_sum));
    }

    // --------------------------------------------------------------- camera (camera.glsl)
    V camF() const { return f3(s.cam_front); }
    V camR() const { return f3(s.cam_right); }
    V camU() const { return f3(s.cam_up); }
    V camP() const { return f3(s.cam_pos); }
    void cameraRay(W uv, R u0, R u1, R u2, R u3, V& o, V& d)
    {
        W ts(R(1) / R(s.film_width), R(1) / R(s.film_height));
        W bc(uv.x + ts.x * u0, uv.y + ts.y * u1);
        W ndc(bc.x * R(2) - R(1), bc.y * R(2) - R(1));
        W dk = concentric(W(u2, u3));
        V pl(dk.x * R(s.lens_radius), dk.y * R(s.lens_radius), R(0));
        V pf(ndc.x * R(s.aspect) * R(s.focal_dist) * R(s.tan_fov), ndc.y * R(1) * R(s.focal_dist) * R(s.tan_fov),
             R(s.focal_dist));
        V dir = pf - pl;
        d = normalize((camR() * dir.x + camU() * dir.y) + camF() * dir.z);
        o = (camP() + camR() * pl.x) + camU() * pl.y;
    }
    W rasterPos(V o, V d)
    {
        R ct = dot(d, camF());
        R df = R(s.focal_dist) / ct;
        M3<R> mi{f3(s.cam_mat_inv), f3(s.cam_mat_inv + 3), f3(s.cam_mat_inv + 6)};
        V pf = mul(mi, (o + d * df) - camP());
        R aspect = R(s.film_width) / R(s.film_height);
        V den = V(aspect * R(s.tan_fov), R(1) * R(s.tan_fov), R(1)) * R(s.focal_dist);
        pf = pf / den;
        return W((pf.x + R(1)) * R(0.5f), (pf.y + R(1)) * R(0.5f));
    }
    static bool inFilm(W uv) { return uv.x >= R(0) && uv.x <= R(1) && uv.y >= R(0) && uv.y <= R(1); }
    bool camDelta() const { return s.lens_radius <= 1e-6f; }
    V camIe(V o, V d)
    {
        R ct = dot(d, camF());
        if (ct < R(1e-6f)) return V(0);
        if (!inFilm(rasterPos(o, d))) return V(0);
        R tinv = R(1) / R(s.tan_fov), c2 = ct * ct;
        R la = camDelta() ? R(1) : M<R>::Pi * R(s.lens_radius) * R(s.lens_radius);
        return V(R(0.25f)) * sq(tinv / c2) / (la * R(s.aspect));
    }
    struct Ii { V wi, I; R dist; W uv; R pdf; };
    Ii camSampleIi(V ref, W u)
    {
        W dk = concentric(u);
        V pl(dk.x * R(s.lens_radius), dk.y * R(s.lens_radius), R(0));
        V y = ((camP() + camR() * pl.x) + camU() * pl.y) + camF() * pl.z;
        R dist = length(ref - y);
        V wi = normalize(y - ref);
        R ct = satDot(camF(), -wi);
        if (ct < R(1e-6f)) return Ii{V(0), V(0), R(0), W(0, 0), R(0)};
        V I = camIe(y, -wi);
        W uv = rasterPos(y, -wi);
        R la = camDelta() ? R(1) : M<R>::Pi * R(s.lens_radius) * R(s.lens_radius);
        return Ii{wi, I, dist, uv, dist * dist / (ct * la)};
    }
    void camPdfIe(V o, V d, R& pPos, R& pDir)
    {
        R ct = dot(camF(), d);
        pPos = pDir = R(0);
        if (ct < R(1e-6f)) return;
        if (!inFilm(rasterPos(o, d))) return;
        pPos = camDelta() ? R(1) : R(1) / (M<R>::Pi * R(s.lens_radius) * R(s.lens_radius));
        pDir = R(1) / (ct * ct * ct);
    }

    // --------------------------------------------------------------- integrators
    // path_integ_naive.glsl:35-143
    V naive(V o, V d)
    {
        R primDist;
        int id = bvhHit(o, d, primDist);
        V pos = o + d * primDist;
        if (id == -1) return envLe(d);
        if (id - s.obj_prim_count >= 0) return lightLe(id - s.obj_prim_count, pos, -d);
        V wo = -d, result(0), tp(1);
        for (int bounce = 1; bounce <= p.max_depth; bounce++) {
            Surf sf = surface(id, pos);
            cn.bytes += 4;
            int mat = s.mat_tex_indices[id] & 0xffff;
            uint32_t t = matType(mat);
            if (t != 3 && t != 4 && dot(sf.ns, wo) < R(0)) { sf.ns = -sf.ns; sf.ng = -sf.ng; }
            P mp = loadMat(t, mat, s.mat_tex_indices[id] >> 16, sf.uv);
            if (p.sample_light) {
                R ud = sample1D();
                R u0 = sample1D(), u1 = sample1D(), u2 = sample1D(), u3 = sample1D();
                Li ls = sampleLightEnv(pos, ud, u0, u1, u2, u3);
                if (ls.pdf > R(0)) {
                    V f = evalF(t, mp, wo, ls.wi, sf.ns, 0);
                    R fp = evalPdf(t, mp, wo, ls.wi, sf.ns);
                    R wgt = ls.pdf * ls.pdf / (ls.pdf * ls.pdf + fp * fp);
                    result = result + f * tp * satDot(sf.ns, ls.wi) * ls.coef * wgt;
                }
            }
            R a0 = sample1D(), a1 = sample1D(), a2 = sample1D();
            S bs = sampleF(t, mp, sf.ns, wo, 0, V(a0, a1, a2));
            bool delta = bs.flag == (1u << 3) || bs.flag == (1u << 4);
            if (bs.pdf < R(1e-8f)) break;
            tp = tp * (bs.f / bs.pdf * (delta ? R(1) : absDot(sf.ns, bs.wi)));
            V no = pos + bs.wi * R(1e-4f);
            R dist;
            int next = bvhHit(no, bs.wi, dist);
            int lid = next - s.obj_prim_count;
            V np = no + bs.wi * dist;
            if (next == -1) {
                V rad = envLe(bs.wi);
                R wgt = R(1);
                if (p.sample_light && !delta) {
                    R ep = envPdfLi(bs.wi) * pdfSelEnv();
                    wgt = (ep <= R(0)) ? R(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + ep * ep);
                }
                result = result + rad * tp * wgt;
                break;
            } else if (lid >= 0) {
                V rad = lightLe(lid, np, -bs.wi);
                R wgt = R(1);
                if (p.sample_light && !delta) {
                    R lp = lightPdfLi(lid, pos, np) * pdfSelLight(lid);
                    wgt = (lp <= R(0)) ? R(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + lp * lp);
                }
                result = result + rad * tp * wgt;
                break;
            }
            if (p.russian_roulette) {
                R cp = gmin(maxComp(bs.f / bs.pdf), R(0.95f));
                if (sample1D() >= cp) break;
                tp = tp / cp;
            }
            id = next;
            pos = np;
            wo = -bs.wi;
        }
        return result;
    }

    // ---------------------------------------------------------- global-queue path tracer
    // GlobalQueuePathIntegrator::renderOnePass (src/integrator/GlobalQueuePath.cpp:100-155)
    // over one pixel and one pass.  The primary ray (pt_global_queue_primary.glsl:40-86)
    // takes Sobol dims 0-3; the vertex of depth d = 1..maxDepth
    // (pt_global_queue_streaming.glsl:89-208) takes dims 4 + 9 d onwards -- NEE selector
    // and 4 light dims always drawn, then 3 BSDF dims, then the Russian-roulette dim -- and
    // restarts the pixel's seeds (randSeed, sampleSeed).  Every contribution is added to
    // the pixel when produced (accumulateImage, :51-55), with no NaN test.
    void globalQueuePass(int x, int y, int spp, uint32_t fc, R* acc)
    {
        auto add = [&](V c) { acc[0] = acc[0] + c.x; acc[1] = acc[1] + c.y; acc[2] = acc[2] + c.z; };
        seedPixel(x, y, spp, fc);
        cn.bytes += 64 + 32;
        R u0 = sample1D(), u1 = sample1D(), u2 = sample1D(), u3 = sample1D();
        W scr(R(x) / R(s.film_width), R(y) / R(s.film_height));
        V o, d;
        cameraRay(scr, u0, u1, u2, u3, o, d);
        R primDist;
        int id = bvhHit(o, d, primDist);
        V pos = o + d * primDist;
        if (id == -1) { add(envLe(d)); return; }
        if (id - s.obj_prim_count >= 0) { add(lightLe(id - s.obj_prim_count, pos, -d)); return; }
        V wo = -d, tp(1);
        for (int depth = 1; depth <= p.max_depth; depth++) {
            seedPixel(x, y, spp, fc);
            dim = 4 + 9 * depth;
            Surf sf = surface(id, pos);
            int mat = s.mat_tex_indices[id] & 0xffff;
            uint32_t t = matType(mat);
            if (t != 3 && t != 4 && dot(sf.ns, wo) < R(0)) { sf.ns = -sf.ns; sf.ng = -sf.ng; }
            P mp = loadMat(t, mat, s.mat_tex_indices[id] >> 16, sf.uv);
            R ud = sample1D();
            R v0 = sample1D(), v1 = sample1D(), v2 = sample1D(), v3 = sample1D();
            if (p.sample_light) {
                Li ls = sampleLightEnv(pos, ud, v0, v1, v2, v3);
                if (ls.pdf > R(0)) {
                    V f = evalF(t, mp, wo, ls.wi, sf.ns, 0);
                    R fp = evalPdf(t, mp, wo, ls.wi, sf.ns);
                    R wgt = ls.pdf * ls.pdf / (ls.pdf * ls.pdf + fp * fp);
                    add(f * tp * satDot(sf.ns, ls.wi) * ls.coef * wgt);
                }
            }
            R a0 = sample1D(), a1 = sample1D(), a2 = sample1D();
            S bs = sampleF(t, mp, sf.ns, wo, 0, V(a0, a1, a2));
            bool delta = bs.flag == (1u << 3) || bs.flag == (1u << 4);
            if (bs.pdf < R(1e-8f)) return;
            tp = tp * (bs.f / bs.pdf * (delta ? R(1) : absDot(sf.ns, bs.wi)));
            V no = pos + bs.wi * R(1e-4f);
            R dist;
            int next = bvhHit(no, bs.wi, dist);
            int lid = next - s.obj_prim_count;
            V np = no + bs.wi * dist;
            if (next == -1) {
                V rad = envLe(bs.wi);
                R wgt = R(1);
                if (p.sample_light && !delta) {
                    R ep = envPdfLi(bs.wi) * pdfSelEnv();
                    wgt = (ep <= R(0)) ? R(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + ep * ep);
                }
                add(rad * tp * wgt);
                return;
            } else if (lid >= 0) {
                V rad = lightLe(lid, np, -bs.wi);
                R wgt = R(1);
                if (p.sample_light && !delta) {
                    R lp = lightPdfLi(lid, pos, np) * pdfSelLight(lid);
                    wgt = (lp <= R(0)) ? R(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + lp * lp);
                }
                add(rad * tp * wgt);
                return;
            }
            if (p.russian_roulette) {
                R cp = gmin(maxComp(bs.f / bs.pdf), R(0.95f));
                if (sample1D() >= cp) return;
                tp = tp / cp;
            }
            id = next;
            pos = np;
            wo = -bs.wi;
        }
    }

    // ---------------------------------------------------------- differentiable naive
    // The naive estimator with sampling at the scene (theta0) and BSDF values of
    // differentiable materials (principled, Lambertian) and environment radiance at
    // the evaluation point: material records `ce.s.materials`, FP32 env texels `ev`.
    static bool differentiable(uint32_t t) { return t != 2 && t != 3 && t != 4; }
    V envAtEval(W uv, const float* ev, RepTap* tap)
    {
        int Wd = s.env_width, H = s.env_height;
        R u = uv.x * R(Wd) - R(0.5f), v = uv.y * R(H) - R(0.5f);
        R fu = std::floor(u), fv = std::floor(v);
        R al = u - fu, be = v - fv;
        int i0 = wrap((int)fu, Wd), j0 = wrap((int)fv, H), i1 = wrap((int)fu + 1, Wd), j1 = wrap((int)fv + 1, H);
        R w00 = (R(1) - al) * (R(1) - be), w10 = al * (R(1) - be), w01 = (R(1) - al) * be, w11 = al * be;
        int id[4] = {j0 * Wd + i0, j0 * Wd + i1, j1 * Wd + i0, j1 * Wd + i1};
        R ww[4] = {w00, w10, w01, w11};
        if (tap)
            for (int k = 0; k < 4; k++) { tap->idx[k] = id[k]; tap->w[k] = (double)ww[k]; }
        auto T = [&](int k) { return V(R(ev[3 * (size_t)id[k]]), R(ev[3 * (size_t)id[k] + 1]), R(ev[3 * (size_t)id[k] + 2])); };
        return ((T(0) * w00 + T(1) * w10) + T(2) * w01) + T(3) * w11;
    }
    V envLeEval(V wi, const float* ev, RepTap* tap) { return envAtEval(sphereToPlane(rotZ(wi, rotCN, rotSN)), ev, tap); }

    V naiveReplay(V o, V d, Ctx& ce, const float* ev, RepSample* rec)
    {
        auto put3 = [](float* dst, V x) { dst[0] = (float)x.x; dst[1] = (float)x.y; dst[2] = (float)x.z; };
        R primDist;
        int id = bvhHit(o, d, primDist);
        V pos = o + d * primDist;
        if (id == -1) {
            RepTerm t = {};
            t.kind = 0;
            V le = envLeEval(d, ev, &t.tap);
            if (rec) rec->ts.push_back(t);
            return le;
        }
        if (id - s.obj_prim_count >= 0) return lightLe(id - s.obj_prim_count, pos, -d);
        V wo = -d, result(0), tp(1);
        for (int bounce = 1; bounce <= p.max_depth; bounce++) {
            Surf sf = surface(id, pos);
            int mat = s.mat_tex_indices[id] & 0xffff;
            uint32_t t = matType(mat);
            if (t != 3 && t != 4 && dot(sf.ns, wo) < R(0)) { sf.ns = -sf.ns; sf.ng = -sf.ng; }
            P mp = loadMat(t, mat, s.mat_tex_indices[id] >> 16, sf.uv);
            const bool diff = differentiable(t);
            P mpE = diff ? ce.loadMat(t, mat, s.mat_tex_indices[id] >> 16, sf.uv) : mp;
            const int vi = rec ? (int)rec->vs.size() : 0;
            if (p.sample_light) {
                R ud = sample1D();
                R u0 = sample1D(), u1 = sample1D(), u2 = sample1D(), u3 = sample1D();
                R pl = pdfLightSel();
                bool sl = ud < pl;
                R sel = sl ? pl : R(1) - pl;
                RepTerm term = {};
                Li ls{V(0), V(0), R(0)};
                R envPdf = 0;
                if (sl) {
                    ls = lightSampleLi(lightPick(W(u0, u1)), pos, W(u2, u3));
                } else {
                    const int w = s.env_width, h = s.env_height;
                    int rx = (int)(R(h) * u0);
                    int row = (u1 < envProbAt(w, rx)) ? rx : envAliasAt(w, rx);
                    int cx = (int)(R(w) * u2);
                    int col = (u3 < envProbAt(cx, row)) ? cx : envAliasAt(cx, row);
                    W uv((R(col) + R(0.5f)) / R(w), (R(row) + R(0.5f)) / R(h));
                    V we = rotZ(planeToSphere(uv), rotC, rotS);
                    R pdf = envPortion(we) * R(w) * R(h) * R(0.5f) * sq(M<R>::PiInv);
                    if (!bvhTest(pos + we * R(1e-4f), we, R(1e8f)) && pdf != R(0)) {
                        envPdf = pdf;
                        ls = Li{we, envLeEval(we, ev, &term.tap) / pdf, pdf};
                    }
                }
                ls.coef = ls.coef / sel;
                ls.pdf = ls.pdf * sel;
                if (ls.pdf > R(0)) {
                    V f = evalF(t, mpE, wo, ls.wi, sf.ns, 0);
                    R fp = evalPdf(t, mp, wo, ls.wi, sf.ns);
                    R wgt = ls.pdf * ls.pdf / (ls.pdf * ls.pdf + fp * fp);
                    result = result + f * tp * satDot(sf.ns, ls.wi) * ls.coef * wgt;
                    if (rec) {
                        term.kind = sl ? 1 : 2;
                        term.v = vi;
                        term.mat = mat; term.type = t; term.diff = diff;
                        put3(term.wo, wo); put3(term.ns, sf.ns); put3(term.f, f);
                        put3(term.wl, ls.wi);
                        term.satDot = (float)satDot(sf.ns, ls.wi);
                        term.weight = (float)wgt;
                        term.envPdf = (float)envPdf;
                        term.sel = (float)sel;
                        put3(term.coef, ls.coef);
                        rec->ts.push_back(term);
                    }
                }
            }
            R a0 = sample1D(), a1 = sample1D(), a2 = sample1D();
            S bs = sampleF(t, mp, sf.ns, wo, 0, V(a0, a1, a2));
            bool delta = bs.flag == (1u << 3) || bs.flag == (1u << 4);
            if (bs.pdf < R(1e-8f)) break;
            V fE = diff ? evalF(t, mpE, wo, bs.wi, sf.ns, 0) : bs.f;
            const R cosf = delta ? R(1) : absDot(sf.ns, bs.wi);
            tp = tp * (fE / bs.pdf * cosf);
            if (rec) {
                RepVertex vx = {};
                vx.mat = mat; vx.type = t; vx.diff = diff;
                put3(vx.wo, wo); put3(vx.wi, bs.wi); put3(vx.ns, sf.ns);
                vx.pdf = (float)bs.pdf; vx.cosf = (float)cosf;
                put3(vx.fConst, bs.f);
                put3(vx.fE, fE);
                vx.rr = 1.0f;
                rec->vs.push_back(vx);
            }
            V no = pos + bs.wi * R(1e-4f);
            R dist;
            int next = bvhHit(no, bs.wi, dist);
            int lid = next - s.obj_prim_count;
            V np = no + bs.wi * dist;
            if (next == -1) {
                RepTerm term = {};
                V rad = envLeEval(bs.wi, ev, &term.tap);
                R wgt = R(1);
                if (p.sample_light && !delta) {
                    R ep = envPdfLi(bs.wi) * pdfSelEnv();
                    wgt = (ep <= R(0)) ? R(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + ep * ep);
                }
                result = result + rad * tp * wgt;
                if (rec) { term.kind = 3; term.v = vi; term.weight = (float)wgt; rec->ts.push_back(term); }
                break;
            } else if (lid >= 0) {
                V rad = lightLe(lid, np, -bs.wi);
                R wgt = R(1);
                if (p.sample_light && !delta) {
                    R lp = lightPdfLi(lid, pos, np) * pdfSelLight(lid);
                    wgt = (lp <= R(0)) ? R(0) : bs.pdf * bs.pdf / (bs.pdf * bs.pdf + lp * lp);
                }
                result = result + rad * tp * wgt;
                if (rec) {
                    RepTerm term = {};
                    term.kind = 4; term.v = vi; term.weight = (float)wgt; put3(term.le, rad);
                    rec->ts.push_back(term);
                }
                break;
            }
            if (p.russian_roulette) {
                R cp = gmin(maxComp(bs.f / bs.pdf), R(0.95f));
                if (sample1D() >= cp) break;
                tp = tp / cp;
                if (rec) rec->vs.back().rr = (float)cp;
            }
            id = next;
            pos = np;
            wo = -bs.wi;
        }
        return result;
    }

    static R remap(R x) { return x < R(1e-8f) ? R(1) : x * x; }

    // triple_path_pass_pt.glsl:59-195
    V tripleCamera(V o, V d)
    {
        R primDist;
        int id = bvhHit(o, d, primDist);
        V pos = o + d * primDist;
        if (id == -1) return envLe(d);
        if (id - s.obj_prim_count >= 0) return lightLe(id - s.obj_prim_count, pos, -d);
        Surf sf = surface(id, pos);
        V wo = -d, prevNorm = camF(), result(0), tp(1);
        R pPos, pDir;
        camPdfIe(o, d, pPos, pDir);
        R prim = remap(pPos) / remap(pDir * absDot(sf.ns, d) / sq(primDist));
        R t1s0 = prim, t1s1 = prim;
        for (int bounce = 1; bounce <= p.max_depth; bounce++) {
            if (bounce > 1) sf = surface(id, pos);
            cn.bytes += 4;
            int mat = s.mat_tex_indices[id] & 0xffff;
            uint32_t t = matType(mat);
            if (t != 3 && t != 4 && dot(sf.ns, wo) < R(0)) { sf.ns = -sf.ns; sf.ng = -sf.ng; }
            P mp = loadMat(t, mat, s.mat_tex_indices[id] >> 16, sf.uv);
            {
                R l0 = sample1D(), l1 = sample1D();
                int light = lightPick(W(l0, l1));
                int tri = light + s.obj_prim_count;
                R pdfSource = lightPdfPick(light);
                R q0 = sample1D(), q1 = sample1D();
                V a, b, c;
                triVerts(tri, a, b, c);
                V pLit = triUniform(a, b, c, W(q0, q1));
                V wi = normalize(pLit - pos);
                V Le = lightLe(light, pLit, -wi);
                if (!isBlack(Le) && visible(pos, pLit)) {
                    V nLit = surface(tri, pLit).ng;
                    R pA = pdfSource / areaOf(tri);
                    V dd = pos - pLit;
                    R dist2 = dot(dd, dd);
                    R pS = pA * dist2 / absDot(nLit, wi);
                    V f = evalF(t, mp, wo, wi, sf.ns, 0);
                    R fp = evalPdf(t, mp, wo, wi, sf.ns);
                    R pdfRev = evalPdf(t, mp, wi, wo, sf.ns);
                    R pdfPLit = remap(pA);
                    const R emitPdf = (gVariant & kVarEmitCos) ? satDot(nLit, -wi) * M<R>::PiInv
                                                                : R(0.5f) * M<R>::PiInv;
                    R cSurf = remap(emitPdf * absDot(sf.ns, wi));
                    R cLight = remap(fp * satDot(nLit, -wi));
                    R cPrev = (bounce == 1 || (gVariant & kVarPlantPrev)) ? R(1)
                                                                        : remap(pdfRev * absDot(prevNorm, wo));
                    R cDist = remap(dist2);
                    R s1s0 = pdfPLit * cDist / cLight;
                    R t1s1x = t1s1 * cSurf * cPrev / cDist;
                    R wgt = s1s0 / (R(1) + s1s0 + s1s0 * t1s1x);
                    result = result + Le * f * tp * absDot(sf.ns, wi) / pS * wgt;
                }
            }
            R a0 = sample1D(), a1 = sample1D(), a2 = sample1D();
            S bs = sampleF(t, mp, sf.ns, wo, 0, V(a0, a1, a2));
            bool delta = bs.flag == (1u << 3) || bs.flag == (1u << 4);
            if (bs.pdf < R(1e-8f)) break;
            tp = tp * (bs.f / bs.pdf * (delta ? R(1) : absDot(sf.ns, bs.wi)));
            V no = pos + bs.wi * R(1e-4f);
            R dist;
            int next = bvhHit(no, bs.wi, dist);
            int lid = next - s.obj_prim_count;
            V np = no + bs.wi * dist;
            R toNext = evalPdf(t, mp, wo, bs.wi, sf.ns);
            R toPrev = evalPdf(t, mp, bs.wi, wo, sf.ns);
            if (next == -1) break;
            if (lid >= 0) {
                V nLit = surface(next, np).ng;
                // lightPdfLe (light.glsl:268-277)
                V nl = surface(lid + s.obj_prim_count, np).ng;
                R lpPos = R(1) / areaOf(lid + s.obj_prim_count);
                R lpDir = (dot(nl, -bs.wi) <= R(0)) ? R(0)
                          : (gVariant & kVarEmitCos) ? dot(nl, -bs.wi) * M<R>::PiInv : R(0.5f) * M<R>::PiInv;
                R pdfPLit = remap(lpPos * lightPdfPick(lid));
                R cLight = remap(toNext * satDot(nLit, -bs.wi));
                R cSurf = remap(lpDir * absDot(sf.ns, bs.wi));
                R cPrev = (bounce == 1) ? R(1) : remap(toPrev * absDot(prevNorm, wo));
                R cDist = remap(dist * dist);
                R s1s0 = pdfPLit * cDist / cLight;
                R t1s0x = t1s0 * cSurf * pdfPLit * cPrev / cLight;
                R wgt = std::isnan(t1s0) ? R(0) : R(1) / (R(1) + s1s0 + t1s0x);
                result = result + lightLe(lid, np, -bs.wi) * tp * wgt;
                break;
            }
            if (p.russian_roulette) {
                R cp = gmin(maxComp(bs.f / bs.pdf), R(0.95f));
                if (sample1D() >= cp) break;
                tp = tp / cp;
            }
            R coef = ((bounce == 1) ? R(1) : remap(toPrev * absDot(prevNorm, wo))) /
                     remap(toNext * absDot(shadingNormal(next, np), bs.wi));
            t1s0 = t1s0 * coef;
            t1s1 = t1s1 * coef;
            prevNorm = sf.ns;
            pos = np;
            wo = -bs.wi;
            id = next;
        }
        return result;
    }

    // film splat (light_path_integ.glsl:34-43)
    void splat(std::vector<double>& film, W uv, V res)
    {
        if (!inFilm(uv)) return;
        int ix = (int)(uv.x * R(s.film_width)), iy = (int)(uv.y * R(s.film_height));
        if (ix >= s.film_width || iy >= s.film_height) return;
        size_t o = (size_t)iy * 3 * s.film_width + 3 * (size_t)ix;
        film[o] += (double)res.x;
        film[o + 1] += (double)res.y;
        film[o + 2] += (double)res.z;
    }

    // light_path_integ.glsl:45-146
    void lightPath(std::vector<double>& film)
    {
        R l0 = sample1D(), l1 = sample1D();
        int light = lightPick(W(l0, l1));
        R pdfSource = lightPdfPick(light);
        int tri = light + s.obj_prim_count;
        V o, d, wo, tp;
        {
            R q0 = sample1D(), q1 = sample1D();
            V a, b, c;
            triVerts(tri, a, b, c);
            V pLit = triUniform(a, b, c, W(q0, q1));
            R c0 = sample1D(), c1 = sample1D();
            Ii ci = camSampleIi(pLit, W(c0, c1));
            if (ci.pdf > R(0)) {
                V pCam = pLit + ci.wi * ci.dist;
                R pdfPos = R(1) / areaOf(tri);
                if (visible(pLit, pCam)) {
                    V Le = lightLe(light, pLit, ci.wi);
                    V contrib = Le * ci.I / (ci.pdf * pdfPos * pdfSource);
                    if (!isBlack(contrib)) { splat(film, ci.uv, contrib); cn.bytes += 12; }
                }
            }
            R e0 = sample1D(), e1 = sample1D(), e2 = sample1D(), e3 = sample1D();
            // lightSampleOneLe (light.glsl:111-120)
            V b0, b1, b2;
            triVerts(tri, b0, b1, b2);
            V ori = triUniform(b0, b1, b2, W(e0, e1));
            V nrmL = surface(tri, ori).ng;
            R pdfDir;
            V dir = cosineSample(nrmL, W(e2, e3), &pdfDir);
            V leO = ori + dir * R(1e-4f);
            V Le = lightLe(light, ori, dir);
            R pdfPos = R(1) / areaOf(tri);
            V nl = surface(tri, leO).ng;
            wo = -dir;
            o = leO + dir * R(1e-4f);
            d = dir;
            tp = Le * absDot(nl, -wo) / (pdfSource * pdfPos * pdfDir);
        }
        for (int bounce = 1; bounce <= p.max_depth; bounce++) {
            R dist;
            int id = bvhHit(o, d, dist);
            if (id == -1) break;
            if (id - s.obj_prim_count >= 0) break;
            V pos = o + d * dist;
            Surf sf = surface(id, pos);
            cn.bytes += 4;
            int mat = s.mat_tex_indices[id] & 0xffff;
            uint32_t t = matType(mat);
            if (t != 3 && t != 4 && dot(sf.ns, wo) < R(0)) { sf.ns = -sf.ns; sf.ng = -sf.ng; }
            P mp = loadMat(t, mat, s.mat_tex_indices[id] >> 16, sf.uv);
            {
                R c0 = sample1D(), c1 = sample1D();
                Ii ci = camSampleIi(pos, W(c0, c1));
                if (ci.pdf > R(0)) {
                    V pCam = pos + ci.wi * ci.dist;
                    if (visible(pos, pCam)) {
                        V f = evalF(t, mp, wo, ci.wi, sf.ns, 1);
                        R cosWi = satDot(sf.ng, ci.wi) * std::fabs(dot(sf.ns, wo) / dot(sf.ng, wo));
                        V res = ci.I * f * tp * cosWi / ci.pdf;
                        if (!hasNan(res) && !std::isnan(ci.pdf) && ci.pdf > R(1e-8f) && !isBlack(res)) {
                            splat(film, ci.uv, res);
                            cn.bytes += 12;
                        }
                    }
                }
            }
            R a0 = sample1D(), a1 = sample1D(), a2 = sample1D();
            S bs = sampleF(t, mp, sf.ns, wo, 1, V(a0, a1, a2));
            bool delta = bs.flag == (1u << 3) || bs.flag == (1u << 4);
            if (bs.pdf < R(1e-8f) || std::isnan(bs.pdf)) break;
            if (p.russian_roulette) {
                R cp = gmin(maxComp(bs.f / bs.pdf), R(1));
                if (sample1D() >= cp) break;
                tp = tp / cp;
            }
            R cosWi = delta ? R(1) : std::fabs(dot(sf.ng, bs.wi) * dot(sf.ns, wo) / dot(sf.ng, wo));
            tp = tp * (bs.f * cosWi / bs.pdf);
            o = pos + bs.wi * R(1e-4f);
            d = bs.wi;
            wo = -bs.wi;
        }
    }

    // triple_path_pass_lpt.glsl:58-182
    void tripleLight(std::vector<double>& film, R scale)
    {
        R l0 = sample1D(), l1 = sample1D();
        int light = lightPick(W(l0, l1));
        R pdfSource = lightPdfPick(light);
        int tri = light + s.obj_prim_count;
        V o, d, wo, tp, prevNorm;
        R prevPdfDir, s0t1, s1t1;
        {
            R q0 = sample1D(), q1 = sample1D();
            V a, b, c;
            triVerts(tri, a, b, c);
            (void)q0; (void)q1;
            R e0 = sample1D(), e1 = sample1D(), e2 = sample1D(), e3 = sample1D();
            V b0, b1, b2;
            triVerts(tri, b0, b1, b2);
            V ori = triUniform(b0, b1, b2, W(e0, e1));
            V nrmL = surface(tri, ori).ng;
            R pdfDir;
            V dir = cosineSample(nrmL, W(e2, e3), &pdfDir);
            V leO = ori + dir * R(1e-4f);
            V Le = lightLe(light, ori, dir);
            R pdfPos = R(1) / areaOf(tri);
            V nl = surface(tri, leO).ng;
            wo = -dir;
            o = leO + dir * R(1e-4f);
            d = dir;
            tp = Le * absDot(nl, -wo) / (pdfSource * pdfPos * pdfDir);
            prevNorm = nl;
            prevPdfDir = pdfDir;
            s0t1 = R(1) / remap(pdfPos * pdfSource);
            s1t1 = R(1);
        }
        for (int bounce = 1; bounce <= p.max_depth; bounce++) {
            R dist;
            int id = bvhHit(o, d, dist);
            if (id == -1) break;
            if (id - s.obj_prim_count >= 0) break;
            V pos = o + d * dist;
            Surf sf = surface(id, pos);
            cn.bytes += 4;
            int mat = s.mat_tex_indices[id] & 0xffff;
            uint32_t t = matType(mat);
            if (t != 3 && t != 4 && dot(sf.ns, wo) < R(0)) { sf.ns = -sf.ns; sf.ng = -sf.ng; }
            R cPos = remap(prevPdfDir * absDot(sf.ns, wo));
            s0t1 = s0t1 / cPos;
            s1t1 = s1t1 / (cPos / (bounce == 1 ? remap(dist * dist) : R(1)));
            P mp = loadMat(t, mat, s.mat_tex_indices[id] >> 16, sf.uv);
            {
                R c0 = sample1D(), c1 = sample1D();
                Ii ci = camSampleIi(pos, W(c0, c1));
                if (ci.pdf > R(0)) {
                    V pCam = pos + ci.wi * ci.dist;
                    if (visible(pos, pCam)) {
                        R cosWi = satDot(sf.ng, ci.wi) * std::fabs(dot(sf.ns, wo) / dot(sf.ng, wo));
                        V f = evalF(t, mp, wo, ci.wi, sf.ns, 1);
                        V contrib = ci.I * f * tp * cosWi / ci.pdf;
                        R pp, pd;
                        camPdfIe(pCam, -ci.wi, pp, pd);
                        R cSurf = remap(pd * satDot(sf.ns, ci.wi));
                        R cPrev = remap(evalPdf(t, mp, ci.wi, wo, sf.ns) * absDot(prevNorm, wo));
                        R cDist = remap(ci.dist * ci.dist);
                        R k0 = cSurf * cPrev / cDist;
                        R k1 = ((bounce == 1) ? R(1) : cPrev) * cSurf / cDist;
                        R wgt = R(1) / (s0t1 * k0 + s1t1 * k1 + R(1));
                        V res = contrib * wgt;
                        if (!hasNan(res) && !std::isnan(ci.pdf) && ci.pdf > R(1e-8f) && !isBlack(res)) {
                            splat(film, ci.uv, res * scale);
                            cn.bytes += 12;
                        }
                    }
                }
            }
            R a0 = sample1D(), a1 = sample1D(), a2 = sample1D();
            S bs = sampleF(t, mp, sf.ns, wo, 1, V(a0, a1, a2));
            bool delta = bs.flag == (1u << 3) || bs.flag == (1u << 4);
            if (bs.pdf < R(1e-8f) || std::isnan(bs.pdf)) break;
            if (p.russian_roulette) {
                R cp = gmin(maxComp(bs.f / bs.pdf), R(1));
                if (sample1D() >= cp) break;
                tp = tp / cp;
            }
            R cPrev = remap(evalPdf(t, mp, bs.wi, wo, sf.ns) * absDot(prevNorm, wo));
            s0t1 = s0t1 * cPrev;
            s1t1 = s1t1 * ((bounce == 1) ? R(1) : cPrev);
            prevPdfDir = evalPdf(t, mp, wo, bs.wi, sf.ns);
            prevNorm = sf.ns;
            R cosWi = delta ? R(1) : std::fabs(dot(sf.ng, bs.wi) * dot(sf.ns, wo) / dot(sf.ng, wo));
            tp = tp * (bs.f * cosWi / bs.pdf);
            o = pos + bs.wi * R(1e-4f);
            d = bs.wi;
            wo = -bs.wi;
        }
    }

    // per-pixel seeding (path_integ_naive.glsl:152-164)
    void seedPixel(int x, int y, int spp, uint32_t fc)
    {
        W scr(R(x) / R(s.film_width), R(y) / R(s.film_height));
        W nc = noiseAt(noiseAt(scr));
        R tx = R(s.film_width) * nc.x, ty = R(s.film_height) * nc.y;
        uint32_t ux = (uint32_t)tx, uy = (uint32_t)ty;
        randSeed = ux * fc + uy;
        sampleSeed = ux * uy;
        int off = (int)((uint32_t)spp * 256u);  // uSpp * uSampleDim, wrapping as GLSL int arithmetic
        if (off > 131072 * 256) off -= 131072 * 256;
        sampleOffset = off;
        dim = 0;
    }
};

template <class R>
static void renderRegion(const orc_scene& s, const orc_params& p, int x0, int y0, int w, int h, float* out,
                         Cnt& cn, int rowBegin, int rowEnd)
{
    const int samplerType = s.sampler ? 1 : 0;
    for (int yy = rowBegin; yy < rowEnd; yy++) {
        for (int xx = 0; xx < w; xx++) {
            const int x = x0 + xx, y = y0 + yy;
            R acc[3] = {0, 0, 0};
            for (int k = 0; k < p.n_passes; k++) {
                const int spp = p.first_pass + k;
                Ctx<R> c(s, p, cn, samplerType);
                if (p.integrator == 3) {
                    c.globalQueuePass(x, y, spp, (uint32_t)(p.free_counter + k), acc);
                    continue;
                }
                c.seedPixel(x, y, spp, (uint32_t)(p.free_counter + k));
                cn.bytes += 64 + 32;
                R u0 = c.sample1D(), u1 = c.sample1D(), u2 = c.sample1D(), u3 = c.sample1D();
                V2<R> scr(R(x) / R(s.film_width), R(y) / R(s.film_height));
                V3<R> o, d;
                c.cameraRay(scr, u0, u1, u2, u3, o, d);
                V3<R> r = (p.integrator == 2) ? c.tripleCamera(o, d) : c.naive(o, d);
                if (!hasNan(r)) {
                    // naive: the frame is read back only when uSpp != 0 (path_integ_naive.glsl:170)
                    const bool fresh = (p.integrator != 2 && spp == 0);
                    acc[0] = (fresh ? R(0) : acc[0]) + r.x;
                    acc[1] = (fresh ? R(0) : acc[1]) + r.y;
                    acc[2] = (fresh ? R(0) : acc[2]) + r.z;
                }
            }
            float* o3 = out + 3 * ((size_t)yy * w + xx);
            o3[0] = (float)acc[0]; o3[1] = (float)acc[1]; o3[2] = (float)acc[2];
        }
    }
}

// ---------------------------------------------------------------- differentiable naive
// FP64 value and Jacobian of a differentiable BSDF with respect to the five record
// parameters (slots 0, 1, 2, 3, 5), by central differences of the FP64 evaluation.
static const int kRepSlots[5] = {0, 1, 2, 3, 5};
static V3<double> bsdf64At(const double* r, uint32_t t, V3<double> wo, V3<double> wi, V3<double> n)
{
    using D = Ctx<double>;
    D::P p;
    p.base = V3<double>(r[0], r[1], r[2]);
    if (t == 1) {
        p.rough = mixs(0.0134, 1.0, r[3]);
        p.ss = r[4]; p.metal = r[5]; p.spec = r[6]; p.specTint = r[7];
        p.sheen = r[8]; p.sheenTint = r[9]; p.cc = r[10]; p.ccGloss = r[11];
    }
    return D::evalF(t == 1 ? 1u : 0u, p, wo, wi, n, 0);
}
static void bsdf64(const float* rec, uint32_t t, const float* wo, const float* wi, const float* ns, double f[3],
                   double J[3][5])
{
    double r[16];
    for (int i = 0; i < 16; i++) r[i] = rec[i];
    V3<double> WO(wo[0], wo[1], wo[2]), WI(wi[0], wi[1], wi[2]), N(ns[0], ns[1], ns[2]);
    V3<double> v = bsdf64At(r, t, WO, WI, N);
    f[0] = v.x; f[1] = v.y; f[2] = v.z;
    for (int k = 0; k < 5; k++) {
        const double h = 1e-6;
        double rp[16], rm[16];
        for (int i = 0; i < 16; i++) rp[i] = rm[i] = r[i];
        rp[kRepSlots[k]] += h;
        rm[kRepSlots[k]] -= h;
        V3<double> a = bsdf64At(rp, t, WO, WI, N), b = bsdf64At(rm, t, WO, WI, N);
        J[0][k] = (a.x - b.x) / (2 * h);
        J[1][k] = (a.y - b.y) / (2 * h);
        J[2][k] = (a.z - b.z) / (2 * h);
    }
}

// Gradient of delta . C for one recorded sample, in FP64.  Each term of C is
//   (product of the vertex factors before it) x (local factor),
// a vertex factor being f / pdf * cos / rr, with f evaluated at the evaluation point
// for differentiable materials; the gradient is taken factor by factor, with
// df/dtheta from FP64 central differences.  fp32Values = false: f in FP64 (the exact
// gradient of valueSample); true: f as the FP32 estimator computed it, so that the
// result is the gradient of the FP32 estimator (FP32 and FP64 BSDF values differ
// where the BSDF is ill-conditioned, e.g. pow5(1 - cos) for cos near 1).
static void adjointSample(const float* evMat, const float* evEnv, const RepSample& rs, const double delta[3],
                          double* dMat, double* dEnv, bool fp32Values)
{
    const int n = (int)rs.vs.size();
    std::vector<std::array<double, 3>> g(n), gf(n);
    std::vector<std::array<std::array<double, 5>, 3>> dg(n), dgf(n);
    for (int j = 0; j < n; j++) {
        const RepVertex& v = rs.vs[j];
        double f[3], J[3][5] = {};
        if (v.diff) bsdf64(evMat + 16 * (size_t)v.mat, v.type, v.wo, v.wi, v.ns, f, J);
        else for (int c = 0; c < 3; c++) f[c] = v.fConst[c];
        if (fp32Values) for (int c = 0; c < 3; c++) f[c] = v.fE[c];
        for (int c = 0; c < 3; c++) {
            g[j][c] = f[c] / v.pdf * v.cosf;
            gf[j][c] = g[j][c] / v.rr;
            for (int k = 0; k < 5; k++) {
                dg[j][c][k] = J[c][k] / v.pdf * v.cosf;
                dgf[j][c][k] = dg[j][c][k] / v.rr;
            }
        }
    }
    auto envLe64 = [&](const RepTap& t, int c) {
        double x = 0;
        for (int k = 0; k < 4; k++) x += t.w[k] * evEnv[3 * (size_t)t.idx[k] + c];
        return x;
    };
    for (const RepTerm& t : rs.ts) {
        const bool nee = t.kind == 1 || t.kind == 2;
        const bool withG = t.kind == 3 || t.kind == 4;  // the hit vertex's own BSDF factor
        const int m = (t.kind == 0) ? 0 : t.v;           // full factors 0..m-1
        double fn[3] = {0, 0, 0}, Jn[3][5] = {};
        if (nee) {
            if (t.diff) bsdf64(evMat + 16 * (size_t)t.mat, t.type, t.wo, t.wl, t.ns, fn, Jn);
            else for (int c = 0; c < 3; c++) fn[c] = t.f[c];
            if (fp32Values) for (int c = 0; c < 3; c++) fn[c] = t.f[c];
        }
        for (int c = 0; c < 3; c++) {
            // local factor and its derivative with respect to the env radiance of channel c
            double rest = 0, local = 0, dLe = 0;
            switch (t.kind) {
            case 0: local = envLe64(t.tap, c); dLe = 1; break;
            case 1: rest = (double)t.satDot * (double)t.coef[c] * t.weight; local = fn[c] * rest; break;
            case 2:
                rest = (double)t.satDot * envLe64(t.tap, c) / t.envPdf / t.sel * t.weight;
                local = fn[c] * rest;
                dLe = fn[c] * t.satDot / t.envPdf / t.sel * t.weight;
                break;
            case 3: local = envLe64(t.tap, c) * t.weight; dLe = t.weight; break;
            default: local = (double)t.le[c] * t.weight; break;
            }
            const int nf = m + (withG ? 1 : 0);
            auto factor = [&](int i) { return (withG && i == m) ? g[i][c] : gf[i][c]; };
            double P = 1;
            for (int i = 0; i < nf; i++) P *= factor(i);
            const double wc = delta[c];
            for (int i = 0; i < nf; i++) {
                const RepVertex& v = rs.vs[i];
                if (!v.diff) continue;
                double ex = 1;
                for (int l = 0; l < nf; l++) if (l != i) ex *= factor(l);
                for (int k = 0; k < 5; k++) {
                    const double d = (withG && i == m) ? dg[i][c][k] : dgf[i][c][k];
                    dMat[5 * (size_t)v.mat + k] += wc * ex * d * local;
                }
            }
            if (nee && t.diff)
                for (int k = 0; k < 5; k++) dMat[5 * (size_t)t.mat + k] += wc * P * Jn[c][k] * rest;
            if (t.kind == 0 || t.kind == 2 || t.kind == 3)
                for (int q = 0; q < 4; q++)
                    if (t.tap.w[q] != 0) dEnv[3 * (size_t)t.tap.idx[q] + c] += wc * P * dLe * t.tap.w[q];
        }
    }
}

// delta . C for one recorded sample, evaluated in FP64 from the records (the value whose
// gradient adjointSample takes).
static double valueSample(const float* evMat, const float* evEnv, const RepSample& rs, const double delta[3])
{
    const int n = (int)rs.vs.size();
    std::vector<std::array<double, 3>> g(n), gf(n);
    for (int j = 0; j < n; j++) {
        const RepVertex& v = rs.vs[j];
        double f[3], J[3][5];
        if (v.diff) bsdf64(evMat + 16 * (size_t)v.mat, v.type, v.wo, v.wi, v.ns, f, J);
        else for (int c = 0; c < 3; c++) f[c] = v.fConst[c];
        for (int c = 0; c < 3; c++) { g[j][c] = f[c] / v.pdf * v.cosf; gf[j][c] = g[j][c] / v.rr; }
    }
    double sum = 0;
    for (const RepTerm& t : rs.ts) {
        double fn[3] = {0, 0, 0}, Jn[3][5];
        const bool nee = t.kind == 1 || t.kind == 2;
        if (nee) {
            if (t.diff) bsdf64(evMat + 16 * (size_t)t.mat, t.type, t.wo, t.wl, t.ns, fn, Jn);
            else for (int c = 0; c < 3; c++) fn[c] = t.f[c];
        }
        const int m = (t.kind == 0) ? 0 : t.v;
        const bool withG = t.kind == 3 || t.kind == 4;
        for (int c = 0; c < 3; c++) {
            double le = 0;
            if (t.kind == 0 || t.kind == 2 || t.kind == 3)
                for (int k = 0; k < 4; k++) le += t.tap.w[k] * evEnv[3 * (size_t)t.tap.idx[k] + c];
            double local = 0;
            switch (t.kind) {
            case 0: local = le; break;
            case 1: local = fn[c] * ((double)t.satDot * (double)t.coef[c] * t.weight); break;
            case 2: local = fn[c] * ((double)t.satDot * le / t.envPdf / t.sel * t.weight); break;
            case 3: local = le * t.weight; break;
            default: local = (double)t.le[c] * t.weight; break;
            }
            double P = 1;
            for (int i = 0; i < m; i++) P *= gf[i][c];
            if (withG) P *= g[m][c];
            sum += delta[c] * P * local;
        }
    }
    return sum;
}

// The replay estimator over an image region: frame sums (dFrame == null, FP32 like
// renderRegion) or the FP64 gradient of sum(dFrame . frame) into dMat / dEnv.
static void replayRegion(const orc_scene& s, const orc_params& p, const float* evMat, const float* evEnv, int x0,
                         int y0, int w, int h, float* out, const float* dFrame, double* dMat, double* dEnv, Cnt& cn,
                         int rowBegin, int rowEnd, double* loss = nullptr, bool fp32Values = false)
{
    const int samplerType = s.sampler ? 1 : 0;
    orc_scene se = s;
    se.materials = evMat;
    for (int yy = rowBegin; yy < rowEnd; yy++) {
        for (int xx = 0; xx < w; xx++) {
            const int x = x0 + xx, y = y0 + yy;
            float acc[3] = {0, 0, 0};
            double delta[3] = {0, 0, 0};
            if (dFrame)
                for (int c = 0; c < 3; c++) delta[c] = dFrame[3 * ((size_t)yy * w + xx) + c];
            for (int k = 0; k < p.n_passes; k++) {
                const int spp = p.first_pass + k;
                Ctx<float> c(s, p, cn, samplerType), ce(se, p, cn, samplerType);
                c.seedPixel(x, y, spp, (uint32_t)(p.free_counter + k));
                float u0 = c.sample1D(), u1 = c.sample1D(), u2 = c.sample1D(), u3 = c.sample1D();
                V2<float> scr(float(x) / float(s.film_width), float(y) / float(s.film_height));
                V3<float> o, d;
                c.cameraRay(scr, u0, u1, u2, u3, o, d);
                RepSample rs;
                V3<float> r = c.naiveReplay(o, d, ce, evEnv, dFrame ? &rs : nullptr);
                if (hasNan(r)) continue;
                if (loss) {
                    *loss += valueSample(evMat, evEnv, rs, delta);
                } else if (dFrame) {
                    adjointSample(evMat, evEnv, rs, delta, dMat, dEnv, fp32Values);
                } else {
                    const bool fresh = spp == 0;
                    acc[0] = (fresh ? 0.0f : acc[0]) + r.x;
                    acc[1] = (fresh ? 0.0f : acc[1]) + r.y;
                    acc[2] = (fresh ? 0.0f : acc[2]) + r.z;
                }
            }
            if (out) {
                float* o3 = out + 3 * ((size_t)yy * w + xx);
                o3[0] = acc[0]; o3[1] = acc[1]; o3[2] = acc[2];
            }
        }
    }
}

template <class R>
static void lightRange(const orc_scene& s, const orc_params& p, int64_t begin, int64_t end, std::vector<double>& film,
                       Cnt& cn, bool triple)
{
    const int64_t perPass = 1536LL * p.blocks_one_pass;
    const int loops = p.loops_per_pass > 0 ? p.loops_per_pass : 1;
    const R scale = R((float)(s.film_width * s.film_height) / (float)(p.blocks_one_pass * loops * 1536));
    for (int64_t g = begin; g < end; g++) {
        const int k = (int)(g / perPass);
        const uint32_t id = (uint32_t)(g % perPass);
        const int spp = p.first_pass + k;
        Ctx<R> c(s, p, cn, 0);
        const uint32_t mul = triple ? (uint32_t)(perPass * loops) : (uint32_t)perPass;
        c.randSeed = (uint32_t)spp * mul + id + (uint32_t)(p.free_counter + k);
        c.sampleSeed = 0;
        c.sampleOffset = 0;
        c.dim = 0;
        if (triple) { for (int i = 0; i < loops; i++) c.tripleLight(film, scale); }
        else c.lightPath(film);
    }
}

}  // namespace orc

using namespace orc;

extern "C" {

int orc_build_mtbvh(const float* V, int64_t nv, const uint32_t* I, int64_t nt, float* bounds, int32_t* table)
{
    if (!V || !I || nt < 1) return 1;
    for (int64_t i = 0; i < 3 * nt; i++) if ((int64_t)I[i] >= nv) return 1;
    std::vector<float> b;
    std::vector<int32_t> t;
    buildBvh(V, I, nt, b, t);
    std::memcpy(bounds, b.data(), b.size() * 4);
    std::memcpy(table, t.data(), t.size() * 4);
    return 0;
}

int orc_alias_build(const float* pdf, int32_t n, int32_t* alias, float* prob)
{
    for (int i = 0; i < n; i++) prob[i] = pdf[i];
    aliasCore(prob, alias, n, 1, true);
    return 0;
}

float orc_env_alias_build(const float* rgb, int32_t w, int32_t h, int32_t* alias, float* prob)
{
    const size_t st = (size_t)w + 1;
    for (int i = 0; i < h; i++) {
        float sv = std::sin((float)(i + 0.5f) / (float)h * 3.141592653589793f);
        for (int j = 0; j < w; j++) {
            const float* q = rgb + 3 * ((size_t)i * w + j);
            prob[i * st + j] = (0.2126f * q[0] + 0.7152f * q[1] + 0.0722f * q[2]) * sv;
        }
        alias[i * st + w] = 0;
    }
    for (int i = 0; i < h; i++) prob[i * st + w] = aliasCore(prob + i * st, alias + i * st, w, 1, false);
    return aliasCore(prob + w, alias + w, h, (int)st, false);
}

void orc_sobol_matrices(uint32_t* out)
{
    for (int d = 0; d < 256; d++) for (int k = 0; k < 32; k++) out[d * 32 + k] = sobol().col[d][k];
}
uint32_t orc_sobol_sample(uint32_t index, int32_t dim) { return sobol().sample(index, dim); }
uint32_t orc_hash(uint32_t x) { return hashU(x); }
void orc_hash_n(const uint32_t* in, uint32_t* out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) out[i] = hashU(in[i]);
}
// envSampleWi for n 4-tuples of uniforms: direction and pdf (4 floats each) and the
// sampled texel (row, column).
int orc_env_sample(const orc_scene* s, const orc_params* p, const float* u, int64_t n, float* wi_pdf, int32_t* texel)
{
    orc::Cnt cn;
    orc::Ctx<float> c(*s, *p, cn, s->sampler ? 1 : 0);
    for (int64_t i = 0; i < n; i++) {
        float pdf;
        int row, col;
        orc::V3<float> wi = c.envSampleWi(u[4 * i], u[4 * i + 1], u[4 * i + 2], u[4 * i + 3], pdf, &row, &col);
        wi_pdf[4 * i] = wi.x; wi_pdf[4 * i + 1] = wi.y; wi_pdf[4 * i + 2] = wi.z; wi_pdf[4 * i + 3] = pdf;
        texel[2 * i] = row; texel[2 * i + 1] = col;
    }
    return 0;
}
// envPdfLi (light.glsl:175-179) for n directions.
int orc_env_pdf(const orc_scene* s, const orc_params* p, const float* wi, int64_t n, float* pdf)
{
    orc::Cnt cn;
    orc::Ctx<float> c(*s, *p, cn, s->sampler ? 1 : 0);
    for (int64_t i = 0; i < n; i++) pdf[i] = c.envPdfLi(orc::V3<float>(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
    return 0;
}
// texture2DArray(uTextures, vec3(fract(uv) * uvScale, layer)) for n uv pairs (3 floats each out).
int orc_tex_lookup(const orc_scene* s, const orc_params* p, int layer, const float* uv, int64_t n, float* out)
{
    if (!s->tex_rgb8 || layer < 0 || layer >= s->tex_layers) return 1;
    orc::Cnt cn;
    orc::Ctx<float> c(*s, *p, cn, s->sampler ? 1 : 0);
    for (int64_t i = 0; i < n; i++) {
        orc::V3<float> v = c.texAt(layer, orc::V2<float>(uv[2 * i], uv[2 * i + 1]));
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
    return 0;
}
// Test-only estimator variants (kVar* above); 0 restores the reference's estimators.
void orc_set_variant(int v) { gVariant = v; }

// op: 0 sin, 1 cos, 2 atan2(in, in2), 3 log, 4 exp, 5 pow(in, in2)
void orc_fmath(int op, const float* in, const float* in2, float* out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        float s, c;
        switch (op) {
        case 0: sincosF(in[i], &s, &c); out[i] = s; break;
        case 1: sincosF(in[i], &s, &c); out[i] = c; break;
        case 2: out[i] = atan2F(in[i], in2[i]); break;
        case 3: out[i] = logF(in[i]); break;
        case 4: out[i] = expF(in[i]); break;
        default: out[i] = powF(in[i], in2[i]); break;
        }
    }
}

// The FP64 routines, for the tests that pin them (op as orc_fmath).
void orc_dmath(int op, const double* in, const double* in2, double* out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        double s, c;
        switch (op) {
        case 0: sincosD(in[i], &s, &c); out[i] = s; break;
        case 1: sincosD(in[i], &s, &c); out[i] = c; break;
        case 2: out[i] = atan2D(in[i], in2[i]); break;
        case 3: out[i] = logD(in[i]); break;
        case 4: out[i] = expD(in[i]); break;
        default: out[i] = powD(in[i], in2[i]); break;
        }
    }
}

static int threadsOf(int n) { return n > 0 ? n : (int)std::max(1u, std::thread::hardware_concurrency()); }

// The box test alone (intersection.glsl:226-329) on n (box, ray) pairs: boxes n x 6
// {pMin, pMax}, rays n x 6 {o, d}; writes hit and, where hit, the entry distance.
int orc_box_hit(const float* boxes, const float* rays, int64_t n, uint8_t* hit, float* tmin)
{
    orc_scene sc = {};
    sc.bounds = boxes;
    orc_params p = {};
    Cnt cn;
    Ctx<float> c(sc, p, cn, 0);
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + 6 * i;
        float t = 0.0f;
        hit[i] = c.boxHit((int)i, V3<float>(r[0], r[1], r[2]), V3<float>(r[3], r[4], r[5]), t) ? 1 : 0;
        tmin[i] = hit[i] ? t : 0.0f;
    }
    return 0;
}

int orc_trace_closest(const orc_scene* s, const float* rays, int64_t n, int32_t* prim, float* t, uint32_t* bin,
                      uint64_t* counters, int nthreads)
{
    const int T = threadsOf(nthreads);
    orc_params p = {};
    std::vector<Cnt> cs(T);
    std::vector<std::thread> th;
    for (int ti = 0; ti < T; ti++) {
        th.emplace_back([&, ti] {
            Ctx<float> c(*s, p, cs[ti], 0);
            for (int64_t i = ti; i < n; i += T) {
                V3<float> o(rays[i], rays[n + i], rays[2 * n + i]), d(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
                float dist;
                int id = c.bvhHit(o, d, dist);
                prim[i] = id;
                t[i] = dist;
                bin[i] = 0xffffffffu;
                if (id >= 0) {
                    V3<float> a = c.vert(s->indices[3 * id]), b = c.vert(s->indices[3 * id + 1]), cc = c.vert(s->indices[3 * id + 2]);
                    float tt = 0.0f, u = 0.0f, v = 0.0f, det = 1.0f;
                    Ctx<float>::triHit(a, b, cc, o, d, tt, &u, &v, &det);
                    int bu = (int)(16.0f * (u / det)), bv = (int)(16.0f * (v / det));
                    bu = bu > 15 ? 15 : (bu < 0 ? 0 : bu);
                    bv = bv > 15 ? 15 : (bv < 0 ? 0 : bv);
                    bin[i] = (uint32_t)(bu * 16 + bv);
                }
            }
        });
    }
    for (auto& x : th) x.join();
    if (counters) {
        for (auto& c : cs) { counters[0] += c.bytes; counters[1] += c.rows; counters[2] += c.tests; counters[3] += c.trav; }
    }
    return 0;
}

// Camera-path integrators (naive = 0, triple camera pass = 2, global queue = 3) over a
// pixel region:
// out[h][w][3] = per-pixel sums over the passes.  fp64 selects the double instance.
int orc_render_pixels(const orc_scene* s, const orc_params* p, int x0, int y0, int w, int h, float* out,
                      uint64_t* counters, int nthreads, int fp64)
{
    if (!s || !p || !out || w < 1 || h < 1) return 1;
    const int T = std::min(threadsOf(nthreads), h);
    std::vector<Cnt> cs(T);
    std::vector<std::thread> th;
    std::atomic<int> next(0);
    for (int ti = 0; ti < T; ti++) {
        th.emplace_back([&, ti] {
            for (;;) {
                int row = next.fetch_add(1);
                if (row >= h) break;
                if (fp64) renderRegion<double>(*s, *p, x0, y0, w, h, out, cs[ti], row, row + 1);
                else renderRegion<float>(*s, *p, x0, y0, w, h, out, cs[ti], row, row + 1);
            }
        });
    }
    for (auto& x : th) x.join();
    if (counters)
        for (auto& c : cs) { counters[0] += c.bytes; counters[1] += c.rows; counters[2] += c.tests; counters[3] += c.trav; }
    return 0;
}

// post_proc.glsl:1-59 and the RGB8 read-back of captureImage (Application.cpp:370-378):
// sum * scale, clamp(0, 1e30), tone map (0 none, 1 filmic, 2 ACES), pow(1/2.2); bytes
// are round(255 * clamp(x, 0, 1)) with rows written top-down.
static float filmicCalc(float x)
{
    const float A = 0.22f, B = 0.3f, C = 0.1f, D = 0.2f, E = 0.01f, F = 0.3f;
    const float bc = B * C, de = D * E, df = D * F, ef = E / F;
    return (x * (x * A + bc) + de) / (x * (x * A + B) + df) - ef;
}
int orc_post_process(const float* frame, int w, int h, float scale, int tone, float* out, uint8_t* rgb8)
{
    if (!frame || w < 1 || h < 1) return 1;
    const float g = 1.0f / 2.2f;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const float* f = frame + 4 * ((size_t)y * w + x);
            float m[3];
            for (int c = 0; c < 3; c++) {
                float v = gmin(gmax(f[c] * scale, 0.0f), 1e30f);
                if (tone == 1) v = filmicCalc(v * 1.6f) / filmicCalc(11.2f);
                else if (tone == 2) v = (v * (v * 2.51f + 0.03f)) / (v * (v * 2.43f + 0.59f) + 0.14f);
                m[c] = powF(v, g);
            }
            if (out) {
                float* o = out + 4 * ((size_t)y * w + x);
                o[0] = m[0]; o[1] = m[1]; o[2] = m[2]; o[3] = 1.0f;
            }
            if (rgb8)
                for (int c = 0; c < 3; c++)
                    rgb8[3 * ((size_t)(h - 1 - y) * w + x) + c] =
                        std::isnan(m[c]) ? 0 : (uint8_t)(int)(gmin(gmax(m[c], 0.0f), 1.0f) * 255.0f + 0.5f);
        }
    return 0;
}

// Differentiable naive estimator (no counterpart in the reference): the replay
// estimator's frame sums over a region (FP32), and the FP64 gradient of
// sum(d_frame . frame) w.r.t. the material parameters (n x 5: baseColor rgb, raw
// roughness, metallic) and the FP32 env texels (h x w x 3).  eval_mat holds the
// material records at the evaluation point, eval_env the env texels (FP32 RGB).
// fp32_values selects the BSDF values of the products (see adjointSample).
int orc_render_replay(const orc_scene* s, const orc_params* p, const float* eval_mat, const float* eval_env, int x0,
                      int y0, int w, int h, float* out, int nthreads)
{
    if (!s || !p || !out || !eval_mat || !eval_env || w < 1 || h < 1) return 1;
    const int T = std::min(threadsOf(nthreads), h);
    std::vector<Cnt> cs(T);
    std::vector<std::thread> th;
    std::atomic<int> next(0);
    for (int ti = 0; ti < T; ti++)
        th.emplace_back([&, ti] {
            for (int row; (row = next.fetch_add(1)) < h;)
                replayRegion(*s, *p, eval_mat, eval_env, x0, y0, w, h, out, nullptr, nullptr, nullptr, cs[ti], row,
                             row + 1);
        });
    for (auto& x : th) x.join();
    return 0;
}

int orc_adjoint(const orc_scene* s, const orc_params* p, const float* eval_mat, const float* eval_env,
                const float* d_frame, int x0, int y0, int w, int h, double* d_mat, double* d_env, int nthreads,
                int fp32_values)
{
    if (!s || !p || !eval_mat || !eval_env || !d_frame || !d_mat || !d_env || w < 1 || h < 1) return 1;
    const int T = std::min(threadsOf(nthreads), h);
    const size_t nm = (size_t)s->n_materials * 5, ne = (size_t)s->env_width * s->env_height * 3;
    std::vector<std::vector<double>> gm(T, std::vector<double>(nm, 0.0)), ge(T, std::vector<double>(ne, 0.0));
    std::vector<Cnt> cs(T);
    std::vector<std::thread> th;
    std::atomic<int> next(0);
    for (int ti = 0; ti < T; ti++)
        th.emplace_back([&, ti] {
            for (int row; (row = next.fetch_add(1)) < h;)
                replayRegion(*s, *p, eval_mat, eval_env, x0, y0, w, h, nullptr, d_frame, gm[ti].data(),
                             ge[ti].data(), cs[ti], row, row + 1, nullptr, fp32_values != 0);
        });
    for (auto& x : th) x.join();
    for (size_t i = 0; i < nm; i++) { d_mat[i] = 0; for (int ti = 0; ti < T; ti++) d_mat[i] += gm[ti][i]; }
    for (size_t i = 0; i < ne; i++) { d_env[i] = 0; for (int ti = 0; ti < T; ti++) d_env[i] += ge[ti][i]; }
    return 0;
}

// sum(d_frame . frame) of the replay estimator evaluated in FP64 from the recorded
// paths (the function orc_adjoint differentiates).
double orc_replay_loss64(const orc_scene* s, const orc_params* p, const float* eval_mat, const float* eval_env,
                         const float* d_frame, int x0, int y0, int w, int h, int nthreads)
{
    if (!s || !p || !eval_mat || !eval_env || !d_frame || w < 1 || h < 1) return 0.0;
    const int T = std::min(threadsOf(nthreads), h);
    std::vector<double> part(T, 0.0);
    std::vector<Cnt> cs(T);
    std::vector<std::thread> th;
    std::atomic<int> next(0);
    for (int ti = 0; ti < T; ti++)
        th.emplace_back([&, ti] {
            for (int row; (row = next.fetch_add(1)) < h;)
                replayRegion(*s, *p, eval_mat, eval_env, x0, y0, w, h, nullptr, d_frame, nullptr, nullptr, cs[ti],
                             row, row + 1, &part[ti]);
        });
    for (auto& x : th) x.join();
    double sum = 0;
    for (double v : part) sum += v;
    return sum;
}

// Light-path integrators (light = 1, triple light pass = 2): splats of the paths
// [path_begin, path_end) of the pass range into film (3W x H, accumulated in FP64).
int orc_render_light(const orc_scene* s, const orc_params* p, int64_t path_begin, int64_t path_end, double* film,
                     uint64_t* counters, int nthreads, int fp64)
{
    if (!s || !p || !film) return 1;
    const int T = threadsOf(nthreads);
    const size_t fs = (size_t)s->film_width * s->film_height * 3;
    std::vector<std::vector<double>> films(T, std::vector<double>(fs, 0.0));
    std::vector<Cnt> cs(T);
    std::vector<std::thread> th;
    const bool triple = p->integrator == 2;
    const int64_t n = path_end - path_begin;
    for (int ti = 0; ti < T; ti++) {
        th.emplace_back([&, ti] {
            int64_t b = path_begin + n * ti / T, e = path_begin + n * (ti + 1) / T;
            if (fp64) lightRange<double>(*s, *p, b, e, films[ti], cs[ti], triple);
            else lightRange<float>(*s, *p, b, e, films[ti], cs[ti], triple);
        });
    }
    for (auto& x : th) x.join();
    for (int ti = 0; ti < T; ti++)
        for (size_t i = 0; i < fs; i++) film[i] += films[ti][i];
    if (counters)
        for (auto& c : cs) { counters[0] += c.bytes; counters[1] += c.rows; counters[2] += c.tests; counters[3] += c.trav; }
    return 0;
}

// BSDF evaluation for unit tests: mat[16] is one material record, type from mat[13]
// bits.  out = {f.rgb, pdf}; sample_out = {wi.xyz, pdf, f.rgb, flag}.
int orc_bsdf_eval(const float* mat, const float* wo, const float* wi, const float* n, int mode, float* out)
{
    orc_scene s = {};
    s.materials = mat;
    orc_params p = {};
    Cnt cn;
    Ctx<float> c(s, p, cn, 0);
    uint32_t t = c.matType(0);
    auto mp = c.loadMat(t, 0);
    V3<float> WO(wo[0], wo[1], wo[2]), WI(wi[0], wi[1], wi[2]), N(n[0], n[1], n[2]);
    V3<float> f = Ctx<float>::evalF(t, mp, WO, WI, N, (uint32_t)mode);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
    out[3] = Ctx<float>::evalPdf(t, mp, WO, WI, N);
    return 0;
}
// FP64 evaluation of the principled (type 1) or Lambertian (type 0) BSDF from a
// material record given in double, the loader's roughness remap done in double:
// the reference point for derivatives taken by central differences.
int orc_bsdf_eval_f64(const double* rec, int type, const double* wo, const double* wi, const double* n, double* out)
{
    using D = Ctx<double>;
    D::P p;
    p.base = V3<double>(rec[0], rec[1], rec[2]);
    if (type == 1) {
        p.rough = mixs(0.0134, 1.0, rec[3]);
        p.ss = rec[4]; p.metal = rec[5]; p.spec = rec[6]; p.specTint = rec[7];
        p.sheen = rec[8]; p.sheenTint = rec[9]; p.cc = rec[10]; p.ccGloss = rec[11];
    }
    V3<double> WO(wo[0], wo[1], wo[2]), WI(wi[0], wi[1], wi[2]), N(n[0], n[1], n[2]);
    V3<double> f = D::evalF(type == 1 ? 1u : 0u, p, WO, WI, N, 0);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
    return 0;
}

int orc_bsdf_sample(const float* mat, const float* wo, const float* n, int mode, const float* u, uint32_t rng_seed,
                    float* out)
{
    orc_scene s = {};
    s.materials = mat;
    orc_params p = {};
    Cnt cn;
    Ctx<float> c(s, p, cn, 0);
    c.randSeed = rng_seed;
    uint32_t t = c.matType(0);
    auto mp = c.loadMat(t, 0);
    V3<float> WO(wo[0], wo[1], wo[2]), N(n[0], n[1], n[2]);
    auto r = c.sampleF(t, mp, N, WO, (uint32_t)mode, V3<float>(u[0], u[1], u[2]));
    out[0] = r.wi.x; out[1] = r.wi.y; out[2] = r.wi.z; out[3] = r.pdf;
    out[4] = r.f.x; out[5] = r.f.y; out[6] = r.f.z;
    std::memcpy(out + 7, &r.flag, 4);
    return 0;
}

float orc_fresnel(float cos_ti, float eta) { return Ctx<float>::fresnel(cos_ti, eta); }
float orc_ggx(float cos_theta, float alpha) { return Ctx<float>::ggx(cos_theta, alpha); }
float orc_concentric(float u, float v, float* out)
{
    V2<float> r = concentric(V2<float>(u, v));
    out[0] = r.x;
    out[1] = r.y;
    return 0.0f;
}

// Camera round trip: raster position of the ray generated for (uv, u) (camera.glsl:62-91).
int orc_camera_roundtrip(const orc_scene* s, float u, float v, float* uv_out)
{
    orc_params p = {};
    Cnt cn;
    Ctx<float> c(*s, p, cn, 0);
    V3<float> o, d;
    c.cameraRay(V2<float>(u, v), 0.0f, 0.0f, 0.0f, 0.0f, o, d);
    V2<float> r = c.rasterPos(o, d);
    uv_out[0] = r.x;
    uv_out[1] = r.y;
    return 0;
}

}  // extern "C"
