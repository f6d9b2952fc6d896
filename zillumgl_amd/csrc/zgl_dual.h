// zgl_dual.h -- forward-mode derivatives of the differentiable BSDFs, host and device.
//
// The path-replay adjoint needs, at every vertex, the BSDF value f and its Jacobian
// with respect to the vertex material's five parameters
//     theta = (baseColor.r, baseColor.g, baseColor.b, roughness, metallic)
// where roughness is the raw material-record value (the loader remaps it with
// mix(0.0134, 1, r), material_loader.glsl:32) and metallic is record slot 5.  D5 is a
// dual number carrying those five tangents.  Every function here repeats the float
// evaluation of the corresponding BSDF (material.glsl:68-83 Lambertian, :357-555
// principled BRDF) operation by operation, so the value part equals the float
// evaluation bit for bit.  The tangents are the derivatives of the BSDF: those of the
// float program, except in the specular lobe, where the float GGX term loses its digits
// near a smooth lobe's peak and the tangents are formed in FP64 (dPrincipledMetal).  The
// host build is compared with the oracle's evaluation (values) and with FP64 central
// differences (tangents).
#pragma once

#include "zgl_math.h"

namespace zgl {

constexpr int NumMatGrad = 5;  // baseColor rgb, roughness, metallic

struct D5 {
    float v;
    float d[NumMatGrad];
};

ZGL_HD D5 dconst(float v)
{
    D5 r;
    r.v = v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = 0.0f;
    return r;
}
ZGL_HD D5 dvar(float v, int k)
{
    D5 r = dconst(v);
    r.d[k] = 1.0f;
    return r;
}
ZGL_HD D5 operator+(D5 a, D5 b)
{
    D5 r;
    r.v = a.v + b.v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
ZGL_HD D5 operator-(D5 a, D5 b)
{
    D5 r;
    r.v = a.v - b.v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
ZGL_HD D5 operator*(D5 a, D5 b)
{
    D5 r;
    r.v = a.v * b.v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
ZGL_HD D5 operator/(D5 a, D5 b)
{
    D5 r;
    r.v = a.v / b.v;
    const float inv = 1.0f / b.v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
    return r;
}
ZGL_HD D5 operator+(D5 a, float b)
{
    a.v = a.v + b;
    return a;
}
ZGL_HD D5 operator+(float a, D5 b)
{
    b.v = a + b.v;
    return b;
}
ZGL_HD D5 operator-(D5 a, float b)
{
    a.v = a.v - b;
    return a;
}
ZGL_HD D5 operator-(float a, D5 b)
{
    D5 r;
    r.v = a - b.v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = -b.d[k];
    return r;
}
ZGL_HD D5 operator*(D5 a, float b)
{
    D5 r;
    r.v = a.v * b;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = a.d[k] * b;
    return r;
}
ZGL_HD D5 operator*(float a, D5 b)
{
    D5 r;
    r.v = a * b.v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = a * b.d[k];
    return r;
}
ZGL_HD D5 operator/(D5 a, float b)
{
    D5 r;
    r.v = a.v / b;
    const float inv = 1.0f / b;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = a.d[k] * inv;
    return r;
}
ZGL_HD D5 operator/(float a, D5 b)
{
    D5 r;
    r.v = a / b.v;
    const float s = -r.v / b.v;
    for (int k = 0; k < NumMatGrad; k++) r.d[k] = s * b.d[k];
    return r;
}

struct DV3 {
    D5 x, y, z;
};
ZGL_HD DV3 dv3(D5 s) { return DV3{s, s, s}; }
ZGL_HD DV3 dv3c(vec3 c) { return DV3{dconst(c.x), dconst(c.y), dconst(c.z)}; }
ZGL_HD vec3 value(const DV3& a) { return v3(a.x.v, a.y.v, a.z.v); }
ZGL_HD DV3 operator+(const DV3& a, const DV3& b) { return DV3{a.x + b.x, a.y + b.y, a.z + b.z}; }
ZGL_HD DV3 operator-(const DV3& a, const DV3& b) { return DV3{a.x - b.x, a.y - b.y, a.z - b.z}; }
ZGL_HD DV3 operator*(const DV3& a, const DV3& b) { return DV3{a.x * b.x, a.y * b.y, a.z * b.z}; }
ZGL_HD DV3 operator*(const DV3& a, D5 s) { return DV3{a.x * s, a.y * s, a.z * s}; }
ZGL_HD DV3 operator*(const DV3& a, float s) { return DV3{a.x * s, a.y * s, a.z * s}; }
ZGL_HD DV3 operator/(const DV3& a, D5 s) { return DV3{a.x / s, a.y / s, a.z / s}; }
ZGL_HD DV3 operator/(const DV3& a, float s) { return DV3{a.x / s, a.y / s, a.z / s}; }
ZGL_HD DV3 operator*(vec3 a, D5 s) { return DV3{a.x * s, a.y * s, a.z * s}; }
ZGL_HD DV3 operator-(vec3 a, const DV3& b) { return DV3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// mix(x, y, a) = x * (1 - a) + y * a, with the operand kinds of each call site
ZGL_HD DV3 dmix(const DV3& x, const DV3& y, float a) { return x * (1.0f - a) + y * a; }
ZGL_HD DV3 dmix(const DV3& x, const DV3& y, D5 a) { return x * (1.0f - a) + y * a; }
ZGL_HD DV3 dmix1(const DV3& y, float a) { return v3(1.0f) * dconst(1.0f - a) + y * a; }
ZGL_HD D5 dmixf(float x, float y, D5 a) { return x * (1.0f - a) + y * a; }
ZGL_HD D5 dmix1s(D5 y, float a) { return dconst(1.0f * (1.0f - a)) + y * a; }

// Parameters of a differentiable BSDF with their tangents seeded.
struct DParam {
    DV3 baseColor;
    D5 roughness;  // after the loader's mix(0.0134, 1, r)
    D5 metallic;
    float subsurface, specular, specularTint, sheen, sheenTint, clearcoat, clearcoatGloss;
};

// A material record (4 texels of 4 floats, material_loader.glsl:20-36) as a DParam.
ZGL_HD DParam dparamOf(const float* rec)
{
    DParam p;
    p.baseColor = DV3{dvar(rec[0], 0), dvar(rec[1], 1), dvar(rec[2], 2)};
    p.roughness = dmixf(0.0134f, 1.0f, dvar(rec[3], 3));
    p.metallic = dvar(rec[5], 4);
    p.subsurface = rec[4];
    p.specular = rec[6];
    p.specularTint = rec[7];
    p.sheen = rec[8];
    p.sheenTint = rec[9];
    p.clearcoat = rec[10];
    p.clearcoatGloss = rec[11];
    return p;
}

// Scalar pieces that do not depend on the parameters (microfacet.glsl:4-28, :94-101).
ZGL_HD float dSchlickW(float cosTheta) { return pow5(1.0f - cosTheta); }
ZGL_HD float dSchlickGc(float cosTheta, float alpha)
{
    float k = alpha * 0.5f;
    return cosTheta / (cosTheta * (1.0f - k) + k);
}
ZGL_HD float dGtr1(float cosTheta, float alpha)
{
    float a2 = alpha * alpha;
    return (a2 - 1.0f) / (2.0f * Pi * logZ(alpha) * (1.0f + (a2 - 1.0f) * cosTheta * cosTheta));
}

// Parameter-dependent pieces.
ZGL_HD DV3 dSchlickF(float cosTheta, const DV3& F0) { return F0 + (v3(1.0f) - F0) * dconst(pow5(1.0f - cosTheta)); }
ZGL_HD D5 dSchlickG(float cosTheta, D5 alpha)
{
    D5 k = alpha * 0.5f;
    return cosTheta / (cosTheta * (1.0f - k) + k);
}
ZGL_HD D5 dSmithG(vec3 n, vec3 wo, vec3 wi, D5 alpha)
{
    return dSchlickG(absDot(n, wo), alpha) * dSchlickG(absDot(n, wi), alpha);
}
// ggx's float program (microfacet.glsl:4-12): the values of the differentiable BSDF.
ZGL_HD float ggxValue(float cosTheta, float a2)
{
    if (cosTheta < 1e-6f) return 0.0f;
    float den = cosTheta * cosTheta * (a2 - 1.0f) + 1.0f;
    den = den * den * Pi;
    return a2 / den;
}
// GGX D and dD/d(a2) in FP64 at the FP32 directions, for the tangents:
// D = a2 / (pi q^2), dD/da2 = (s2 - a2 c2) / (pi q^3), q = s2 + a2 c2, c = n.h, s2 = 1 - c^2.
// Near a smooth lobe's peak (1 - c and a2 both ~1e-6) the float program keeps few digits
// of q: the FP32 half vector's cos carries an absolute error of ~6e-8 and
// c^2 (a2 - 1) + 1 cancels, so D is off by up to ~10% there; a tangent built on that
// value is off by as much, while the FP64 form is exact to the inputs' precision.
ZGL_HD void ggx64(vec3 wo, vec3 wi, vec3 n, float cosTheta, float a2f, double& D, double& dD)
{
    D = dD = 0.0;
    if (cosTheta < 1e-6f) return;
    const double hx = (double)wo.x + (double)wi.x, hy = (double)wo.y + (double)wi.y,
                 hz = (double)wo.z + (double)wi.z;
    const double c = ((double)n.x * hx + (double)n.y * hy + (double)n.z * hz) / sqrt(hx * hx + hy * hy + hz * hz);
    const double a2 = a2f, c2 = c * c, s2 = 1.0 - c2, q = s2 + a2 * c2;
    const double pi = 3.14159265358979323846;
    D = a2 / (pi * q * q);
    dD = (s2 - a2 * c2) / (pi * q * q * q);
}

// The specular lobe: values are the float program's bits; tangents are formed in FP64
// from the FP64 D and dD/d(a2) of ggx64 and the float Fresnel and Smith factors.
ZGL_HD DV3 dPrincipledMetal(vec3 wo, vec3 wi, vec3 n, const DV3& fm0, D5 alpha)
{
    float cosWo = satDot(n, wo), cosWi = satDot(n, wi);
    vec3 h = normalize(wo + wi);
    if (cosWo < 1e-10f || cosWi < 1e-10f) return dv3c(v3(0.0f));
    DV3 fm = dSchlickF(absDot(h, wo), fm0);
    const float cosH = dot(n, h);
    const D5 a2 = alpha * alpha;
    const float dm = ggxValue(cosH, a2.v);
    D5 gm = dSmithG(n, wo, wi, alpha);
    float denom = 4.0f * cosWi * cosWo;
    if (denom < 1e-7f) return dv3c(v3(0.0f));
    DV3 r = fm * dconst(dm) * gm / denom;
    double D, dD;
    ggx64(wo, wi, n, cosH, a2.v, D, dD);
    const double G = gm.v;
    D5* out[3] = {&r.x, &r.y, &r.z};
    const D5* f[3] = {&fm.x, &fm.y, &fm.z};
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < NumMatGrad; k++)
            out[c]->d[k] = (float)(((double)f[c]->d[k] * D * G +
                                    (double)f[c]->v * (dD * (double)a2.d[k] * G + D * (double)gm.d[k])) /
                                   (double)denom);
    return r;
}
ZGL_HD DV3 dPrincipledClearcoat(vec3 wo, vec3 wi, vec3 n, const DV3& baseColor, float alpha)
{
    float cosWo = satDot(n, wo), cosWi = satDot(n, wi);
    vec3 h = normalize(wo + wi);
    if (cosWo < 1e-6f || cosWi < 1e-6f) return dv3c(v3(0.0f));
    DV3 fc = dSchlickF(absDot(h, wo), baseColor);
    float dc = dGtr1(satDot(n, h), alpha);
    float gc = dSchlickGc(absDot(n, wo), 0.25f) * dSchlickGc(absDot(n, wi), 0.25f);
    float denom = 4.0f * cosWi * cosWo;
    if (denom < 1e-7f) return dv3c(v3(0.0f));
    return fc * dc * gc / denom;
}
ZGL_HD DV3 dPrincipledDiffuse(vec3 wo, vec3 wi, vec3 n, const DV3& baseColor, float subsurface, D5 roughness)
{
    float cosWo = satDot(n, wo), cosWi = satDot(n, wi);
    if (cosWo < 1e-10f || cosWi < 1e-10f) return dv3c(v3(0.0f));
    vec3 h = normalize(wo + wi);
    float hCosWi = dot(h, wi);
    float hCosWi2 = hCosWi * hCosWi;
    float fi = dSchlickW(cosWi), fo = dSchlickW(cosWo);
    D5 fd90 = 0.5f + 2.0f * roughness * hCosWi2;
    D5 fd = dmix1s(fd90, fi) * dmix1s(fd90, fo);
    DV3 baseDiffuse = baseColor * fd * PiInv;
    D5 fss90 = roughness * hCosWi2;
    D5 fss = dmix1s(fss90, fi) * dmix1s(fss90, fo);
    DV3 ss = baseColor * PiInv * 1.25f * (fss * (1.0f / (cosWi + cosWo) - 0.5f) + 0.5f);
    return dmix(baseDiffuse, ss, subsurface);
}
ZGL_HD DV3 dPrincipledTint(const DV3& baseColor)
{
    D5 lum = baseColor.x * 0.299f + baseColor.y * 0.587f + baseColor.z * 0.114f;
    return lum.v > 0.0f ? baseColor / lum : dv3c(v3(1.0f));
}
// principledBRDF (material.glsl:462-479) with tangents.
ZGL_HD DV3 dPrincipledBRDF(vec3 wo, vec3 wi, vec3 n, const DParam& p)
{
    DV3 res = dv3c(v3(0.0f));
    D5 alpha = p.roughness * p.roughness;
    float clearcoatAlpha = mixf(0.1f, 0.001f, p.clearcoatGloss);
    DV3 tintColor = dPrincipledTint(p.baseColor);
    DV3 fm0 = dmix1(tintColor, p.specularTint) * (0.08f * p.specular);
    fm0 = dmix(fm0, p.baseColor, p.metallic);
    float hCosWi = dot(normalize(wo + wi), wi);
    res = res + dPrincipledDiffuse(wo, wi, n, p.baseColor, p.subsurface, p.roughness) * (1.0f - p.metallic);
    res = res + dPrincipledMetal(wo, wi, n, fm0, alpha);
    res = res + dPrincipledClearcoat(wo, wi, n, p.baseColor, clearcoatAlpha) * p.clearcoat * 0.25f;
    res = res + dmix1(tintColor, p.sheenTint) * dSchlickW(hCosWi) * p.sheen * (dot(n, wi) < 0.0f ? 0.0f : 1.0f);
    return res;
}
// lambertian (material.glsl:68-71) with tangents.
ZGL_HD DV3 dLambertian(const DParam& p) { return p.baseColor * PiInv; }

}  // namespace zgl
