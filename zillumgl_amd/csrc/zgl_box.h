// zgl_box.h -- the reference's ray/box test (intersection.glsl:226-329), host and device.
//
// The reference tests a box in one of five ways, picked by the ray direction: a ray
// within 1e-6 of a coordinate axis (|d.i| > 1 - 1e-6) uses that axis' slab plus an
// open containment test of the origin on the other two axes; a ray with a component
// below 1e-6 in magnitude uses the two other slabs when their overlap test passes; any
// other ray uses the three-slab overlap test.  Here every quantity is computed without
// branches and the case is chosen with selects, so a wave evaluates one instruction
// stream whatever its rays' directions.  boxHitGeneral is the three-slab case alone,
// exact for rays whose RayPre::special is false.  The host build of this header is
// compared with the oracle's test on random and degenerate rays and boxes.
#pragma once

#include "zgl_math.h"

namespace zgl {

// Per-ray constants of the box test: 1/d and the case selection depend only on the
// ray, so computing them once gives the same bits as the per-box evaluation.
struct RayPre {
    vec3 o, d, dInv;
    int mode;      // 0 general, 1/2/3: |d.x|/|d.y|/|d.z| > 1 - 1e-6
    bool nx, ny, nz;
    bool special;  // mode != 0, a component below 1e-6 in magnitude, or a non-finite o or d
    int table;     // MTBVH table: cubemapFace(-d)
};

ZGL_HD bool finite3(vec3 v)
{
    const float big = 3.4028235e38f;  // FLT_MAX; false for +-inf and NaN
    return fabsG(v.x) <= big && fabsG(v.y) <= big && fabsG(v.z) <= big;
}

ZGL_HD RayPre prepareRay(vec3 o, vec3 d)
{
    const float eps = 1e-6f;
    RayPre p;
    p.o = o;
    p.d = d;
    p.dInv = v3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const float one = 1.0f - eps;
    p.mode = (fabsG(d.x) > one) ? 1 : (fabsG(d.y) > one) ? 2 : (fabsG(d.z) > one) ? 3 : 0;
    p.nx = fabsG(d.x) < eps;
    p.ny = fabsG(d.y) < eps;
    p.nz = fabsG(d.z) < eps;
    p.special = p.mode != 0 || p.nx || p.ny || p.nz || !finite3(o) || !finite3(d);
    p.table = cubemapFace(-d);
    return p;
}

// Three-slab case only, valid when !p.special and the box bounds are finite (the
// library rejects scenes with non-finite bounds).  Then 1/d is finite and nonzero and
// o is finite, so the slab distances hold no NaN, and IEEE min/max (one instruction
// each, with min3/max3) equal the reference's min/max, (y < x) ? y : x, except for the
// sign of a zero result, which only ever enters comparisons.  Writes the entry
// distance; returns hit.
ZGL_HD bool boxHitGeneral(vec3 pMin, vec3 pMax, const RayPre& p, float& tMin)
{
    const vec3 ta = (pMin - p.o) * p.dInv;
    const vec3 tb = (pMax - p.o) * p.dInv;
    const vec3 mn = v3(__builtin_fminf(ta.x, tb.x), __builtin_fminf(ta.y, tb.y), __builtin_fminf(ta.z, tb.z));
    const vec3 mx = v3(__builtin_fmaxf(ta.x, tb.x), __builtin_fmaxf(ta.y, tb.y), __builtin_fmaxf(ta.z, tb.z));
    const vec3 dt = mx - mn;
    const float tyz = mx.z - mn.y, tzx = mx.x - mn.z, txy = mx.y - mn.x;
    const bool overlap = (dt.y + dt.z > tyz) & (dt.z + dt.x > tzx) & (dt.x + dt.y > txy);
    const float lo = __builtin_fmaxf(__builtin_fmaxf(mn.x, mn.y), mn.z);
    const float hi = __builtin_fminf(__builtin_fminf(mx.x, mx.y), mx.z);
    tMin = lo;
    return overlap & (hi >= 0.0f) & (hi >= lo);
}

// All five cases, selected per ray without branches.
ZGL_HD bool boxHit(vec3 pMin, vec3 pMax, const RayPre& p, float& tMin)
{
    const vec3 o = p.o;
    const vec3 ta = (pMin - o) * p.dInv;
    const vec3 tb = (pMax - o) * p.dInv;
    const vec3 mn = vmin(ta, tb);
    const vec3 mx = vmax(ta, tb);
    const vec3 dt = mx - mn;
    const float tyz = mx.z - mn.y, tzx = mx.x - mn.z, txy = mx.y - mn.x;
    const bool cx = dt.y + dt.z > tyz, cy = dt.z + dt.x > tzx, cz = dt.x + dt.y > txy;
    // a component below 1e-6: the first such axis (x, y, z order) whose overlap passes
    const bool useX = p.nx & cx;
    const bool useY = !useX & p.ny & cy;
    const bool useZ = !useX & !useY & p.nz & cz;
    float lo = fmaxG(fmaxG(mn.x, mn.y), mn.z);
    float hi = fminG(fminG(mx.x, mx.y), mx.z);
    bool ok = cx & cy & cz;
    lo = useZ ? fmaxG(mn.x, mn.y) : lo;
    hi = useZ ? fminG(mx.x, mx.y) : hi;
    lo = useY ? fmaxG(mn.z, mn.x) : lo;
    hi = useY ? fminG(mx.z, mx.x) : hi;
    lo = useX ? fmaxG(mn.y, mn.z) : lo;
    hi = useX ? fminG(mx.y, mx.z) : hi;
    ok = ok | useX | useY | useZ;
    // near-axis rays: one slab plus open containment of the origin on the other axes
    const bool inX = (o.x > pMin.x) & (o.x < pMax.x);
    const bool inY = (o.y > pMin.y) & (o.y < pMax.y);
    const bool inZ = (o.z > pMin.z) & (o.z < pMax.z);
    const int m = p.mode;
    lo = (m == 1) ? mn.x : (m == 2) ? mn.y : (m == 3) ? mn.z : lo;
    hi = (m == 1) ? mx.x : (m == 2) ? mx.y : (m == 3) ? mx.z : hi;
    ok = (m == 1) ? (inY & inZ) : (m == 2) ? (inX & inZ) : (m == 3) ? (inX & inY) : ok;
    tMin = lo;
    return ok & (hi >= 0.0f) & (hi >= lo);
}

}  // namespace zgl
