// zgl_post.h -- display transform of a frame (post_proc.glsl:1-59) and the 8-bit
// conversion of its read-back (Texture2D::readFromDevice, GL_RGB / GL_UNSIGNED_BYTE).
// Host and device.
#pragma once

#include "zgl_math.h"

namespace zgl {

constexpr int ToneNone = 0, ToneFilmic = 1, ToneACES = 2;

ZGL_HD vec3 filmicCurve(vec3 x)
{
    const float A = 0.22f, B = 0.3f, C = 0.1f, D = 0.2f, E = 0.01f, F = 0.3f;
    const float BC = B * C, DE = D * E, DF = D * F, EF = E / F;
    return (x * (x * A + BC) + DE) / (x * (x * A + B) + DF) - EF;
}
ZGL_HD vec3 toneFilmic(vec3 c) { return filmicCurve(c * 1.6f) / filmicCurve(v3(11.2f)); }
ZGL_HD vec3 toneACES(vec3 c) { return (c * (c * 2.51f + 0.03f)) / (c * (c * 2.43f + 0.59f) + 0.14f); }

// sum * scale, clamped to [0, 1e30], tone mapped, raised to 1/2.2.
ZGL_HD vec3 displayColor(vec3 sum, float scale, int toneMapper)
{
    vec3 c = sum * scale;
    c = v3(fclampG(c.x, 0.0f, 1e30f), fclampG(c.y, 0.0f, 1e30f), fclampG(c.z, 0.0f, 1e30f));
    vec3 m = c;
    if (toneMapper == ToneFilmic) m = toneFilmic(c);
    else if (toneMapper == ToneACES) m = toneACES(c);
    const float g = 1.0f / 2.2f;
    return v3(powZ(m.x, g), powZ(m.y, g), powZ(m.z, g));
}

// Normalised unsigned byte of a float channel: clamp to [0, 1], round(x * 255); NaN
// gives 0 (GL's float-to-normalised conversion).
ZGL_HD unsigned char toUnorm8(float x)
{
    if (fisnan(x)) return 0;
    const float c = fclampG(x, 0.0f, 1.0f);
    return (unsigned char)(int)(c * 255.0f + 0.5f);
}

}  // namespace zgl
