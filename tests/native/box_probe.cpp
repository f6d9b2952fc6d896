// Host build of the product's box test (zillumgl_amd/csrc/zgl_box.h) for comparison
// with the oracle's test (oracle/zgl_oracle.cpp orc_box_hit).
#include <cstdint>

#include "../../zillumgl_amd/csrc/zgl_box.h"

extern "C" void probe_box_hit(const float* boxes, const float* rays, int64_t n, int general, uint8_t* hit,
                              float* tmin, uint8_t* special)
{
    using namespace zgl;
    for (int64_t i = 0; i < n; i++) {
        const float* b = boxes + 6 * i;
        const float* r = rays + 6 * i;
        const RayPre p = prepareRay(v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]));
        float t = 0.0f;
        const bool h = general ? boxHitGeneral(v3(b[0], b[1], b[2]), v3(b[3], b[4], b[5]), p, t)
                               : boxHit(v3(b[0], b[1], b[2]), v3(b[3], b[4], b[5]), p, t);
        hit[i] = h ? 1 : 0;
        tmin[i] = h ? t : 0.0f;
        special[i] = p.special ? 1 : 0;
    }
}
