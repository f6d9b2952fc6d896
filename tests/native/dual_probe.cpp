// Host build of the product's forward-mode BSDF derivatives (zillumgl_amd/csrc/zgl_dual.h).
#include <cstdint>

#include "../../zillumgl_amd/csrc/zgl_dual.h"

extern "C" void probe_dual_bsdf(const float* rec, int type, const float* wo, const float* wi, const float* n,
                                float* val, float* jac)
{
    using namespace zgl;
    const DParam p = dparamOf(rec);
    const vec3 WO = v3(wo[0], wo[1], wo[2]), WI = v3(wi[0], wi[1], wi[2]), N = v3(n[0], n[1], n[2]);
    const DV3 f = (type == 1) ? dPrincipledBRDF(WO, WI, N, p) : dLambertian(p);
    const D5* c[3] = {&f.x, &f.y, &f.z};
    for (int i = 0; i < 3; i++) {
        val[i] = c[i]->v;
        for (int k = 0; k < NumMatGrad; k++) jac[i * NumMatGrad + k] = c[i]->d[k];
    }
}
