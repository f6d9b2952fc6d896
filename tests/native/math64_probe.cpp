// Host build of the product's FP64 routines (zillumgl_amd/csrc/zgl_math.h: sincosF64,
// atan2F64, logF64, expF64, powF64, used by the double instance of the device library)
// for comparison with the oracle's double mode (oracle/zgl_oracle.cpp orc_dmath).
#include <cstdint>

#include "../../zillumgl_amd/csrc/zgl_math.h"

extern "C" void probe_dmath(int op, const double* in, const double* in2, double* out, int64_t n)
{
    using namespace zgl;
    for (int64_t i = 0; i < n; i++) {
        double s, c;
        switch (op) {
        case 0: sincosF64(in[i], &s, &c); out[i] = s; break;
        case 1: sincosF64(in[i], &s, &c); out[i] = c; break;
        case 2: out[i] = atan2F64(in[i], in2[i]); break;
        case 3: out[i] = logF64(in[i]); break;
        case 4: out[i] = expF64(in[i]); break;
        default: out[i] = powF64(in[i], in2[i]); break;
        }
    }
}
