// zgl_f64.h -- the double instance of the device rendering library, in namespace
// zgl::f64: the same source as the rendering kernels' FP32 library (zgl_math_body.inc,
// zgl_box_body.inc, zgl_device_body.inc) with real = double.  Scene data, literals and
// random numbers stay the FP32 ones and are widened where they enter a computation;
// sin/cos/atan2/log/exp/pow are the library's double routines.  This mirrors the CPU
// restatement's double mode, and it is what the converged FP64 reference renderer
// (zgl_reference.hip) is built from.
#pragma once

#include "zgl_device.h"

namespace zgl {
namespace f64 {

typedef double real;
#include "zgl_math_body.inc"
#include "zgl_box_body.inc"
#include "zgl_device_body.inc"

}  // namespace f64
}  // namespace zgl
