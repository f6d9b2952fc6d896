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
namespace f32 {
#include "zgl_box_body.inc"
}  // namespace f32
}  // namespace zgl
