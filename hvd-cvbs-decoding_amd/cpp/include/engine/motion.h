// engine/motion.h -- block-matching motion estimation of the hvd-core C++ API
// (hvd-core/src/engine/motion.h), run on the GPU in hvd-core's order of
// operations (bit-identical vectors and confidences).
#ifndef HVD_AMD_ENGINE_MOTION_H_
#define HVD_AMD_ENGINE_MOTION_H_

#include "engine/plane.h"

namespace hvd {

struct MotionField {
  Plane dy;          // per tile, field lines
  Plane dx;          // per tile, samples
  Plane confidence;  // per tile, [0, 1]
  int tile = 32;
};

// Motion of y_cur relative to y_ref per tile (motion.cpp:397-488).
MotionField EstimateMotion(const Plane& y_ref, const Plane& y_cur, int tile = 32, int search = 8);

// (2r + 1)-box mean with truncated edge windows (motion.cpp:21-55).
Plane BoxBlur(Plane a, int r);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_MOTION_H_
