// engine/motion.h -- block-matching motion estimation of the hvd-core C++ API
// (hvd-core/src/engine/motion.h), run on the GPU in hvd-core's order of
// operations (bit-identical vectors and confidences).
#ifndef HVD_AMD_ENGINE_MOTION_H_
#define HVD_AMD_ENGINE_MOTION_H_

#include <vector>

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

// Audit of a predicted motion field (motion.h:66-75, motion.cpp:490-525):
// per tile the rounded prediction (pdy, pdx: tile grid) against the zero
// vector, one SSD each on the r = 2 blurred planes; the winner (zero on
// ties, the prediction at its float value) with its confidence.
MotionField VerifyMotion(const Plane& y_ref, const Plane& y_cur, const Plane& pdy, const Plane& pdx, int tile = 32);

// One pairwise measurement of a multi-offset trajectory fit (motion.h:77-87):
// the signed temporal offset, the half-line parity shift h = (p_k - p_j) / 2.
struct OffsetMotion {
  int offset = 0;
  float parity_shift = 0.0F;
  MotionField field;
};

// Per-tile velocity across the offsets (motion.h:89-95): the median of the
// confident (conf > 0.15) (dy - h) / o and dx / o, 0 where none qualifies.
void FitTrajectory(const std::vector<OffsetMotion>& measurements, Plane* vy, Plane* vx);

// Consensus snap onto the fitted trajectory (motion.h:97-106): where at least
// three offsets agree within 1.5 px, the agreeing ones take o v + h.
void TrajectorySnap(std::vector<OffsetMotion>* measurements);

// (2r + 1)-box mean with truncated edge windows (motion.cpp:21-55).
Plane BoxBlur(Plane a, int r);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_MOTION_H_
