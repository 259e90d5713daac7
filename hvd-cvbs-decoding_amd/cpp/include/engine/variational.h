// engine/variational.h -- the IRLS + conjugate-gradient Y/C arbitration of
// the hvd-core C++ API (hvd-core/src/engine/variational.h), solved on the GPU.
#ifndef HVD_AMD_ENGINE_VARIATIONAL_H_
#define HVD_AMD_ENGINE_VARIATIONAL_H_

#include <vector>

#include "engine/hvd_config.h"
#include "engine/plane.h"

namespace hvd {

struct RefineResult {
  Plane luma;           // Y = S - Re[chi c]
  ComplexPlane chroma;  // chi = V - iU
};

// A motion-compensated temporal neighbour (variational.h; 3-D decoding).
struct NeighborTerm {
  Plane composite;
  ComplexPlane carrier;
  Plane confidence;
};

// The solve from chi0 (variational.cpp:198-556): the 2-D objective and, with
// `neighbors` non-empty and cfg.temporal_strength > 0, the motion-compensated
// temporal data terms at strength cfg.temporal_strength and robust gate
// cfg.temporal_eps (which must then be positive: hvd-core's sequence
// resolves it before the solve).  Errors return an empty result.
RefineResult VariationalRefine(const Plane& s, const ComplexPlane& carrier, const ComplexPlane& chi0,
                               const HvdConfig& cfg, const std::vector<NeighborTerm>& neighbors = {});

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_VARIATIONAL_H_
