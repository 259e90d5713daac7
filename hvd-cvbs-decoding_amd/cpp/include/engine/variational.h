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

// The anchor of the joint pass (variational.h:80-90): a synthesized reference
// (Y_a, chi_a) and its per-pixel weight.
struct AnchorTerm {
  Plane luma;
  ComplexPlane chroma;
  Plane weight;
};

// The pass-2+ solver with Y and chi both free (variational.h:92-109,
// variational.cpp:558-891): the data, neighbour (strength
// cfg.temporal_strength, robust scale cfg.temporal_eps, which must be
// positive when neighbours are given) and anchor (weight cfg.nr_anchor) terms
// and the spatial priors; y0 and chi0 start the solve.  On the GPU
// (hvd_gpu_variational_refine_joint); failure: empty planes.
RefineResult VariationalRefineJoint(const Plane& s, const ComplexPlane& carrier, const Plane& y0,
                                    const ComplexPlane& chi0, const HvdConfig& cfg,
                                    const std::vector<NeighborTerm>& neighbors, const AnchorTerm* anchor);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_VARIATIONAL_H_
