// engine/sequence.h -- the field-granularity decode of the hvd-core C++ API
// (hvd-core/src/engine/sequence.h), 2-D path: every field of a window is
// initialised and solved in one GPU batch.
#ifndef HVD_AMD_ENGINE_SEQUENCE_H_
#define HVD_AMD_ENGINE_SEQUENCE_H_

#include <vector>

#include "engine/hvd_config.h"
#include "engine/ntsc_geometry.h"
#include "engine/plane.h"

namespace hvd {

class Fft2d;
struct FieldInput;

// One field's observation: the active composite and its carrier.
struct FieldObs {
  Plane s;               // composite active picture (IRE)
  ComplexPlane carrier;  // exp(i phi) on the same grid
  int parity = 0;        // 0 top field, 1 bottom field
};

// Lock-in on the full field, active crop, chroma-phase trim and carrier
// (sequence.cpp:229-268).
FieldObs PrepareFieldObs(const FieldInput& field, const FieldGeometry& g, const HvdConfig& cfg, int parity);

struct DecodedField {
  Plane luma;           // Y = S - Re[chi c]
  ComplexPlane chroma;  // chi = V - iU
};

struct SequenceDiagnostics {
  float resolved_strength = 0.0F;
  float ambiguity_ire = 0.0F;
  float sigma_ire = 0.0F;
  float temporal_eps = 0.0F;
  float nr_eps = 0.0F;
};

// The window decode (sequence.cpp:423-1114), 2-D: temporal strength < 0, or
// 0 with no more fields than the same-parity stride (2 NTSC, 4 PAL), or
// cg_iterations == 0.  Anything that would engage 3-D returns an empty vector.
std::vector<DecodedField> DecodeFieldWindow(const std::vector<FieldObs>& fields, const FieldGeometry& g,
                                            const HvdConfig& cfg, Fft2d* fft, SequenceDiagnostics* diag = nullptr);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_SEQUENCE_H_
