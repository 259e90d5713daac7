// engine/sequence.h -- the field-granularity decode of the hvd-core C++ API
// (hvd-core/src/engine/sequence.h): a window's inits run as one GPU batch,
// then the library's window driver (2-D, and 3-D with its motion, gates and
// anchored passes) on the device.
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

// The window decode (sequence.cpp:270-281, :423-1114): HolographicInit per
// field, then noise self-calibration, the adaptive or fixed temporal
// strength, the Gauss-Seidel passes with pair motion, the gated
// motion-compensated neighbour equations, the PSI init, the SynthReference
// anchor and joint pass, and the output-fidelity luma.  All fields must
// share one shape when the window can engage 3-D.  `fft` is not used (the
// inits run on the device).  Failure: an empty vector (GpuLastError()).
std::vector<DecodedField> DecodeFieldWindow(const std::vector<FieldObs>& fields, const FieldGeometry& g,
                                            const HvdConfig& cfg, Fft2d* fft, SequenceDiagnostics* diag = nullptr);

// Drizzle stacking of fields j0 and j0 + 1 onto a (2 lines scale) x width
// grid (sequence.h:105-130, sequence.cpp:283-421): window neighbours matched
// by EstimateMotion (cfg.mc_tile, cfg.mc_search) within 2 cfg.nr_radius
// fields, Geman-McClure weights at cfg.nr_eps (3 when unset), blended with
// the woven interpolation where coverage is thin.  On the GPU; failure:
// empty planes (GpuLastError()).
struct DrizzleResult {
  Plane luma;
  ComplexPlane chroma;
};
DrizzleResult DrizzleFrame(int j0, const std::vector<Plane>& Ys, const std::vector<ComplexPlane>& chis,
                           const std::vector<int>& parities, const HvdConfig& cfg, int scale = 2);

// The same driver from the caller's inits (sequence.h:104-135, :139-142):
// inits index-aligned with fields; an empty luma means Y0 = S - Re[chi0 c].
std::vector<DecodedField> DecodeFieldWindowWithInits(const std::vector<FieldObs>& fields,
                                                     const std::vector<DecodedField>& inits, const FieldGeometry& g,
                                                     const HvdConfig& cfg, SequenceDiagnostics* diag = nullptr);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_SEQUENCE_H_
