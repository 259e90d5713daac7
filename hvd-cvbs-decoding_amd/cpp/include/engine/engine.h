// engine/engine.h -- the hvd-core decoder facade (hvd-core/src/engine/engine.h)
// over the MI355X library: one HvdEngine owns one GPU context (created on the
// calling thread's current HIP device at first use) and, like hvd-core's,
// must not be used from two threads at once.
#ifndef HVD_AMD_ENGINE_ENGINE_H_
#define HVD_AMD_ENGINE_ENGINE_H_

#include <memory>
#include <vector>

#include "engine/fft2d.h"
#include "engine/hvd_config.h"
#include "engine/ntsc_geometry.h"
#include "engine/plane.h"
#include "engine/sequence.h"

namespace hvd {

namespace gpu {
class Session;
}

// One field in IRE (field_height x field_width) and whether its lines are the
// frame's top (even) rows.
struct FieldInput {
  Plane samples;
  bool is_first_field = true;
};

// Woven-frame decode over the active picture (2 * active lines rows):
// luma + chroma == composite element for element.
struct FrameYc {
  Plane luma;
  Plane chroma;                // S - Y (Re[chi c] for DecodeChromaOnly)
  Plane composite;
  ComplexPlane chroma_phasor;  // chi = V - iU
  ComplexPlane carrier;
  float acc_gain = 1.0F;
};

class HvdEngine {
 public:
  HvdEngine();
  ~HvdEngine();

  HvdEngine(const HvdEngine&) = delete;
  HvdEngine& operator=(const HvdEngine&) = delete;

  // 2-D woven-frame decode of two fields (engine.cpp:131-196).
  FrameYc DecodeFrame(const FieldInput& first, const FieldInput& second, const FieldGeometry& g,
                      const HvdConfig& cfg);

  // Chroma-only decode of an already separated chroma channel: lock-in and
  // the holographic crop, no arbitration (engine.cpp:198-236).
  FrameYc DecodeChromaOnly(const FieldInput& first, const FieldInput& second, const FieldGeometry& g,
                           const HvdConfig& cfg);

  // DecodeFieldWindow on this engine's GPU context.
  std::vector<DecodedField> DecodeSequenceWindow(const std::vector<FieldObs>& fields, const FieldGeometry& g,
                                                 const HvdConfig& cfg, SequenceDiagnostics* diag = nullptr);

  // Host FFT threads: kept for the API, no effect on the GPU path.
  void SetFftThreads(int n);

  // The GPU session this engine decodes on (library internal).
  gpu::Session& session();

 private:
  std::unique_ptr<Fft2d> fft_;
  std::unique_ptr<gpu::Session> gpu_;
};

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_ENGINE_H_
