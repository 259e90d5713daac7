// woven.h -- the woven-frame decode shared by HvdEngine::DecodeFrame /
// DecodeChromaOnly and the frame-buffer bridge, with every plane kept on the
// device (engine.cpp:53-236).
#ifndef HVD_AMD_SRC_WOVEN_H_
#define HVD_AMD_SRC_WOVEN_H_

#include "engine/hvd_config.h"
#include "engine/ntsc_geometry.h"
#include "gpu_session.h"

namespace hvd {
namespace gpu {

// Two full fields on the device, field_lines x field_width IRE each, and
// whether each carries the frame's first (top) field.
struct DeviceFields {
  const float* first = nullptr;
  const float* second = nullptr;
  bool first_is_first = true;
  bool second_is_first = false;
  int field_lines = 0;
  int field_width = 0;
};

// Woven active picture (2 * active lines rows) and the decode's planes.
struct WovenResult {
  int rows = 0, width = 0;
  DevBuf<float> s, y, chi, chroma, carrier;
  float acc_gain = 1.0F;
};

// chroma_only: lock-in + holographic crop, chroma = Re[chi c] (DecodeChromaOnly);
// otherwise init + solve, chroma = S - Y (DecodeFrame).
WovenResult WovenDecode(Session& ss, const DeviceFields& f, const FieldGeometry& g, const HvdConfig& cfg,
                        bool chroma_only);

}  // namespace gpu
}  // namespace hvd

#endif  // HVD_AMD_SRC_WOVEN_H_
