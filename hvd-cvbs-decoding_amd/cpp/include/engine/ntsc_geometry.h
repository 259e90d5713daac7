// engine/ntsc_geometry.h -- field geometry, level conversion and carrier
// construction of the hvd-core C++ API (hvd-core/src/engine/ntsc_geometry.h).
// Names, members and defaults are hvd-core's, so a FieldGeometry filled by
// hvd-core's callers means the same thing here.
#ifndef HVD_AMD_ENGINE_NTSC_GEOMETRY_H_
#define HVD_AMD_ENGINE_NTSC_GEOMETRY_H_

#include <cmath>
#include <cstdint>
#include <vector>

#include "engine/plane.h"

namespace hvd {

constexpr double kFscNtsc = 315.0e6 / 88.0;
constexpr double kFs4Fsc = 4.0 * kFscNtsc;
constexpr double kFscPal = 4433618.75;
constexpr double kFs4FscPal = 4.0 * kFscPal;
constexpr double kFscPalM = (909.0 / 4.0) * (525.0 * 30000.0 / 1001.0);
constexpr double kFs4FscPalM = 4.0 * kFscPalM;

enum class VideoStandard { kNtsc, kPal, kPalM };

// One field's sampling grid (ntsc_geometry.h:142-267).
struct FieldGeometry {
  int field_width = 910;
  int field_height = 263;
  int active_video_start = 134;
  int active_video_end = 894;
  int colour_burst_start = 78;
  int colour_burst_end = 110;
  int first_active_field_line = 21;
  int last_active_field_line = 0;  // 0 => field_height
  double sample_rate = kFs4Fsc;
  VideoStandard standard = VideoStandard::kNtsc;
  double subcarrier_hz = 0.0;      // 0 => the standard's nominal fsc

  bool uses_vswitch() const { return standard == VideoStandard::kPal || standard == VideoStandard::kPalM; }
  double fsc() const {
    if (subcarrier_hz > 0.0) return subcarrier_hz;
    return standard == VideoStandard::kPal ? kFscPal : (standard == VideoStandard::kPalM ? kFscPalM : kFscNtsc);
  }
  // Carrier phase advance per sample: a quarter turn on the 4fsc grid.
  double phase_per_sample() const {
    constexpr double kPiD = 3.14159265358979323846;
    return subcarrier_hz <= 0.0 ? 0.5 * kPiD : 2.0 * kPiD * subcarrier_hz / sample_rate;
  }
  // Carrier phase advance per line of the field.
  double line_advance() const {
    constexpr double kPiD = 3.14159265358979323846;
    if (subcarrier_hz <= 0.0) {
      if (standard == VideoStandard::kPal) return 2.0 * kPiD * 0.7516;
      if (standard == VideoStandard::kPalM) return 0.5 * kPiD;
      return kPiD;
    }
    const double cycles = subcarrier_hz * static_cast<double>(field_width) / sample_rate;
    return 2.0 * kPiD * (cycles - std::floor(cycles));
  }
  float nominal_burst_ire() const { return standard == VideoStandard::kPal ? 21.43F : 20.0F; }
  int active_width() const { return active_video_end - active_video_start; }
  int last_active_line() const { return last_active_field_line != 0 ? last_active_field_line : field_height; }
  int active_lines() const { return last_active_line() - first_active_field_line; }
};

// Level conversion in the sample domain (ntsc_geometry.h:295-309).
inline float CodesPerIre(float blanking_level, float white_level) { return (white_level - blanking_level) / 100.0F; }
inline float SampleToIre(float sample, float blanking_level, float white_level) {
  return (sample - blanking_level) / CodesPerIre(blanking_level, white_level);
}
inline float IreToSample(float ire, float blanking_level, float white_level) {
  return blanking_level + ire * CodesPerIre(blanking_level, white_level);
}

// Robust per-field noise (IRE) from the stride-4 horizontal second difference
// (ntsc_geometry.h:278-283, ntsc_geometry.cpp:49-71), computed on the GPU;
// 0 for planes of 8 columns or fewer (and on failure, see GpuLastError()).
float EstimateNoiseIre(const Plane& s);

// Explicit carrier planes, theta.size() x active_width (computed on the GPU).
ComplexPlane MakeCarrier(const std::vector<float>& theta, const FieldGeometry& g);
ComplexPlane MakeCarrierPal(const std::vector<float>& theta, const std::vector<int8_t>& vswitch,
                            const FieldGeometry& g);

// Chroma phase trim applied to the burst-locked line phases
// (ntsc_geometry.h:265-276): theta += -deg * pi / 180, with the sign flipped
// on V-switched lines so chi rotates uniformly on PAL-family signals.
inline void ApplyChromaPhase(float chroma_phase_deg, const FieldGeometry& g, const std::vector<int8_t>& vswitch,
                             std::vector<float>* theta) {
  if (chroma_phase_deg == 0.0F || theta == nullptr) return;
  const float offset = -chroma_phase_deg * 3.14159265358979323846F / 180.0F;
  const bool by_parity = g.uses_vswitch() && vswitch.size() == theta->size();
  for (size_t i = 0; i < theta->size(); ++i) (*theta)[i] += (by_parity && vswitch[i] < 0) ? -offset : offset;
}

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_NTSC_GEOMETRY_H_
