// engine/lockin.h -- colour-burst lock-in of the hvd-core C++ API
// (hvd-core/src/engine/lockin.h), computed by the GPU library.
#ifndef HVD_AMD_ENGINE_LOCKIN_H_
#define HVD_AMD_ENGINE_LOCKIN_H_

#include <cstdint>
#include <vector>

#include "engine/ntsc_geometry.h"
#include "engine/plane.h"

namespace hvd {

// Per-line carrier phase at sample 0 of a field in IRE (lockin.cpp:101-186);
// one entry per plane row.  The plane's own size is the field's size.
std::vector<float> BurstLockinPhase(const Plane& field_ire, const FieldGeometry& g);

// Burst amplitude in IRE for ACC (lockin.cpp:188-255).
float BurstAmplitudeIre(const Plane& field_ire, const FieldGeometry& g);

// PAL-family lock-in: phase plus per-line V-switch sense (lockin.cpp:257-349).
struct PalBurstLockin {
  std::vector<float> theta;
  std::vector<int8_t> vswitch;
};
PalBurstLockin BurstLockinPhasePal(const Plane& field_ire, const FieldGeometry& g);

// (diag(a) + lam L) x = a d with Neumann ends (lockin.cpp:65-99).
std::vector<float> TridiagSmooth(const std::vector<float>& d, const std::vector<float>& a, float lam);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_LOCKIN_H_
