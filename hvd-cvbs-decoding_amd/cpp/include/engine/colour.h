// engine/colour.h -- automatic colour control of the hvd-core C++ API
// (hvd-core/src/engine/colour.h).
#ifndef HVD_AMD_ENGINE_COLOUR_H_
#define HVD_AMD_ENGINE_COLOUR_H_

namespace hvd {

constexpr float kIreBlack = 7.5F;
constexpr float kIreWhite = 100.0F;

// Chroma gain that brings the measured burst to its nominal amplitude,
// clamped to [0.5, 2] (colour.cpp:20-27); the one-argument form uses the
// NTSC nominal of 20 IRE.
float AccGain(float median_burst_amplitude_ire);
float AccGain(float median_burst_amplitude_ire, float nominal_ire);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_COLOUR_H_
