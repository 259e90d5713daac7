// engine/colour.h -- automatic colour control of the hvd-core C++ API
// (hvd-core/src/engine/colour.h).
#ifndef HVD_AMD_ENGINE_COLOUR_H_
#define HVD_AMD_ENGINE_COLOUR_H_

#include <array>
#include <cstdint>

namespace hvd {

constexpr float kIreBlack = 7.5F;
constexpr float kIreWhite = 100.0F;

// Chroma gain that brings the measured burst to its nominal amplitude,
// clamped to [0.5, 2] (colour.cpp:20-27); the one-argument form uses the
// NTSC nominal of 20 IRE.
float AccGain(float median_burst_amplitude_ire);
float AccGain(float median_burst_amplitude_ire, float nominal_ire);

// One pixel of IRE-domain (Y, U, V) to 16-bit RGB (colour.h:36-38,
// colour.cpp:29-42): Y over black_ire .. 100, U and V over 7.5 .. 100, the
// BT.601 matrix, clipped to [0, 1], x 65535 rounded.  A scalar helper of the
// API; the plane-level exports convert on the GPU.
std::array<uint16_t, 3> YuvToRgb16(float y, float u, float v, float black_ire);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_COLOUR_H_
