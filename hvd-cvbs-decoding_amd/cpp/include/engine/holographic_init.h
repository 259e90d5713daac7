// engine/holographic_init.h -- the FFT holographic chroma init of the
// hvd-core C++ API (hvd-core/src/engine/holographic_init.h), computed by the
// GPU library's exact-size LDS FFTs.
#ifndef HVD_AMD_ENGINE_HOLOGRAPHIC_INIT_H_
#define HVD_AMD_ENGINE_HOLOGRAPHIC_INIT_H_

#include "engine/fft2d.h"
#include "engine/hvd_config.h"
#include "engine/ntsc_geometry.h"
#include "engine/plane.h"

namespace hvd {

struct HoloInit {
  Plane luma;           // Y0 (IRE)
  ComplexPlane chroma;  // chi0 = V - iU (IRE)
};

// chi0 and Y0 = S - Re[chi0 c] from the composite S and its carrier
// (holographic_init.cpp:200-317).  `fft` may be null.
HoloInit HolographicInit(const Plane& s, const ComplexPlane& carrier, const FieldGeometry& g, const HvdConfig& cfg,
                         Fft2d* fft);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_HOLOGRAPHIC_INIT_H_
