// engine/fft2d.h -- hvd-core's FFT context (hvd-core/src/engine/fft2d.h): the
// numpy-convention 2-D DFT, forward unnormalised with sign -1, inverse with
// sign +1 divided by h w (fft2d.cpp:88-118).  The transforms run on the GPU
// library's LDS FFT (hvd_gpu_fft2d), whose twiddle tables the library builds
// per call; this object owns no plans.  Planes up to 1024 x 4096; failures
// return an empty plane (GpuLastError() says why).
#ifndef HVD_AMD_ENGINE_FFT2D_H_
#define HVD_AMD_ENGINE_FFT2D_H_

#include "engine/plane.h"

namespace hvd {

class Fft2d {
 public:
  Fft2d() = default;
  Fft2d(const Fft2d&) = delete;
  Fft2d& operator=(const Fft2d&) = delete;

  ComplexPlane Forward(const ComplexPlane& in);
  ComplexPlane Inverse(const ComplexPlane& in);

  // Host FFT threads: no effect on the GPU path.
  void SetThreadCount(int n) { threads_ = n; }
  int thread_count() const { return threads_; }

 private:
  ComplexPlane Run(const ComplexPlane& in, bool inverse);
  int threads_ = 0;
};

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_FFT2D_H_
