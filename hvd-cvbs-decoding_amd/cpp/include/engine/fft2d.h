// engine/fft2d.h -- hvd-core's FFT context type (hvd-core/src/engine/fft2d.h).
// The transforms run inside the GPU library's own kernels, whose twiddle
// tables live in the GPU context; this object only keeps the API's shape
// (HolographicInit and DecodeFieldWindow take an Fft2d*, which may be null).
#ifndef HVD_AMD_ENGINE_FFT2D_H_
#define HVD_AMD_ENGINE_FFT2D_H_

namespace hvd {

class Fft2d {
 public:
  Fft2d() = default;
  // Host FFT threads: no effect on the GPU path.
  void SetThreadCount(int n) { threads_ = n; }
  int thread_count() const { return threads_; }

 private:
  int threads_ = 0;
};

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_FFT2D_H_
