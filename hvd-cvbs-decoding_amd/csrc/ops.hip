// ops.hip -- the forward differences and their exact adjoints of hvd-core's
// engine/gradients.h, which the C++ layer exposes on their own.
#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

// gradients.h:26-381: d = op(a) for op in hvd_gpu_diff_op order (Dx, Dy,
// DxT, DyT, D1, D1T, D2, D2T), Neumann zeros where the difference leaves the
// plane, `wt` (Dx, Dy, D1, D2 only, may be null) multiplying the difference
// (the *WeightedInto forms).  nc = 1 real, 2 interleaved complex (each
// component on its own, as std::complex<float> arithmetic does).
__global__ void k_difference(int op, int nc, const float* a, const float* wt, int h, int w, float* d) {
#pragma clang fp contract(off)
  const size_t n = (size_t)h * w;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int y = (int)(i / w), x = (int)(i % w);
  for (int c = 0; c < nc; ++c) {
    auto A = [&](int yy, int xx) { return a[((size_t)yy * w + xx) * nc + c]; };
    float v = 0.0f;
    bool weighted = false;
    switch (op) {
      case 0:  // Dx
        if (x + 1 < w) {
          v = A(y, x + 1) - A(y, x);
          weighted = true;
        }
        break;
      case 1:  // Dy
        if (y + 1 < h) {
          v = A(y + 1, x) - A(y, x);
          weighted = true;
        }
        break;
      case 2:  // DxT
        if (w >= 2) v = x == 0 ? -A(y, 0) : (x == w - 1 ? A(y, w - 2) : A(y, x - 1) - A(y, x));
        break;
      case 3:  // DyT
        if (h >= 2) v = y == 0 ? -A(0, x) : (y == h - 1 ? A(h - 2, x) : A(y - 1, x) - A(y, x));
        break;
      case 4:  // D1
        if (y + 1 < h && x + 1 < w) {
          v = A(y + 1, x + 1) - A(y, x);
          weighted = true;
        }
        break;
      case 5:  // D1T
        if (y + 1 < h && x + 1 < w) v = v - A(y, x);
        if (y >= 1 && x >= 1) v = v + A(y - 1, x - 1);
        break;
      case 6:  // D2
        if (y + 1 < h && x >= 1) {
          v = A(y + 1, x - 1) - A(y, x);
          weighted = true;
        }
        break;
      default:  // D2T
        if (y + 1 < h && x >= 1) v = v - A(y, x);
        if (y >= 1 && x + 1 < w) v = v + A(y - 1, x + 1);
        break;
    }
    if (weighted && wt) v = wt[i] * v;
    d[i * nc + c] = v;
  }
}

}  // namespace hvdk
