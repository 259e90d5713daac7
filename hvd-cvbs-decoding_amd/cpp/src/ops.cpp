// ops.cpp -- the component operators of the C++ API without a solver of
// their own: gradients.h's differences, EstimateNoiseIre and Fft2d, each a
// round trip through the GPU library.
#include <exception>

#include <algorithm>

#include "engine/colour.h"
#include "engine/fft2d.h"
#include "engine/gradients.h"
#include "engine/ntsc_geometry.h"
#include "gpu_session.h"

namespace hvd {

using gpu::DevBuf;
using gpu::Session;

namespace grad_detail {

bool Run(int op, bool is_complex, const float* a, const float* weight, int h, int w, float* d) {
  try {
    const size_t n = (size_t)h * (size_t)w * (is_complex ? 2 : 1);
    if (n == 0) return true;
    Session& s = gpu::ThreadSession();
    DevBuf<float> da = gpu::Upload(a, n), dd(n), dw;
    if (weight) dw = gpu::Upload(weight, (size_t)h * w);
    s.check(hvd_gpu_difference(s.ctx(), op, is_complex ? 1 : 0, da.get(), weight ? dw.get() : nullptr, h, w,
                               dd.get(), nullptr),
            "hvd_gpu_difference");
    dd.download(d, n);
    return true;
  } catch (const std::exception& e) {
    gpu::Report("gradients", e);
    return false;
  }
}

}  // namespace grad_detail

float EstimateNoiseIre(const Plane& s) {
  try {
    if (s.width() <= 8 || s.height() <= 0) return 0.0F;
    Session& ss = gpu::ThreadSession();
    DevBuf<float> d = gpu::Upload(s.data(), s.size());
    float sigma = 0.0F;
    ss.check(hvd_gpu_estimate_noise_ire(ss.ctx(), d.get(), s.height(), s.width(), &sigma, nullptr),
             "hvd_gpu_estimate_noise_ire");
    return sigma;
  } catch (const std::exception& e) {
    gpu::Report("EstimateNoiseIre", e);
    return 0.0F;
  }
}

std::array<uint16_t, 3> YuvToRgb16(float y, float u, float v, float black_ire) {
  // BT.601 YUV -> RGB, the matrix of hvd-core's colour.cpp:12-17 (the same
  // one the woven RGB export kernel uses)
  static constexpr float m[3][3] = {{1.0F, 0.0F, 1.13983F}, {1.0F, -0.39465F, -0.58060F}, {1.0F, 2.03211F, 0.0F}};
  const float yn = (y - black_ire) / (kIreWhite - black_ire);
  const float un = u / (kIreWhite - kIreBlack);
  const float vn = v / (kIreWhite - kIreBlack);
  std::array<uint16_t, 3> rgb{};
  for (int c = 0; c < 3; ++c) {
    const float lin = m[c][0] * yn + m[c][1] * un + m[c][2] * vn;
    rgb[c] = static_cast<uint16_t>(std::clamp(lin, 0.0F, 1.0F) * 65535.0F + 0.5F);
  }
  return rgb;
}

ComplexPlane Fft2d::Forward(const ComplexPlane& in) { return Run(in, false); }
ComplexPlane Fft2d::Inverse(const ComplexPlane& in) { return Run(in, true); }

ComplexPlane Fft2d::Run(const ComplexPlane& in, bool inverse) {
  try {
    if (in.empty()) return in;
    Session& s = gpu::ThreadSession();
    const size_t n = in.size() * 2;
    DevBuf<float> d = gpu::Upload(reinterpret_cast<const float*>(in.data()), n), out(n);
    s.check(hvd_gpu_fft2d(s.ctx(), d.get(), in.height(), in.width(), inverse ? 1 : 0, out.get(), nullptr),
            "hvd_gpu_fft2d");
    ComplexPlane r(in.height(), in.width());
    out.download(reinterpret_cast<float*>(r.data()), n);
    return r;
  } catch (const std::exception& e) {
    gpu::Report(inverse ? "Fft2d::Inverse" : "Fft2d::Forward", e);
    return {};
  }
}

}  // namespace hvd
