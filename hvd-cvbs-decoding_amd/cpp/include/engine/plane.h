// engine/plane.h -- host image planes of the hvd-core C++ API
// (hvd-core/src/engine/plane.h:25-72): row-major owning buffers, float for
// real planes and std::complex<float> (interleaved re, im) for complex ones.
// The GPU library reads and writes the same storage, so a plane's data()
// pointer is what is copied to and from the device.
#ifndef HVD_AMD_ENGINE_PLANE_H_
#define HVD_AMD_ENGINE_PLANE_H_

#include <cassert>
#include <complex>
#include <cstddef>
#include <vector>

namespace hvd {

using Complex = std::complex<float>;

template <typename T>
class BasicPlane {
 public:
  BasicPlane() = default;
  BasicPlane(int height, int width, T fill = T{})
      : height_(height), width_(width), data_(static_cast<size_t>(height) * static_cast<size_t>(width), fill) {
    assert(height >= 0 && width >= 0);
  }

  int height() const { return height_; }
  int width() const { return width_; }
  size_t size() const { return data_.size(); }
  bool empty() const { return data_.empty(); }

  T* data() { return data_.data(); }
  const T* data() const { return data_.data(); }

  T& at(int y, int x) {
    assert(y >= 0 && y < height_ && x >= 0 && x < width_);
    return data_[static_cast<size_t>(y) * width_ + x];
  }
  const T& at(int y, int x) const {
    assert(y >= 0 && y < height_ && x >= 0 && x < width_);
    return data_[static_cast<size_t>(y) * width_ + x];
  }

  T& operator[](size_t i) { return data_[i]; }
  const T& operator[](size_t i) const { return data_[i]; }

 private:
  int height_ = 0;
  int width_ = 0;
  std::vector<T> data_;
};

using Plane = BasicPlane<float>;
using ComplexPlane = BasicPlane<Complex>;

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_PLANE_H_
