// engine/gradients.h -- the forward differences with Neumann boundaries and
// their exact adjoints of the hvd-core C++ API (hvd-core/src/engine/
// gradients.h), for float (luma) and Complex (chroma) planes: Dx, Dy, DxT,
// DyT returning planes, the *Into forms writing a caller-owned plane of a's
// size, the diagonal D1/D2 pair and the weighted forms of the solver.  Every
// operator runs on the GPU (hvd_gpu_difference); a plane that cannot be
// processed comes back empty (the *Into forms leave d as it was), with the
// reason in GpuLastError().
#ifndef HVD_AMD_ENGINE_GRADIENTS_H_
#define HVD_AMD_ENGINE_GRADIENTS_H_

#include <type_traits>

#include "engine/plane.h"

namespace hvd {

namespace grad_detail {
// d = op(a) on the device; op as hvd_gpu_diff_op (include/hvd_gpu.h).
// Returns false on failure (recorded for GpuLastError()).
bool Run(int op, bool is_complex, const float* a, const float* weight, int h, int w, float* d);

template <typename T>
bool Apply(int op, const BasicPlane<T>& a, const BasicPlane<float>* wt, BasicPlane<T>& d) {
  static_assert(std::is_same<T, float>::value || std::is_same<T, Complex>::value,
                "gradients are defined on float and Complex planes");
  if (d.height() != a.height() || d.width() != a.width()) return false;
  if (wt && (wt->height() != a.height() || wt->width() != a.width())) return false;
  return Run(op, std::is_same<T, Complex>::value, reinterpret_cast<const float*>(a.data()),
             wt ? wt->data() : nullptr, a.height(), a.width(), reinterpret_cast<float*>(d.data()));
}

template <typename T>
BasicPlane<T> Make(int op, const BasicPlane<T>& a) {
  BasicPlane<T> d(a.height(), a.width());
  if (!Apply(op, a, nullptr, d)) return {};
  return d;
}
}  // namespace grad_detail

// d[:, :-1] = a[:, 1:] - a[:, :-1]; d[:, -1] = 0  (gradients.h:26-42)
template <typename T>
BasicPlane<T> Dx(const BasicPlane<T>& a) { return grad_detail::Make(0, a); }
// d[:-1] = a[1:] - a[:-1]; d[-1] = 0  (:44-63)
template <typename T>
BasicPlane<T> Dy(const BasicPlane<T>& a) { return grad_detail::Make(1, a); }
// the adjoint of Dx (:65-83)
template <typename T>
BasicPlane<T> DxT(const BasicPlane<T>& a) { return grad_detail::Make(2, a); }
// the adjoint of Dy (:85-104)
template <typename T>
BasicPlane<T> DyT(const BasicPlane<T>& a) { return grad_detail::Make(3, a); }

template <typename T>
void DxInto(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(0, a, nullptr, d); }
template <typename T>
void DyInto(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(1, a, nullptr, d); }
template <typename T>
void DxTInto(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(2, a, nullptr, d); }
template <typename T>
void DyTInto(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(3, a, nullptr, d); }

// +45 deg: d[:-1, :-1] = a[1:, 1:] - a[:-1, :-1], last row and column 0
// (:196-214), and its adjoint (:216-236)
template <typename T>
void D1Into(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(4, a, nullptr, d); }
template <typename T>
void D1TInto(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(5, a, nullptr, d); }
// -45 deg: d[:-1, 1:] = a[1:, :-1] - a[:-1, 1:], last row and first column 0
// (:238-256), and its adjoint (:258-278)
template <typename T>
void D2Into(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(6, a, nullptr, d); }
template <typename T>
void D2TInto(const BasicPlane<T>& a, BasicPlane<T>& d) { grad_detail::Apply(7, a, nullptr, d); }

// d = w * D a with the same zeros (:290-381)
template <typename T>
void DxWeightedInto(const BasicPlane<T>& a, const BasicPlane<float>& w, BasicPlane<T>& d) {
  grad_detail::Apply(0, a, &w, d);
}
template <typename T>
void DyWeightedInto(const BasicPlane<T>& a, const BasicPlane<float>& w, BasicPlane<T>& d) {
  grad_detail::Apply(1, a, &w, d);
}
template <typename T>
void D1WeightedInto(const BasicPlane<T>& a, const BasicPlane<float>& w, BasicPlane<T>& d) {
  grad_detail::Apply(4, a, &w, d);
}
template <typename T>
void D2WeightedInto(const BasicPlane<T>& a, const BasicPlane<float>& w, BasicPlane<T>& d) {
  grad_detail::Apply(6, a, &w, d);
}

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_GRADIENTS_H_
