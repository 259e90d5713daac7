// ref_fb_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// hvd-core's AmbiguousBoxes (frame_bridge.cpp:399-621) is internal to its
// translation unit, so this shim compiles that unit itself (included from
// the reference checkout where it lies, never copied) and exposes the one
// function through a C entry point for the oracle's pin.  Nothing else of
// the unit is called: its other dependencies (the engine, the FFT) stay
// unresolved in the shared library and are never reached.
#include "frame_bridge.cpp"

namespace hvd {

// boxes: cap x 4 ints (fy0, fy1, x0, x1) of the valid boxes; returns their
// count, 0 for none (stay 2-D), -1 for the single invalid box (full 3-D).
static int RefAmbiguousBoxes(const float* y, const float* chi, int h, int w, float max_area, int* boxes, int cap) {
  Plane Y(h, w);
  ComplexPlane C(h, w);
  for (int r = 0; r < h; ++r)
    for (int x = 0; x < w; ++x) {
      const size_t i = static_cast<size_t>(r) * w + x;
      Y.at(r, x) = y[i];
      C.at(r, x) = Complex(chi[2 * i], chi[2 * i + 1]);
    }
  const std::vector<SelBox> b = AmbiguousBoxes(Y, C, max_area);
  if (b.empty()) return 0;
  if (!b[0].valid) return -1;
  const int n = static_cast<int>(b.size());
  for (int k = 0; k < n && k < cap; ++k) {
    boxes[4 * k] = b[k].fy0;
    boxes[4 * k + 1] = b[k].fy1;
    boxes[4 * k + 2] = b[k].x0;
    boxes[4 * k + 3] = b[k].x1;
  }
  return n;
}

}  // namespace hvd

extern "C" int ref_ambiguous_boxes(const float* y, const float* chi, int h, int w, float max_area, int* boxes,
                                   int cap) {
  return hvd::RefAmbiguousBoxes(y, chi, h, w, max_area, boxes, cap);
}
