// temporal.cpp -- hvd-core's temporal-neighbour warps (temporal.cpp) over
// the GPU library: planes are copied to the device, the C ABI does the
// work, results are copied back.
#include "engine/temporal.h"

#include "engine/gpu_status.h"
#include "gpu_session.h"

namespace hvd {

using gpu::DevBuf;
using gpu::Session;

namespace {

struct DevVectors {
  DevBuf<float> dy, dx;
};

DevVectors Vectors(Session& s, const Plane& mdy, const Plane& mdx, int tile, int h, int w) {
  if (mdy.height() != mdx.height() || mdy.width() != mdx.width()) throw gpu::Error("motion planes differ in size");
  DevBuf<float> ty = gpu::Upload(mdy.data(), mdy.size()), tx = gpu::Upload(mdx.data(), mdx.size());
  DevVectors v{DevBuf<float>((size_t)h * w), DevBuf<float>((size_t)h * w)};
  s.check(hvd_gpu_vectors_per_pixel(s.ctx(), ty.get(), tx.get(), mdy.height(), mdy.width(), tile, h, w, v.dy.get(),
                                    v.dx.get(), nullptr),
          "hvd_gpu_vectors_per_pixel");
  return v;
}

DevVectors Vectors(const PerPixelMotion& v, int h, int w) {
  if (v.dy.height() != h || v.dy.width() != w || v.dx.height() != h || v.dx.width() != w)
    throw gpu::Error("per-pixel vectors differ in size from the plane");
  return DevVectors{gpu::Upload(v.dy.data(), v.dy.size()), gpu::Upload(v.dx.data(), v.dx.size())};
}

template <typename P>
P Warp(Session& s, const P& a, const DevVectors& v, int ch) {
  P out(a.height(), a.width());
  if (a.empty()) return out;
  DevBuf<float> da = gpu::Upload(reinterpret_cast<const float*>(a.data()), a.size() * ch), o(a.size() * ch);
  s.check(hvd_gpu_warp_by_tiles(s.ctx(), da.get(), a.height(), a.width(), ch, v.dy.get(), v.dx.get(), o.get(),
                                nullptr),
          "hvd_gpu_warp_by_tiles");
  o.download(reinterpret_cast<float*>(out.data()), a.size() * ch);
  return out;
}

Plane Bilinear(Session& s, const Plane& a, const DevVectors& v, float row_offset, int out_h, int out_w) {
  Plane out(out_h, out_w);
  DevBuf<float> da = gpu::Upload(a.data(), a.size()), o(out.size());
  s.check(hvd_gpu_warp_bilinear(s.ctx(), da.get(), a.height(), a.width(), v.dy.get(), v.dx.get(), row_offset, out_h,
                                out_w, o.get(), nullptr),
          "hvd_gpu_warp_bilinear");
  o.download(out.data(), out.size());
  return out;
}

Plane Upsample(const Plane& conf, int tile, int out_h, int out_w, bool fast) {
  Session& s = gpu::ThreadSession();
  Plane out(out_h, out_w);
  DevBuf<float> c = gpu::Upload(conf.data(), conf.size()), o(out.size());
  s.check(hvd_gpu_upsample_confidence(s.ctx(), c.get(), conf.height(), conf.width(), tile, out_h, out_w, fast ? 1 : 0,
                                      o.get(), nullptr),
          "hvd_gpu_upsample_confidence");
  o.download(out.data(), out.size());
  return out;
}

}  // namespace

PerPixelMotion VectorsPerPixel(const Plane& mdy, const Plane& mdx, int tile, int out_h, int out_w) {
  try {
    Session& s = gpu::ThreadSession();
    DevVectors v = Vectors(s, mdy, mdx, tile, out_h, out_w);
    PerPixelMotion out{Plane(out_h, out_w), Plane(out_h, out_w)};
    v.dy.download(out.dy.data(), out.dy.size());
    v.dx.download(out.dx.data(), out.dx.size());
    return out;
  } catch (const std::exception& e) {
    gpu::Report("VectorsPerPixel", e);
    return {};
  }
}

Plane WarpByTiles(const Plane& a, const Plane& mdy, const Plane& mdx, int tile) {
  try {
    Session& s = gpu::ThreadSession();
    return Warp(s, a, Vectors(s, mdy, mdx, tile, a.height(), a.width()), 1);
  } catch (const std::exception& e) {
    gpu::Report("WarpByTiles", e);
    return {};
  }
}

Plane WarpByTiles(const Plane& a, const PerPixelMotion& v) {
  try {
    return Warp(gpu::ThreadSession(), a, Vectors(v, a.height(), a.width()), 1);
  } catch (const std::exception& e) {
    gpu::Report("WarpByTiles", e);
    return {};
  }
}

ComplexPlane WarpByTilesComplex(const ComplexPlane& a, const Plane& mdy, const Plane& mdx, int tile) {
  try {
    Session& s = gpu::ThreadSession();
    return Warp(s, a, Vectors(s, mdy, mdx, tile, a.height(), a.width()), 2);
  } catch (const std::exception& e) {
    gpu::Report("WarpByTilesComplex", e);
    return {};
  }
}

ComplexPlane WarpByTilesComplex(const ComplexPlane& a, const PerPixelMotion& v) {
  try {
    return Warp(gpu::ThreadSession(), a, Vectors(v, a.height(), a.width()), 2);
  } catch (const std::exception& e) {
    gpu::Report("WarpByTilesComplex", e);
    return {};
  }
}

Plane WarpBilinearTiles(const Plane& a, const Plane& dyf, const Plane& dxf, int tile, float row_offset, int out_h,
                        int out_w) {
  try {
    Session& s = gpu::ThreadSession();
    return Bilinear(s, a, Vectors(s, dyf, dxf, tile, out_h, out_w), row_offset, out_h, out_w);
  } catch (const std::exception& e) {
    gpu::Report("WarpBilinearTiles", e);
    return {};
  }
}

Plane WarpBilinearTiles(const Plane& a, const PerPixelMotion& v, float row_offset, int out_h, int out_w) {
  try {
    return Bilinear(gpu::ThreadSession(), a, Vectors(v, out_h, out_w), row_offset, out_h, out_w);
  } catch (const std::exception& e) {
    gpu::Report("WarpBilinearTiles", e);
    return {};
  }
}

Plane ComplexCoherence(const ComplexPlane& z1, const ComplexPlane& z2, int r) {
  try {
    if (z1.height() != z2.height() || z1.width() != z2.width()) throw gpu::Error("phasor fields differ in size");
    Session& s = gpu::ThreadSession();
    Plane out(z1.height(), z1.width());
    if (z1.empty()) return out;
    DevBuf<float> a = gpu::Upload(reinterpret_cast<const float*>(z1.data()), z1.size() * 2);
    DevBuf<float> b = gpu::Upload(reinterpret_cast<const float*>(z2.data()), z2.size() * 2);
    DevBuf<float> o(out.size());
    s.check(hvd_gpu_complex_coherence(s.ctx(), a.get(), b.get(), z1.height(), z1.width(), r, o.get(), nullptr),
            "hvd_gpu_complex_coherence");
    o.download(out.data(), out.size());
    return out;
  } catch (const std::exception& e) {
    gpu::Report("ComplexCoherence", e);
    return {};
  }
}

Plane UpsampleConfidence(const Plane& conf, int tile, int out_h, int out_w) {
  try {
    return Upsample(conf, tile, out_h, out_w, false);
  } catch (const std::exception& e) {
    gpu::Report("UpsampleConfidence", e);
    return {};
  }
}

Plane UpsampleConfidenceFast(const Plane& conf, int tile, int out_h, int out_w) {
  try {
    return Upsample(conf, tile, out_h, out_w, true);
  } catch (const std::exception& e) {
    gpu::Report("UpsampleConfidenceFast", e);
    return {};
  }
}

MotionCompensatedResult MotionCompensatePrev(const NeighborRawState& prev, const Plane& y_cur_init, int tile,
                                             int search, const MotionField* motion, bool fast,
                                             const PerPixelMotion* vpix) {
  try {
    MotionField local;
    const MotionField* mf = motion;
    if (!mf) {
      local = EstimateMotion(prev.luma, y_cur_init, tile, search);
      if (local.dy.empty()) throw gpu::Error("motion estimation failed: " + GpuLastError());
      mf = &local;
    }
    const int h = prev.composite.height(), w = prev.composite.width();
    if (prev.carrier.height() != h || prev.carrier.width() != w) throw gpu::Error("carrier differs from composite");
    Session& s = gpu::ThreadSession();
    const DevVectors v = vpix ? Vectors(*vpix, h, w) : Vectors(s, mf->dy, mf->dx, tile, h, w);
    MotionCompensatedResult r;
    r.composite = Warp(s, prev.composite, v, 1);
    r.carrier = Warp(s, prev.carrier, v, 2);
    r.confidence = Upsample(mf->confidence, tile, prev.luma.height(), prev.luma.width(), fast);
    return r;
  } catch (const std::exception& e) {
    gpu::Report("MotionCompensatePrev", e);
    return {};
  }
}

}  // namespace hvd
