// components.cpp -- hvd-core's component functions (lock-in, carrier, init,
// solve, motion, colour) over the GPU library.  Host planes are copied to
// the device, the C ABI does the work, results are copied back; on any
// failure the function returns an empty result (GpuLastError() says why).
#include <algorithm>

#include "engine/colour.h"
#include "engine/holographic_init.h"
#include "engine/lockin.h"
#include "engine/motion.h"
#include "engine/ntsc_geometry.h"
#include "engine/variational.h"
#include "gpu_session.h"

namespace hvd {

using gpu::DevBuf;
using gpu::Session;

namespace {

// The lock-in reads the plane's own size as the field's size.
hvd_field_geometry LockinGeometry(const Plane& field, const FieldGeometry& g) {
  hvd_field_geometry a = gpu::ToAbi(g);
  a.field_width = field.width();
  a.field_height = field.height();
  return a;
}

struct LockinOut {
  std::vector<float> theta;
  std::vector<int8_t> vswitch;
};

LockinOut RunLockin(const Plane& field, const FieldGeometry& g) {
  Session& s = gpu::ThreadSession();
  const hvd_field_geometry a = LockinGeometry(field, g);
  const int h = field.height();
  DevBuf<float> d = gpu::Upload(field.data(), field.size());
  DevBuf<float> th(h);
  DevBuf<int8_t> vs(h);
  const float* fp = d.get();
  float* tp = th.get();
  int8_t* vp = vs.get();
  s.check(hvd_gpu_burst_lockin_phase(s.ctx(), &fp, 1, &a, &tp, &vp, nullptr), "hvd_gpu_burst_lockin_phase");
  LockinOut o;
  o.theta.resize(h);
  th.download(o.theta.data(), h);
  o.vswitch.assign(h, 1);
  if (g.uses_vswitch()) vs.download(o.vswitch.data(), h);
  return o;
}

}  // namespace

// NTSC lock-in (lockin.cpp:101-186).  A PAL-family geometry is locked with
// the joint phase + V-switch estimate, whose theta is returned.
std::vector<float> BurstLockinPhase(const Plane& field_ire, const FieldGeometry& g) {
  try {
    if (field_ire.empty()) return {};
    return RunLockin(field_ire, g).theta;
  } catch (const std::exception& e) {
    gpu::Report("BurstLockinPhase", e);
    return {};
  }
}

PalBurstLockin BurstLockinPhasePal(const Plane& field_ire, const FieldGeometry& g) {
  try {
    if (field_ire.empty()) return {};
    LockinOut o = RunLockin(field_ire, g);
    return PalBurstLockin{std::move(o.theta), std::move(o.vswitch)};
  } catch (const std::exception& e) {
    gpu::Report("BurstLockinPhasePal", e);
    return {};
  }
}

float BurstAmplitudeIre(const Plane& field_ire, const FieldGeometry& g) {
  try {
    if (field_ire.empty()) return g.nominal_burst_ire();
    Session& s = gpu::ThreadSession();
    const hvd_field_geometry a = LockinGeometry(field_ire, g);
    DevBuf<float> d = gpu::Upload(field_ire.data(), field_ire.size());
    DevBuf<float> amp(1);
    const float* fp = d.get();
    s.check(hvd_gpu_burst_amplitude_ire(s.ctx(), &fp, 1, &a, amp.get(), nullptr), "hvd_gpu_burst_amplitude_ire");
    float r = 0.0F;
    amp.download(&r, 1);
    return r;
  } catch (const std::exception& e) {
    gpu::Report("BurstAmplitudeIre", e);
    return 0.0F;
  }
}

std::vector<float> TridiagSmooth(const std::vector<float>& d, const std::vector<float>& a, float lam) {
  try {
    if (d.size() != a.size()) throw gpu::Error("d and a differ in length");
    const int n = static_cast<int>(d.size());
    if (n == 0) return {};
    Session& s = gpu::ThreadSession();
    DevBuf<float> dd = gpu::Upload(d.data(), d.size()), da = gpu::Upload(a.data(), a.size()), out(d.size());
    s.check(hvd_gpu_tridiag_smooth(s.ctx(), dd.get(), da.get(), n, lam, out.get(), nullptr), "hvd_gpu_tridiag_smooth");
    std::vector<float> x(n);
    out.download(x.data(), n);
    return x;
  } catch (const std::exception& e) {
    gpu::Report("TridiagSmooth", e);
    return {};
  }
}

float AccGain(float median_burst_amplitude_ire) { return AccGain(median_burst_amplitude_ire, 20.0F); }

float AccGain(float median_burst_amplitude_ire, float nominal_ire) {
  return std::clamp(nominal_ire / std::max(median_burst_amplitude_ire, 1.0F), 0.5F, 2.0F);
}

namespace {

ComplexPlane Carrier(const std::vector<float>& theta, const int8_t* vswitch_host, const FieldGeometry& g) {
  const int lines = static_cast<int>(theta.size());
  const int width = g.active_width();
  ComplexPlane c(lines, std::max(width, 0));
  if (c.empty()) return c;
  Session& s = gpu::ThreadSession();
  DevBuf<float> th = gpu::Upload(theta.data(), theta.size());
  DevBuf<int8_t> vs;
  if (vswitch_host) vs = gpu::Upload(vswitch_host, theta.size());
  DevBuf<float> out(c.size() * 2);
  const hvd_field_geometry a = gpu::ToAbi(g);
  s.check(hvd_gpu_make_carrier(s.ctx(), th.get(), vswitch_host ? vs.get() : nullptr, lines, &a, out.get(), nullptr),
          "hvd_gpu_make_carrier");
  out.download(reinterpret_cast<float*>(c.data()), c.size() * 2);
  return c;
}

}  // namespace

ComplexPlane MakeCarrier(const std::vector<float>& theta, const FieldGeometry& g) {
  try {
    return Carrier(theta, nullptr, g);
  } catch (const std::exception& e) {
    gpu::Report("MakeCarrier", e);
    return {};
  }
}

ComplexPlane MakeCarrierPal(const std::vector<float>& theta, const std::vector<int8_t>& vswitch,
                            const FieldGeometry& g) {
  try {
    if (vswitch.size() != theta.size()) throw gpu::Error("theta and vswitch differ in length");
    return Carrier(theta, vswitch.data(), g);
  } catch (const std::exception& e) {
    gpu::Report("MakeCarrierPal", e);
    return {};
  }
}

namespace {

hvd_gpu_plane PlaneOf(const Plane& s, const DevBuf<float>& ds, const DevBuf<float>& dc, const FieldGeometry& g,
                      float* chi, float* y) {
  hvd_gpu_plane p{};
  p.lines = s.height();
  p.width = s.width();
  p.s = ds.get();
  p.carrier = dc.get();
  p.active_video_start = g.active_video_start;
  p.phase_per_sample = g.phase_per_sample();
  p.sample_rate = g.sample_rate;
  p.chi = chi;
  p.y = y;
  return p;
}

}  // namespace

HoloInit HolographicInit(const Plane& s, const ComplexPlane& carrier, const FieldGeometry& g, const HvdConfig& cfg,
                         Fft2d* /*fft*/) {
  try {
    if (carrier.height() != s.height() || carrier.width() != s.width()) throw gpu::Error("carrier size differs from S");
    HoloInit out;
    out.luma = Plane(s.height(), s.width());
    out.chroma = ComplexPlane(s.height(), s.width());
    if (s.empty()) return out;
    Session& ss = gpu::ThreadSession();
    DevBuf<float> ds = gpu::Upload(s.data(), s.size());
    DevBuf<float> dc = gpu::Upload(reinterpret_cast<const float*>(carrier.data()), carrier.size() * 2);
    DevBuf<float> chi(s.size() * 2), y(s.size());
    const hvd_gpu_plane p = PlaneOf(s, ds, dc, g, chi.get(), y.get());
    ss.check(hvd_gpu_holographic_init(ss.ctx(), &p, 1, cfg.symmetry_variant ? 1 : 0, nullptr),
             "hvd_gpu_holographic_init");
    y.download(out.luma.data(), s.size());
    chi.download(reinterpret_cast<float*>(out.chroma.data()), s.size() * 2);
    return out;
  } catch (const std::exception& e) {
    gpu::Report("HolographicInit", e);
    return {};
  }
}

RefineResult VariationalRefine(const Plane& s, const ComplexPlane& carrier, const ComplexPlane& chi0,
                               const HvdConfig& cfg, const std::vector<NeighborTerm>& neighbors) {
  try {
    if (carrier.height() != s.height() || carrier.width() != s.width() || chi0.height() != s.height() ||
        chi0.width() != s.width())
      throw gpu::Error("carrier / chi0 size differs from S");
    for (const NeighborTerm& nb : neighbors)
      if (nb.composite.height() != s.height() || nb.composite.width() != s.width() ||
          nb.carrier.height() != s.height() || nb.carrier.width() != s.width() ||
          nb.confidence.height() != s.height() || nb.confidence.width() != s.width())
        throw gpu::Error("neighbour plane size differs from S");
    RefineResult out;
    out.luma = Plane(s.height(), s.width());
    out.chroma = ComplexPlane(s.height(), s.width());
    if (s.empty()) return out;
    Session& ss = gpu::ThreadSession();
    DevBuf<float> ds = gpu::Upload(s.data(), s.size());
    DevBuf<float> dc = gpu::Upload(reinterpret_cast<const float*>(carrier.data()), carrier.size() * 2);
    DevBuf<float> chi = gpu::Upload(reinterpret_cast<const float*>(chi0.data()), chi0.size() * 2);
    DevBuf<float> y(s.size());
    FieldGeometry g;
    const hvd_gpu_plane p = PlaneOf(s, ds, dc, g, chi.get(), y.get());
    const hvd_solver_params sp = gpu::ToAbi(cfg, cfg.is_pal);
    if (neighbors.empty()) {
      ss.check(hvd_gpu_variational_refine(ss.ctx(), &p, 1, &sp, nullptr, nullptr, nullptr),
               "hvd_gpu_variational_refine");
    } else {
      // the 3-D solve: nu = temporal_strength (variational.cpp:221)
      std::vector<DevBuf<float>> bufs;
      std::vector<hvd_gpu_neighbor> nb(neighbors.size());
      bufs.reserve(3 * neighbors.size());
      for (size_t j = 0; j < neighbors.size(); ++j) {
        const NeighborTerm& t = neighbors[j];
        bufs.push_back(gpu::Upload(t.composite.data(), t.composite.size()));
        nb[j].composite = bufs.back().get();
        bufs.push_back(gpu::Upload(reinterpret_cast<const float*>(t.carrier.data()), t.carrier.size() * 2));
        nb[j].carrier = bufs.back().get();
        bufs.push_back(gpu::Upload(t.confidence.data(), t.confidence.size()));
        nb[j].confidence = bufs.back().get();
      }
      ss.check(hvd_gpu_variational_refine_temporal(ss.ctx(), &p, nb.data(), (int)nb.size(), &sp,
                                                   cfg.temporal_strength, cfg.temporal_eps, nullptr, nullptr),
               "hvd_gpu_variational_refine_temporal");
    }
    y.download(out.luma.data(), s.size());
    chi.download(reinterpret_cast<float*>(out.chroma.data()), s.size() * 2);
    return out;
  } catch (const std::exception& e) {
    gpu::Report("VariationalRefine", e);
    return {};
  }
}

MotionField EstimateMotion(const Plane& y_ref, const Plane& y_cur, int tile, int search) {
  try {
    if (y_ref.height() != y_cur.height() || y_ref.width() != y_cur.width()) throw gpu::Error("planes differ in size");
    if (tile < 1) throw gpu::Error("tile must be positive");
    const int h = y_ref.height(), w = y_ref.width();
    const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile;
    MotionField m;
    m.tile = tile;
    m.dy = Plane(th, tw);
    m.dx = Plane(th, tw);
    m.confidence = Plane(th, tw);
    Session& s = gpu::ThreadSession();
    DevBuf<float> r = gpu::Upload(y_ref.data(), y_ref.size()), c = gpu::Upload(y_cur.data(), y_cur.size());
    DevBuf<float> dy(m.dy.size()), dx(m.dx.size()), cf(m.confidence.size());
    const float* rp = r.get();
    const float* cp = c.get();
    float* dyp = dy.get();
    float* dxp = dx.get();
    float* cfp = cf.get();
    s.check(hvd_gpu_estimate_motion(s.ctx(), &rp, &cp, 1, h, w, tile, search, &dyp, &dxp, &cfp, nullptr),
            "hvd_gpu_estimate_motion");
    dy.download(m.dy.data(), m.dy.size());
    dx.download(m.dx.data(), m.dx.size());
    cf.download(m.confidence.data(), m.confidence.size());
    return m;
  } catch (const std::exception& e) {
    gpu::Report("EstimateMotion", e);
    return {};
  }
}

RefineResult VariationalRefineJoint(const Plane& s, const ComplexPlane& carrier, const Plane& y0,
                                    const ComplexPlane& chi0, const HvdConfig& cfg,
                                    const std::vector<NeighborTerm>& neighbors, const AnchorTerm* anchor) {
  try {
    auto same = [&](int h, int w) { return h == s.height() && w == s.width(); };
    if (!same(carrier.height(), carrier.width()) || !same(chi0.height(), chi0.width()) ||
        !same(y0.height(), y0.width()))
      throw gpu::Error("carrier / y0 / chi0 size differs from S");
    for (const NeighborTerm& nb : neighbors)
      if (!same(nb.composite.height(), nb.composite.width()) || !same(nb.carrier.height(), nb.carrier.width()) ||
          !same(nb.confidence.height(), nb.confidence.width()))
        throw gpu::Error("neighbour plane size differs from S");
    if (anchor && (!same(anchor->luma.height(), anchor->luma.width()) ||
                   !same(anchor->chroma.height(), anchor->chroma.width()) ||
                   !same(anchor->weight.height(), anchor->weight.width())))
      throw gpu::Error("anchor plane size differs from S");
    RefineResult out;
    out.luma = Plane(s.height(), s.width());
    out.chroma = ComplexPlane(s.height(), s.width());
    if (s.empty()) return out;
    Session& ss = gpu::ThreadSession();
    DevBuf<float> ds = gpu::Upload(s.data(), s.size());
    DevBuf<float> dc = gpu::Upload(reinterpret_cast<const float*>(carrier.data()), carrier.size() * 2);
    DevBuf<float> chi = gpu::Upload(reinterpret_cast<const float*>(chi0.data()), chi0.size() * 2);
    DevBuf<float> y = gpu::Upload(y0.data(), y0.size());
    FieldGeometry g;
    const hvd_gpu_plane p = PlaneOf(s, ds, dc, g, chi.get(), y.get());
    const hvd_solver_params sp = gpu::ToAbi(cfg, cfg.is_pal);
    std::vector<DevBuf<float>> bufs;
    std::vector<hvd_gpu_neighbor> nb(neighbors.size());
    bufs.reserve(3 * neighbors.size() + 3);
    for (size_t j = 0; j < neighbors.size(); ++j) {
      const NeighborTerm& t = neighbors[j];
      bufs.push_back(gpu::Upload(t.composite.data(), t.composite.size()));
      nb[j].composite = bufs.back().get();
      bufs.push_back(gpu::Upload(reinterpret_cast<const float*>(t.carrier.data()), t.carrier.size() * 2));
      nb[j].carrier = bufs.back().get();
      bufs.push_back(gpu::Upload(t.confidence.data(), t.confidence.size()));
      nb[j].confidence = bufs.back().get();
    }
    hvd_gpu_anchor an{};
    if (anchor) {
      bufs.push_back(gpu::Upload(anchor->luma.data(), anchor->luma.size()));
      an.luma = bufs.back().get();
      bufs.push_back(gpu::Upload(reinterpret_cast<const float*>(anchor->chroma.data()), anchor->chroma.size() * 2));
      an.chroma = bufs.back().get();
      bufs.push_back(gpu::Upload(anchor->weight.data(), anchor->weight.size()));
      an.weight = bufs.back().get();
    }
    const int nn = (int)nb.size();
    // nu = temporal_strength with neighbours, nu_a = nr_anchor with an
    // anchor (variational.cpp:581-582)
    ss.check(hvd_gpu_variational_refine_joint(ss.ctx(), &p, 1, nb.empty() ? nullptr : nb.data(), &nn,
                                              anchor ? &an : nullptr, &sp, cfg.temporal_strength, cfg.temporal_eps,
                                              anchor ? cfg.nr_anchor : 0.0F, nullptr, nullptr),
             "hvd_gpu_variational_refine_joint");
    y.download(out.luma.data(), s.size());
    chi.download(reinterpret_cast<float*>(out.chroma.data()), s.size() * 2);
    return out;
  } catch (const std::exception& e) {
    gpu::Report("VariationalRefineJoint", e);
    return {};
  }
}

MotionField VerifyMotion(const Plane& y_ref, const Plane& y_cur, const Plane& pdy, const Plane& pdx, int tile) {
  try {
    if (y_ref.height() != y_cur.height() || y_ref.width() != y_cur.width() || y_ref.empty())
      throw gpu::Error("reference and current planes differ in size or are empty");
    if (tile < 1) throw gpu::Error("tile must be positive");
    const int th = (y_ref.height() + tile - 1) / tile, tw = (y_ref.width() + tile - 1) / tile;
    if (pdy.height() != th || pdy.width() != tw || pdx.height() != th || pdx.width() != tw)
      throw gpu::Error("prediction planes are not on the tile grid");
    Session& s = gpu::ThreadSession();
    DevBuf<float> a = gpu::Upload(y_ref.data(), y_ref.size()), b = gpu::Upload(y_cur.data(), y_cur.size());
    DevBuf<float> py = gpu::Upload(pdy.data(), pdy.size()), px = gpu::Upload(pdx.data(), pdx.size());
    DevBuf<float> dy((size_t)th * tw), dx((size_t)th * tw), cf((size_t)th * tw);
    const float* ra[1] = {a.get()};
    const float* rb[1] = {b.get()};
    const float* rpy[1] = {py.get()};
    const float* rpx[1] = {px.get()};
    float* ody[1] = {dy.get()};
    float* odx[1] = {dx.get()};
    float* ocf[1] = {cf.get()};
    s.check(hvd_gpu_verify_motion(s.ctx(), ra, rb, 1, y_ref.height(), y_ref.width(), rpy, rpx, tile, ody, odx, ocf,
                                  nullptr),
            "hvd_gpu_verify_motion");
    MotionField m;
    m.tile = tile;
    m.dy = Plane(th, tw);
    m.dx = Plane(th, tw);
    m.confidence = Plane(th, tw);
    dy.download(m.dy.data(), m.dy.size());
    dx.download(m.dx.data(), m.dx.size());
    cf.download(m.confidence.data(), m.confidence.size());
    return m;
  } catch (const std::exception& e) {
    gpu::Report("VerifyMotion", e);
    return {};
  }
}

namespace {

// The measurements' planes on the device for the trajectory entry points.
struct TrajectoryBufs {
  std::vector<DevBuf<float>> dy, dx, cf;
  std::vector<int> off;
  std::vector<float> shift;
  int th = 0, tw = 0;
  explicit TrajectoryBufs(const std::vector<OffsetMotion>& ms) {
    th = ms.front().field.dy.height();
    tw = ms.front().field.dy.width();
    for (const OffsetMotion& m : ms) {
      const MotionField& f = m.field;
      if (f.dy.height() != th || f.dy.width() != tw || f.dx.height() != th || f.dx.width() != tw ||
          f.confidence.height() != th || f.confidence.width() != tw)
        throw gpu::Error("measurements on different tile grids");
      dy.push_back(gpu::Upload(f.dy.data(), f.dy.size()));
      dx.push_back(gpu::Upload(f.dx.data(), f.dx.size()));
      cf.push_back(gpu::Upload(f.confidence.data(), f.confidence.size()));
      off.push_back(m.offset);
      shift.push_back(m.parity_shift);
    }
  }
  template <typename P>
  std::vector<P> ptrs(std::vector<DevBuf<float>>& v) const {
    std::vector<P> p;
    for (auto& b : v) p.push_back(b.get());
    return p;
  }
};

}  // namespace

void FitTrajectory(const std::vector<OffsetMotion>& measurements, Plane* vy, Plane* vx) {
  if (measurements.empty() || !vy || !vx) return;
  try {
    TrajectoryBufs B(measurements);
    Session& s = gpu::ThreadSession();
    DevBuf<float> oy((size_t)B.th * B.tw), ox((size_t)B.th * B.tw);
    auto pdy = B.ptrs<const float*>(B.dy), pdx = B.ptrs<const float*>(B.dx), pcf = B.ptrs<const float*>(B.cf);
    s.check(hvd_gpu_trajectory_fit(s.ctx(), (int)measurements.size(), B.off.data(), B.shift.data(), pdy.data(),
                                   pdx.data(), pcf.data(), B.th, B.tw, oy.get(), ox.get(), nullptr),
            "hvd_gpu_trajectory_fit");
    *vy = Plane(B.th, B.tw);
    *vx = Plane(B.th, B.tw);
    oy.download(vy->data(), vy->size());
    ox.download(vx->data(), vx->size());
  } catch (const std::exception& e) {
    gpu::Report("FitTrajectory", e);
  }
}

void TrajectorySnap(std::vector<OffsetMotion>* measurements) {
  if (!measurements || measurements->size() < 2) return;
  try {
    TrajectoryBufs B(*measurements);
    Session& s = gpu::ThreadSession();
    auto pdy = B.ptrs<float*>(B.dy), pdx = B.ptrs<float*>(B.dx);
    auto pcf = B.ptrs<const float*>(B.cf);
    s.check(hvd_gpu_trajectory_snap(s.ctx(), (int)measurements->size(), B.off.data(), B.shift.data(), pdy.data(),
                                    pdx.data(), pcf.data(), B.th, B.tw, nullptr),
            "hvd_gpu_trajectory_snap");
    for (size_t j = 0; j < measurements->size(); ++j) {
      MotionField& f = (*measurements)[j].field;
      B.dy[j].download(f.dy.data(), f.dy.size());
      B.dx[j].download(f.dx.data(), f.dx.size());
    }
  } catch (const std::exception& e) {
    gpu::Report("TrajectorySnap", e);
  }
}

Plane BoxBlur(Plane a, int r) {
  try {
    if (a.empty() || r <= 0) return a;
    Session& s = gpu::ThreadSession();
    DevBuf<float> d = gpu::Upload(a.data(), a.size()), out(a.size());
    s.check(hvd_gpu_box_blur(s.ctx(), d.get(), a.height(), a.width(), r, out.get(), nullptr), "hvd_gpu_box_blur");
    out.download(a.data(), a.size());
    return a;
  } catch (const std::exception& e) {
    gpu::Report("BoxBlur", e);
    return {};
  }
}

}  // namespace hvd
