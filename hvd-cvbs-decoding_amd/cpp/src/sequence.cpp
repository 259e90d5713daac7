// sequence.cpp -- PrepareFieldObs, DecodeFieldWindow and
// DecodeFieldWindowWithInits over the GPU library: the window's inits run as
// one batch, then the library's window driver (2-D and 3-D).
#include <algorithm>

#include "engine/engine.h"
#include "engine/holographic_init.h"
#include "engine/lockin.h"
#include "engine/sequence.h"
#include "gpu_session.h"

namespace hvd {

using gpu::DevBuf;
using gpu::Session;

FieldObs PrepareFieldObs(const FieldInput& field, const FieldGeometry& g, const HvdConfig& cfg, int parity) {
  // sequence.cpp:229-268: lock-in on the whole field, phase trim, active crop
  // and the carrier of the active lines.
  std::vector<float> theta;
  std::vector<int8_t> vswitch;
  if (g.uses_vswitch()) {
    PalBurstLockin l = BurstLockinPhasePal(field.samples, g);
    theta = std::move(l.theta);
    vswitch = std::move(l.vswitch);
  } else {
    theta = BurstLockinPhase(field.samples, g);
  }
  FieldObs obs;
  obs.parity = parity;
  const int fal = g.first_active_field_line, lal = g.last_active_line();
  const int a0 = g.active_video_start, a1 = g.active_video_end;
  const int lines = std::max(0, lal - fal), width = std::max(0, a1 - a0);
  if (static_cast<int>(theta.size()) < lal || field.samples.width() < a1 || fal < 0 || a0 < 0) return obs;
  ApplyChromaPhase(cfg.chroma_phase_deg, g, vswitch, &theta);
  obs.s = Plane(lines, width);
  for (int y = 0; y < lines; ++y)
    std::copy(&field.samples.at(fal + y, a0), &field.samples.at(fal + y, a0) + width, &obs.s.at(y, 0));
  const std::vector<float> theta_active(theta.begin() + fal, theta.begin() + lal);
  if (g.uses_vswitch()) {
    const std::vector<int8_t> sw_active(vswitch.begin() + fal, vswitch.begin() + lal);
    obs.carrier = MakeCarrierPal(theta_active, sw_active, g);
  } else {
    obs.carrier = MakeCarrier(theta_active, g);
  }
  return obs;
}

namespace {

// The window decode on the device.  inits == nullptr: every field's init is
// HolographicInit's (sequence.cpp:270-281); else the caller's chi0 and,
// where given, Y0 (DecodeFieldWindowWithInits).  A window of one field
// shape runs the library's window driver (sequence.cpp:423-1114: noise
// calibration, adaptive strength, the 3-D passes, output fidelity); fields
// of differing shapes, which hvd-core does not accept in one window, can
// only take the 2-D path and run as one solve batch.
std::vector<DecodedField> WindowOn(Session& ss, const std::vector<FieldObs>& fields,
                                   const std::vector<DecodedField>* inits, const FieldGeometry& g,
                                   const HvdConfig& cfg, SequenceDiagnostics* diag) {
  const int nf = static_cast<int>(fields.size());
  const bool pal = g.uses_vswitch();
  if (diag) *diag = SequenceDiagnostics{};
  std::vector<DecodedField> out(nf);
  if (nf == 0) return out;
  if (inits && static_cast<int>(inits->size()) != nf)
    throw gpu::Error("inits must be index-aligned with fields");
  bool one_shape = true;
  for (int j = 1; j < nf; ++j)
    one_shape = one_shape && fields[j].s.height() == fields[0].s.height() &&
                fields[j].s.width() == fields[0].s.width();
  if (!one_shape && (inits || gpu::WantsTemporal(cfg, nf, pal)))
    throw gpu::Error("all fields of a window must share one shape");
  std::vector<DevBuf<float>> s(nf), car(nf), chi(nf), y(nf);
  std::vector<hvd_gpu_plane> planes(nf);
  std::vector<int> parities(nf), y_given(nf, 1);
  for (int j = 0; j < nf; ++j) {
    const FieldObs& f = fields[j];
    if (f.carrier.height() != f.s.height() || f.carrier.width() != f.s.width())
      throw gpu::Error("field " + std::to_string(j) + ": carrier size differs from S");
    if (f.s.empty()) throw gpu::Error("field " + std::to_string(j) + ": empty plane");
    s[j] = gpu::Upload(f.s.data(), f.s.size());
    car[j] = gpu::Upload(reinterpret_cast<const float*>(f.carrier.data()), f.carrier.size() * 2);
    chi[j].resize(f.s.size() * 2);
    y[j].resize(f.s.size());
    if (inits) {
      const DecodedField& in = (*inits)[j];
      if (in.chroma.height() != f.s.height() || in.chroma.width() != f.s.width())
        throw gpu::Error("init " + std::to_string(j) + ": chroma size differs from S");
      chi[j].upload(reinterpret_cast<const float*>(in.chroma.data()), f.s.size() * 2);
      if (in.luma.empty()) {
        y_given[j] = 0;
      } else {
        if (in.luma.height() != f.s.height() || in.luma.width() != f.s.width())
          throw gpu::Error("init " + std::to_string(j) + ": luma size differs from S");
        y[j].upload(in.luma.data(), f.s.size());
      }
    }
    parities[j] = f.parity;
    hvd_gpu_plane& p = planes[j];
    p = hvd_gpu_plane{};
    p.lines = f.s.height();
    p.width = f.s.width();
    p.s = s[j].get();
    p.carrier = car[j].get();
    p.active_video_start = g.active_video_start;
    p.phase_per_sample = g.phase_per_sample();
    p.sample_rate = g.sample_rate;
    p.chi = chi[j].get();
    p.y = y[j].get();
  }
  ss.autotune(planes[0].lines, planes[0].width, pal, nf);
  // sequence.cpp:270-281: the inits, Y0 with them
  if (!inits)
    ss.check(hvd_gpu_holographic_init(ss.ctx(), planes.data(), nf, cfg.symmetry_variant ? 1 : 0, nullptr),
             "hvd_gpu_holographic_init");
  const hvd_solver_params sp = gpu::ToAbi(cfg, pal);
  if (one_shape) {
    const hvd_window_params wp = gpu::WindowOf(cfg);
    hvd_window_diag d{};
    ss.check(hvd_gpu_decode_window(ss.ctx(), planes.data(), nf, parities.data(), pal ? 1 : 0, y_given.data(), &sp,
                                   &wp, &d, nullptr),
             "hvd_gpu_decode_window");
    if (diag) {
      diag->resolved_strength = d.resolved_strength;
      diag->ambiguity_ire = d.ambiguity_ire;
      diag->sigma_ire = d.sigma_ire;
      diag->temporal_eps = d.temporal_eps;
      diag->nr_eps = d.nr_eps;
    }
  } else {
    // the decoupled per-field solve with data-fidelity luma (:1046, :1102-1114)
    ss.check(hvd_gpu_variational_refine(ss.ctx(), planes.data(), nf, &sp, nullptr, nullptr, nullptr),
             "hvd_gpu_variational_refine");
  }
  for (int j = 0; j < nf; ++j) {
    const Plane& sj = fields[j].s;
    out[j].luma = Plane(sj.height(), sj.width());
    out[j].chroma = ComplexPlane(sj.height(), sj.width());
    y[j].download(out[j].luma.data(), sj.size());
    chi[j].download(reinterpret_cast<float*>(out[j].chroma.data()), sj.size() * 2);
  }
  return out;
}

}  // namespace

std::vector<DecodedField> DecodeFieldWindow(const std::vector<FieldObs>& fields, const FieldGeometry& g,
                                            const HvdConfig& cfg, Fft2d* /*fft*/, SequenceDiagnostics* diag) {
  try {
    return WindowOn(gpu::ThreadSession(), fields, nullptr, g, cfg, diag);
  } catch (const std::exception& e) {
    gpu::Report("DecodeFieldWindow", e);
    return {};
  }
}

std::vector<DecodedField> DecodeFieldWindowWithInits(const std::vector<FieldObs>& fields,
                                                     const std::vector<DecodedField>& inits, const FieldGeometry& g,
                                                     const HvdConfig& cfg, SequenceDiagnostics* diag) {
  try {
    return WindowOn(gpu::ThreadSession(), fields, &inits, g, cfg, diag);
  } catch (const std::exception& e) {
    gpu::Report("DecodeFieldWindowWithInits", e);
    return {};
  }
}

DrizzleResult DrizzleFrame(int j0, const std::vector<Plane>& Ys, const std::vector<ComplexPlane>& chis,
                           const std::vector<int>& parities, const HvdConfig& cfg, int scale) {
  try {
    const int n = static_cast<int>(Ys.size());
    if (j0 < 0 || j0 + 1 >= n || chis.size() != Ys.size() || parities.size() != Ys.size())
      throw gpu::Error("j0, j0 + 1 must index a window of equally many Y, chi and parities");
    const int L = Ys[j0].height(), W = Ys[j0].width();
    for (int k = 0; k < n; ++k)
      if (Ys[k].height() != L || Ys[k].width() != W || chis[k].height() != L || chis[k].width() != W)
        throw gpu::Error("all fields of the window must share one shape");
    Session& ss = gpu::ThreadSession();
    std::vector<DevBuf<float>> dy(n), dc(n);
    std::vector<const float*> py(n), pc(n);
    for (int k = 0; k < n; ++k) {
      dy[k] = gpu::Upload(Ys[k].data(), Ys[k].size());
      dc[k] = gpu::Upload(reinterpret_cast<const float*>(chis[k].data()), chis[k].size() * 2);
      py[k] = dy[k].get();
      pc[k] = dc[k].get();
    }
    const size_t nfine = (size_t)2 * L * scale * W;
    DevBuf<float> oy(nfine), oc(2 * nfine);
    ss.check(hvd_gpu_drizzle_frame(ss.ctx(), j0, n, py.data(), pc.data(), parities.data(), L, W, cfg.mc_tile,
                                   cfg.mc_search, cfg.nr_radius, cfg.nr_eps, scale, oy.get(), oc.get(), nullptr),
             "hvd_gpu_drizzle_frame");
    DrizzleResult r;
    r.luma = Plane(2 * L * scale, W);
    r.chroma = ComplexPlane(2 * L * scale, W);
    oy.download(r.luma.data(), nfine);
    oc.download(reinterpret_cast<float*>(r.chroma.data()), 2 * nfine);
    return r;
  } catch (const std::exception& e) {
    gpu::Report("DrizzleFrame", e);
    return {};
  }
}

std::vector<DecodedField> HvdEngine::DecodeSequenceWindow(const std::vector<FieldObs>& fields,
                                                          const FieldGeometry& g, const HvdConfig& cfg,
                                                          SequenceDiagnostics* diag) {
  try {
    return WindowOn(session(), fields, nullptr, g, cfg, diag);
  } catch (const std::exception& e) {
    gpu::Report("HvdEngine::DecodeSequenceWindow", e);
    return {};
  }
}

}  // namespace hvd
