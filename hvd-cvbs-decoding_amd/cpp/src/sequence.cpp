// sequence.cpp -- PrepareFieldObs and the 2-D DecodeFieldWindow over the GPU
// library: the whole window's inits and solves run as one batch each.
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

std::vector<DecodedField> WindowOn(Session& ss, const std::vector<FieldObs>& fields, const FieldGeometry& g,
                                   const HvdConfig& cfg, SequenceDiagnostics* diag) {
  const int nf = static_cast<int>(fields.size());
  if (gpu::WantsTemporal(cfg, nf, g.uses_vswitch()))
    throw gpu::Error("this window would engage 3-D temporal decoding, which this library does not run; "
                     "set temporal_strength < 0 for the 2-D path");
  if (diag) *diag = SequenceDiagnostics{};
  std::vector<DecodedField> out(nf);
  if (nf == 0) return out;
  std::vector<DevBuf<float>> s(nf), car(nf), chi(nf), y(nf);
  std::vector<hvd_gpu_plane> planes(nf);
  for (int j = 0; j < nf; ++j) {
    const FieldObs& f = fields[j];
    if (f.carrier.height() != f.s.height() || f.carrier.width() != f.s.width())
      throw gpu::Error("field " + std::to_string(j) + ": carrier size differs from S");
    if (f.s.empty()) throw gpu::Error("field " + std::to_string(j) + ": empty plane");
    s[j] = gpu::Upload(f.s.data(), f.s.size());
    car[j] = gpu::Upload(reinterpret_cast<const float*>(f.carrier.data()), f.carrier.size() * 2);
    chi[j].resize(f.s.size() * 2);
    y[j].resize(f.s.size());
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
  ss.autotune(planes[0].lines, planes[0].width, g.uses_vswitch(), nf);
  // sequence.cpp:423-431 inits, then the decoupled per-field solve with the
  // standard flag from the geometry (:446) and data-fidelity luma (:1102-1114)
  ss.check(hvd_gpu_holographic_init(ss.ctx(), planes.data(), nf, cfg.symmetry_variant ? 1 : 0, nullptr),
           "hvd_gpu_holographic_init");
  const hvd_solver_params sp = gpu::ToAbi(cfg, g.uses_vswitch());
  ss.check(hvd_gpu_variational_refine(ss.ctx(), planes.data(), nf, &sp, nullptr, nullptr, nullptr),
           "hvd_gpu_variational_refine");
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
    return WindowOn(gpu::ThreadSession(), fields, g, cfg, diag);
  } catch (const std::exception& e) {
    gpu::Report("DecodeFieldWindow", e);
    return {};
  }
}

std::vector<DecodedField> HvdEngine::DecodeSequenceWindow(const std::vector<FieldObs>& fields,
                                                          const FieldGeometry& g, const HvdConfig& cfg,
                                                          SequenceDiagnostics* diag) {
  try {
    return WindowOn(session(), fields, g, cfg, diag);
  } catch (const std::exception& e) {
    gpu::Report("HvdEngine::DecodeSequenceWindow", e);
    return {};
  }
}

}  // namespace hvd
