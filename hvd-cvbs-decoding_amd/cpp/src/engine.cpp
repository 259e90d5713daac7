// engine.cpp -- HvdEngine over the GPU library: the woven-frame decode
// (WeaveAndBuildCarrier + HolographicInit + VariationalRefine, engine.cpp:53-236
// of hvd-core) with the fields woven, locked, initialised and solved on the
// device.
#include "engine/engine.h"

#include <algorithm>
#include <utility>

#include "engine/colour.h"
#include "woven.h"

namespace hvd {
namespace gpu {

WovenResult WovenDecode(Session& ss, const DeviceFields& f, const FieldGeometry& g_in, const HvdConfig& cfg,
                        bool chroma_only) {
  // engine.cpp:53-60: order the fields, top field on the even rows
  const float* top = f.first;
  const float* bot = f.second;
  if (!f.first_is_first && f.second_is_first) std::swap(top, bot);
  FieldGeometry g = g_in;
  g.field_height = f.field_lines;
  g.field_width = f.field_width;
  const int fal = g.first_active_field_line, lal = g.last_active_line();
  const int a0 = g.active_video_start, a1 = g.active_video_end;
  const int lines = std::max(0, lal - fal), width = std::max(0, a1 - a0);
  if (lines < 1 || width < 1 || fal < 0 || lal > f.field_lines || a0 < 0 || a1 > f.field_width)
    throw Error("active picture outside the fields");
  const bool pal = g.uses_vswitch();
  const hvd_field_geometry ag = ToAbi(g);
  hvd_gpu_ctx* ctx = ss.ctx();

  // per-field lock-in (engine.cpp:64-81)
  DevBuf<float> th(2 * (size_t)f.field_lines);
  DevBuf<int8_t> vs(2 * (size_t)f.field_lines);
  const float* fields[2] = {top, bot};
  float* thp[2] = {th.get(), th.get() + f.field_lines};
  int8_t* vsp[2] = {vs.get(), vs.get() + f.field_lines};
  const hvd_field_geometry ags[2] = {ag, ag};
  ss.check(hvd_gpu_burst_lockin_phase(ctx, fields, 2, ags, thp, vsp, nullptr), "hvd_gpu_burst_lockin_phase");
  std::vector<float> theta_f(2 * (size_t)f.field_lines);
  std::vector<int8_t> vs_f(2 * (size_t)f.field_lines, 1);
  th.download(theta_f.data(), theta_f.size());
  if (pal) vs.download(vs_f.data(), vs_f.size());

  // weave the active pictures (engine.cpp:83-107) on the device
  WovenResult w;
  w.rows = 2 * lines;
  w.width = width;
  const size_t np = (size_t)w.rows * width;
  w.s.resize(np);
  for (int k = 0; k < 2; ++k)
    HipCheck(hipMemcpy2D(w.s.get() + (size_t)k * width, 2 * width * sizeof(float),
                         fields[k] + (size_t)fal * f.field_width + a0, f.field_width * sizeof(float),
                         width * sizeof(float), lines, hipMemcpyDeviceToDevice),
             "hipMemcpy2D (weave)");
  std::vector<float> theta(w.rows);
  std::vector<int8_t> vswitch(w.rows, 1);
  for (int y = 0; y < lines; ++y) {
    theta[2 * y] = theta_f[fal + y];
    theta[2 * y + 1] = theta_f[f.field_lines + fal + y];
    if (pal) {
      vswitch[2 * y] = vs_f[fal + y];
      vswitch[2 * y + 1] = vs_f[f.field_lines + fal + y];
    }
  }
  // chroma phase trim on the reference itself (engine.cpp:109-133)
  ApplyChromaPhase(cfg.chroma_phase_deg, g, vswitch, &theta);
  DevBuf<float> dth = Upload(theta.data(), theta.size());
  DevBuf<int8_t> dvs = Upload(vswitch.data(), vswitch.size());
  w.carrier.resize(np * 2);
  ss.check(hvd_gpu_make_carrier(ctx, dth.get(), pal ? dvs.get() : nullptr, w.rows, &ag, w.carrier.get(), nullptr),
           "hvd_gpu_make_carrier");

  // holographic init, then the solve (engine.cpp:151-176)
  w.y.resize(np);
  w.chi.resize(np * 2);
  hvd_gpu_plane p{};
  p.lines = w.rows;
  p.width = width;
  p.s = w.s.get();
  p.carrier = w.carrier.get();
  p.active_video_start = a0;
  p.phase_per_sample = g.phase_per_sample();
  p.sample_rate = g.sample_rate;
  p.chi = w.chi.get();
  p.y = w.y.get();
  ss.check(hvd_gpu_holographic_init(ctx, &p, 1, cfg.symmetry_variant ? 1 : 0, nullptr), "hvd_gpu_holographic_init");
  w.chroma.resize(np);
  if (chroma_only) {
    ss.check(hvd_gpu_chroma_plane(ctx, 1, nullptr, nullptr, w.chi.get(), w.carrier.get(), np, w.chroma.get(), nullptr),
             "hvd_gpu_chroma_plane");
  } else {
    if (cfg.cg_iterations > 0 && !cfg.monochrome) {
      const hvd_solver_params sp = ToAbi(cfg, pal);
      ss.check(hvd_gpu_variational_refine(ctx, &p, 1, &sp, nullptr, nullptr, nullptr), "hvd_gpu_variational_refine");
    }
    if (cfg.monochrome) {
      // chi = 0, and luma = S for the split to hold (engine.cpp:178-182)
      HipCheck(hipMemset(w.chi.get(), 0, np * 2 * sizeof(float)), "hipMemset");
      HipCheck(hipMemcpy(w.y.get(), w.s.get(), np * sizeof(float), hipMemcpyDeviceToDevice), "hipMemcpy");
    }
    ss.check(hvd_gpu_chroma_plane(ctx, 0, w.s.get(), w.y.get(), nullptr, nullptr, np, w.chroma.get(), nullptr),
             "hvd_gpu_chroma_plane");
  }
  // ACC from both fields' bursts (engine.cpp:189-194)
  if (cfg.acc) {
    DevBuf<float> amp(2);
    ss.check(hvd_gpu_burst_amplitude_ire(ctx, fields, 2, ags, amp.get(), nullptr), "hvd_gpu_burst_amplitude_ire");
    float a[2] = {0.0F, 0.0F};
    amp.download(a, 2);
    w.acc_gain = AccGain(0.5F * (a[0] + a[1]), g.nominal_burst_ire());
  }
  return w;
}

namespace {

FrameYc ToHost(const WovenResult& w, bool chroma_only) {
  FrameYc out;
  const size_t np = (size_t)w.rows * w.width;
  out.composite = Plane(w.rows, w.width);
  out.luma = Plane(w.rows, w.width);
  out.chroma = Plane(w.rows, w.width);
  out.chroma_phasor = ComplexPlane(w.rows, w.width);
  out.carrier = ComplexPlane(w.rows, w.width);
  w.s.download(out.composite.data(), np);
  if (!chroma_only) w.y.download(out.luma.data(), np);
  w.chroma.download(out.chroma.data(), np);
  w.chi.download(reinterpret_cast<float*>(out.chroma_phasor.data()), np * 2);
  w.carrier.download(reinterpret_cast<float*>(out.carrier.data()), np * 2);
  out.acc_gain = w.acc_gain;
  return out;
}

FrameYc DecodeHostFields(Session& ss, const FieldInput& first, const FieldInput& second, const FieldGeometry& g,
                         const HvdConfig& cfg, bool chroma_only) {
  if (first.samples.height() != second.samples.height() || first.samples.width() != second.samples.width())
    throw Error("the two fields differ in size");
  DevBuf<float> a = Upload(first.samples.data(), first.samples.size());
  DevBuf<float> b = Upload(second.samples.data(), second.samples.size());
  DeviceFields f;
  f.first = a.get();
  f.second = b.get();
  f.first_is_first = first.is_first_field;
  f.second_is_first = second.is_first_field;
  f.field_lines = first.samples.height();
  f.field_width = first.samples.width();
  return ToHost(WovenDecode(ss, f, g, cfg, chroma_only), chroma_only);
}

}  // namespace
}  // namespace gpu

HvdEngine::HvdEngine() : fft_(std::make_unique<Fft2d>()), gpu_(std::make_unique<gpu::Session>()) {}
HvdEngine::~HvdEngine() = default;

gpu::Session& HvdEngine::session() { return *gpu_; }

void HvdEngine::SetFftThreads(int n) { fft_->SetThreadCount(n); }

FrameYc HvdEngine::DecodeFrame(const FieldInput& first, const FieldInput& second, const FieldGeometry& g,
                               const HvdConfig& cfg) {
  try {
    return gpu::DecodeHostFields(*gpu_, first, second, g, cfg, false);
  } catch (const std::exception& e) {
    gpu::Report("HvdEngine::DecodeFrame", e);
    return {};
  }
}

FrameYc HvdEngine::DecodeChromaOnly(const FieldInput& first, const FieldInput& second, const FieldGeometry& g,
                                    const HvdConfig& cfg) {
  try {
    return gpu::DecodeHostFields(*gpu_, first, second, g, cfg, true);
  } catch (const std::exception& e) {
    gpu::Report("HvdEngine::DecodeChromaOnly", e);
    return {};
  }
}

}  // namespace hvd
