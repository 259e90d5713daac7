// frame_bridge.cpp -- hvd-core's frame-buffer bridge (frame_bridge.cpp) over
// the GPU library.  Frames are uploaded once; split, field order, decode,
// weave and requantisation all run on the device.
#include "frame_bridge.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "woven.h"

namespace hvd {

using gpu::DevBuf;
using gpu::Session;

FieldGeometry FieldGeometryFromParams(const FrameParams& fp) {
  // frame_bridge.cpp:43-58: field geometry from the frame's, with the woven
  // active-line convention halved.
  FieldGeometry g;
  g.field_width = fp.frame_width;
  g.field_height = fp.field1_lines;
  g.active_video_start = fp.active_video_start;
  g.active_video_end = fp.active_video_end;
  g.colour_burst_start = fp.colour_burst_start;
  g.colour_burst_end = fp.colour_burst_end;
  g.first_active_field_line = fp.first_active_frame_line / 2;
  g.last_active_field_line = fp.last_active_frame_line != 0 ? fp.last_active_frame_line / 2 : 0;
  g.sample_rate = fp.sample_rate;
  g.standard = fp.standard;
  g.subcarrier_hz = fp.subcarrier_hz;
  return g;
}

namespace {

size_t FrameSamples(const FrameParams& fp) { return static_cast<size_t>(FrameLineOffset(fp, fp.frame_height)); }

void CheckFrame(const FrameParams& fp) {
  if (fp.frame_width <= 0 || fp.frame_height <= 0 || fp.field1_lines <= 0 || fp.field1_lines > fp.frame_height)
    throw gpu::Error("bad frame geometry");
}

struct DeviceSplit {
  DevBuf<float> block1, block2;
};

DeviceSplit Split(Session& ss, const int16_t* dframe, const FrameParams& fp, float zero) {
  DeviceSplit d;
  const size_t n = (size_t)fp.field1_lines * fp.frame_width;
  d.block1.resize(n);
  d.block2.resize(n);
  const hvd_frame_params a = gpu::ToAbi(fp);
  ss.check(hvd_gpu_split_frame_fields(ss.ctx(), dframe, &a, zero, d.block1.get(), d.block2.get(), nullptr),
           "hvd_gpu_split_frame_fields");
  return d;
}

int Vote(Session& ss, const float* b1, const float* b2, const FrameParams& fp) {
  const hvd_frame_params a = gpu::ToAbi(fp);
  int v = 0;
  ss.check(hvd_gpu_field_parity_vote(ss.ctx(), b1, b2, &a, &v, nullptr), "hvd_gpu_field_parity_vote");
  return v;
}

YcFrameS16 Output(const FrameParams& fp, size_t n, const DevBuf<int16_t>& lum, const DevBuf<int16_t>& chr,
                  const DevBuf<double>& u, const DevBuf<double>& v) {
  YcFrameS16 out;
  out.width = fp.frame_width;
  out.height = fp.frame_height;
  out.chroma_dc = fp.chroma_dc;
  out.luma.resize(n);
  out.chroma.resize(n);
  out.u_plane.resize(n);
  out.v_plane.resize(n);
  lum.download(out.luma.data(), n);
  chr.download(out.chroma.data(), n);
  u.download(out.u_plane.data(), n);
  v.download(out.v_plane.data(), n);
  return out;
}

// Staging bytes of one output frame of n samples: U and V (double), luma and
// chroma (int16), rounded up so every frame's U plane starts 64-byte aligned.
size_t StagedFrameBytes(size_t n) { return (n * (2 * sizeof(double) + 2 * sizeof(int16_t)) + 63) & ~size_t(63); }

// The outputs of a frame window through page-locked staging (`staging`
// holds ncore frames of U, V, luma and chroma planes): every frame's four
// copies are queued on the null stream behind the decode with an event after
// them, and one host thread per frame waits for its event and builds that
// frame's YcFrameS16 from the staging area while later frames' copies are in
// flight.
std::vector<YcFrameS16> StagedOutputs(const FrameParams& fp, size_t n, const std::vector<DevBuf<int16_t>>& lum,
                                      const std::vector<DevBuf<int16_t>>& chr, const std::vector<DevBuf<double>>& u,
                                      const std::vector<DevBuf<double>>& v, char* staging) {
  const int ncore = static_cast<int>(lum.size());
  const size_t fb = StagedFrameBytes(n);
  std::vector<hipEvent_t> ev(ncore, nullptr);
  auto release = [&] {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  try {
    for (int k = 0; k < ncore; ++k) {
      char* b = staging + fb * k;
      gpu::HipCheck(hipMemcpyAsync(b, u[k].get(), n * sizeof(double), hipMemcpyDeviceToHost, nullptr),
                    "hipMemcpyAsync D2H");
      gpu::HipCheck(hipMemcpyAsync(b + n * sizeof(double), v[k].get(), n * sizeof(double), hipMemcpyDeviceToHost,
                                   nullptr), "hipMemcpyAsync D2H");
      gpu::HipCheck(hipMemcpyAsync(b + 2 * n * sizeof(double), lum[k].get(), n * sizeof(int16_t),
                                   hipMemcpyDeviceToHost, nullptr), "hipMemcpyAsync D2H");
      gpu::HipCheck(hipMemcpyAsync(b + 2 * n * sizeof(double) + n * sizeof(int16_t), chr[k].get(),
                                   n * sizeof(int16_t), hipMemcpyDeviceToHost, nullptr), "hipMemcpyAsync D2H");
      gpu::HipCheck(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming), "hipEventCreateWithFlags");
      gpu::HipCheck(hipEventRecord(ev[k], nullptr), "hipEventRecord");
    }
  } catch (...) {
    (void)hipDeviceSynchronize();
    release();
    throw;
  }
  std::vector<YcFrameS16> out(ncore);
  std::vector<std::string> err(ncore);
  std::vector<std::thread> th;
  th.reserve(ncore);
  for (int k = 0; k < ncore; ++k)
    th.emplace_back([&, k] {
      const hipError_t e = hipEventSynchronize(ev[k]);
      if (e != hipSuccess) {
        err[k] = std::string("hipEventSynchronize: ") + hipGetErrorString(e);
        return;
      }
      try {
        const char* b = staging + fb * k;
        const double* pu = reinterpret_cast<const double*>(b);
        const int16_t* pl = reinterpret_cast<const int16_t*>(b + 2 * n * sizeof(double));
        YcFrameS16& o = out[k];
        o.width = fp.frame_width;
        o.height = fp.frame_height;
        o.chroma_dc = fp.chroma_dc;
        o.u_plane.assign(pu, pu + n);
        o.v_plane.assign(pu + n, pu + 2 * n);
        o.luma.assign(pl, pl + n);
        o.chroma.assign(pl + n, pl + 2 * n);
      } catch (const std::exception& x) {
        err[k] = x.what();
      }
    });
  for (std::thread& t : th) t.join();
  release();
  for (const std::string& e : err)
    if (!e.empty()) throw gpu::Error(e);
  return out;
}

// The woven decode of one frame re-woven into the flat layout
// (frame_bridge.cpp:161-255 and :263-377).
YcFrameS16 WovenFrame(Session& ss, const int16_t* passthrough_host, const gpu::DeviceFields& f, const FrameParams& fp,
                      const HvdConfig& cfg, bool chroma_only) {
  const FieldGeometry g = FieldGeometryFromParams(fp);
  gpu::WovenResult w = gpu::WovenDecode(ss, f, g, cfg, chroma_only);
  const size_t nflat = (size_t)fp.frame_width * fp.frame_height;
  DevBuf<int16_t> src = gpu::Upload(passthrough_host, FrameSamples(fp));
  DevBuf<int16_t> lum(nflat), chr(nflat);
  DevBuf<double> u(nflat), v(nflat);
  const hvd_yc_frame o{lum.get(), chr.get(), u.get(), v.get()};
  const hvd_frame_params a = gpu::ToAbi(fp);
  ss.check(hvd_gpu_weave_frame_s16(ss.ctx(), src.get(), &a, w.rows, w.width, chroma_only ? nullptr : w.y.get(),
                                   w.chroma.get(), w.chi.get(), cfg.chroma_gain * w.acc_gain, &o, nullptr),
           "hvd_gpu_weave_frame_s16");
  return Output(fp, nflat, lum, chr, u, v);
}

}  // namespace

void SplitFrameFields(const int16_t* frame, const FrameParams& fp, FieldInput* top, FieldInput* bottom,
                      bool field1_is_bottom) {
  try {
    CheckFrame(fp);
    Session& ss = gpu::ThreadSession();
    DevBuf<int16_t> df = gpu::Upload(frame, FrameSamples(fp));
    DeviceSplit d = Split(ss, df.get(), fp, fp.blanking_level);
    top->samples = Plane(fp.field1_lines, fp.frame_width);
    bottom->samples = Plane(fp.field1_lines, fp.frame_width);
    d.block1.download(top->samples.data(), top->samples.size());
    d.block2.download(bottom->samples.data(), bottom->samples.size());
    top->is_first_field = !field1_is_bottom;
    bottom->is_first_field = field1_is_bottom;
  } catch (const std::exception& e) {
    gpu::Report("SplitFrameFields", e);
    top->samples = Plane();
    bottom->samples = Plane();
  }
}

int DetectFieldParity(const FieldInput& block1, const FieldInput& block2, const FrameParams& fp) {
  try {
    CheckFrame(fp);
    const size_t n = (size_t)fp.field1_lines * fp.frame_width;
    if (block1.samples.size() != n || block2.samples.size() != n) throw gpu::Error("blocks do not match the frame");
    Session& ss = gpu::ThreadSession();
    DevBuf<float> a = gpu::Upload(block1.samples.data(), n), b = gpu::Upload(block2.samples.data(), n);
    const int v = Vote(ss, a.get(), b.get(), fp);
    return v > 0 ? 1 : (v < 0 ? -1 : 0);
  } catch (const std::exception& e) {
    gpu::Report("DetectFieldParity", e);
    return 0;
  }
}

bool ResolveField1IsBottom(const FieldInput& block1, const FieldInput& block2, const FrameParams& fp,
                           const HvdConfig& cfg) {
  if (cfg.field_order == 1) return false;
  if (cfg.field_order == 2) return true;
  return DetectFieldParity(block1, block2, fp) < 0;
}

YcFrameS16 DecodeFrameBuffer(const int16_t* frame, const FrameParams& fp, const HvdConfig& cfg, HvdEngine& engine) {
  try {
    CheckFrame(fp);
    Session& ss = engine.session();
    DevBuf<int16_t> df = gpu::Upload(frame, FrameSamples(fp));
    DeviceSplit d = Split(ss, df.get(), fp, fp.blanking_level);
    bool f1_bottom = cfg.field_order == 2;
    if (cfg.field_order == 0) f1_bottom = Vote(ss, d.block1.get(), d.block2.get(), fp) < 0;
    gpu::DeviceFields f;
    f.first = d.block1.get();
    f.second = d.block2.get();
    f.first_is_first = !f1_bottom;
    f.second_is_first = f1_bottom;
    f.field_lines = fp.field1_lines;
    f.field_width = fp.frame_width;
    return WovenFrame(ss, frame, f, fp, cfg, false);
  } catch (const std::exception& e) {
    gpu::Report("DecodeFrameBuffer", e);
    return {};
  }
}

YcFrameS16 DecodeFrameBuffer(const int16_t* frame, const FrameParams& fp, const HvdConfig& cfg) {
  HvdEngine engine;
  return DecodeFrameBuffer(frame, fp, cfg, engine);
}

YcFrameS16 DecodeYcFrameBuffer(const int16_t* luma, const int16_t* chroma, const FrameParams& fp,
                               const HvdConfig& cfg, HvdEngine& engine) {
  try {
    CheckFrame(fp);
    Session& ss = engine.session();
    DevBuf<int16_t> dc = gpu::Upload(chroma, FrameSamples(fp));
    // chroma re-centred on chroma_dc, on the IRE scale (frame_bridge.cpp:276-297)
    DeviceSplit d = Split(ss, dc.get(), fp, fp.chroma_dc);
    gpu::DeviceFields f;
    f.first = d.block1.get();
    f.second = d.block2.get();
    f.first_is_first = cfg.field_order != 2;
    f.second_is_first = cfg.field_order == 2;
    f.field_lines = fp.field1_lines;
    f.field_width = fp.frame_width;
    return WovenFrame(ss, luma, f, fp, cfg, true);
  } catch (const std::exception& e) {
    gpu::Report("DecodeYcFrameBuffer", e);
    return {};
  }
}

YcFrameS16 DecodeYcFrameBuffer(const int16_t* luma, const int16_t* chroma, const FrameParams& fp,
                               const HvdConfig& cfg) {
  HvdEngine engine;
  return DecodeYcFrameBuffer(luma, chroma, fp, cfg, engine);
}

std::vector<YcFrameS16> DecodeFrameSequenceWindow(const std::vector<const int16_t*>& frames, int core_begin,
                                                  int core_end, const FrameParams& fp, const HvdConfig& cfg,
                                                  HvdEngine& engine, int64_t /*base_frame_id*/) {
  const int nframes = static_cast<int>(frames.size());
  if (nframes == 0 || core_begin < 0 || core_end > nframes || core_begin >= core_end) return {};
  try {
    CheckFrame(fp);
    Session& ss = engine.session();
    const size_t nsrc = FrameSamples(fp), nflat = (size_t)fp.frame_width * fp.frame_height;
    const int ncore = core_end - core_begin;
    // Transfers go through the session's page-locked staging area: the frames
    // in (each frame's host copy overlaps the previous frame's DMA), then each
    // core frame's U, V, luma and chroma planes out.
    const size_t in_bytes = nsrc * sizeof(int16_t);
    const size_t out_bytes = StagedFrameBytes(nflat);
    const size_t in_total = (in_bytes * nframes + 63) & ~size_t(63);
    char* pin = ss.pinned(in_total + out_bytes * ncore);
    std::vector<DevBuf<int16_t>> df;
    std::vector<const int16_t*> dptr;
    for (int t = 0; t < nframes; ++t) {
      df.emplace_back(nsrc);
      dptr.push_back(df.back().get());
      char* st = pin + in_bytes * t;
      std::memcpy(st, frames[t], in_bytes);
      gpu::HipCheck(hipMemcpyAsync(df.back().get(), st, in_bytes, hipMemcpyHostToDevice, nullptr),
                    "hipMemcpyAsync H2D");
    }
    std::vector<DevBuf<int16_t>> lum, chr;
    std::vector<DevBuf<double>> u, v;
    std::vector<hvd_yc_frame> outs(ncore);
    for (int k = 0; k < ncore; ++k) {
      lum.emplace_back(nflat);
      chr.emplace_back(nflat);
      u.emplace_back(nflat);
      v.emplace_back(nflat);
      outs[k] = hvd_yc_frame{lum[k].get(), chr[k].get(), u[k].get(), v[k].get()};
    }
    const FieldGeometry fg = FieldGeometryFromParams(fp);
    ss.autotune(fg.active_lines(), fg.active_width(), fp.uses_vswitch(), 2 * nframes);
    const hvd_frame_params a = gpu::ToAbi(fp);
    const hvd_solver_params sp = gpu::ToAbi(cfg, fp.uses_vswitch());
    const hvd_bridge_params bp = gpu::BridgeOf(cfg);
    // the selective window is this entry's alone (frame_bridge.cpp:800-805)
    hvd_window_params wp = gpu::WindowOf(cfg);
    wp.selective_3d = cfg.selective_3d ? 1 : 0;
    wp.selective_max_area = cfg.selective_max_area;
    ss.check(hvd_gpu_decode_frame_window_ex_s16(ss.ctx(), dptr.data(), nframes, core_begin, core_end, &a, &sp, &bp,
                                                &wp, outs.data(), nullptr, nullptr, nullptr),
             "hvd_gpu_decode_frame_window_ex_s16");
    return StagedOutputs(fp, nflat, lum, chr, u, v, pin + in_total);
  } catch (const std::exception& e) {
    gpu::Report("DecodeFrameSequenceWindow", e);
    return {};
  }
}

}  // namespace hvd
