// gpu_session.cpp -- see gpu_session.h.
#include "gpu_session.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <utility>

#include "engine/gpu_status.h"
#include "frame_bridge.h"

namespace hvd {

namespace {
thread_local std::string g_last_error;
}  // namespace

std::string GpuLastError() { return g_last_error; }
void ClearGpuError() { g_last_error.clear(); }

namespace gpu {

Session::~Session() {
  if (pinned_) (void)hipHostFree(pinned_);
  if (ctx_) hvd_gpu_destroy(ctx_);
}

char* Session::pinned(size_t bytes) {
  if (bytes > pinned_cap_) {
    if (pinned_) HipCheck(hipHostFree(pinned_), "hipHostFree");
    pinned_ = nullptr;
    pinned_cap_ = 0;
    void* p = nullptr;
    HipCheck(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
    pinned_ = static_cast<char*>(p);
    pinned_cap_ = bytes;
  }
  return pinned_;
}

hvd_gpu_ctx* Session::ctx() {
  if (!ctx_) {
    int dev = 0;
    HipCheck(hipGetDevice(&dev), "hipGetDevice");
    hvd_gpu_ctx* c = nullptr;
    const int st = hvd_gpu_create(dev, &c);
    if (st != HVD_GPU_OK || !c) throw Error("no usable HIP device for the GPU library (status " + std::to_string(st) + ")");
    ctx_ = c;
  }
  return ctx_;
}

void Session::check(int status, const char* what) {
  if (status != HVD_GPU_OK)
    throw Error(std::string(what) + ": " + (ctx_ ? hvd_gpu_last_error(ctx_) : "no context") + " (status " +
                std::to_string(status) + ")");
}

void Session::sync() { check(hvd_gpu_synchronize(ctx(), nullptr), "hvd_gpu_synchronize"); }

void Session::autotune(int lines, int width, bool pal, int nfields) {
  static const bool on = [] {
    const char* e = std::getenv("HVD_AUTOTUNE");
    return e && *e && std::atoi(e) != 0;
  }();
  if (!on || nfields <= 0) return;
  const std::array<int, 4> key{lines, width, pal ? 1 : 0, nfields};
  if (!tuned_.insert(key).second) return;
  hvd_gpu_solver_tuning t{};
  const int st = hvd_gpu_autotune_solver(ctx(), lines, width, pal ? 1 : 0, nfields, 8, &t, nullptr);
  if (st != HVD_GPU_OK)
    std::fprintf(stderr, "hvd-gpu: autotune %dx%d x%d: %s\n", lines, width, nfields, hvd_gpu_last_error(ctx_));
}

Session& ThreadSession() {
  thread_local Session s;
  return s;
}

namespace {

struct BlockCache {
  std::mutex mu;
  std::map<std::pair<int, size_t>, std::vector<void*>> free;  // (device, class) -> blocks
  std::map<int, size_t> held;                                // cached bytes per device
};

BlockCache& Cache() {
  static BlockCache* c = new BlockCache;  // outlives every thread's buffers
  return *c;
}

size_t SizeClass(size_t bytes) {
  if (bytes <= 4096) return 4096;
  size_t p2 = 4096;
  while (p2 < bytes) p2 <<= 1;
  const size_t step = p2 / 16;  // classes of 1/8 of the power of two below
  return (bytes + step - 1) / step * step;
}

constexpr size_t kMaxCached = size_t(8) << 30;

}  // namespace

void* CachedMalloc(size_t bytes, size_t* cap, int* device) {
  int dev = 0;
  HipCheck(hipGetDevice(&dev), "hipGetDevice");
  *device = dev;
  const size_t c = SizeClass(bytes);
  BlockCache& bc = Cache();
  {
    std::lock_guard<std::mutex> lk(bc.mu);
    auto it = bc.free.find({dev, c});
    if (it != bc.free.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      bc.held[dev] -= c;
      *cap = c;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, c);
  if (e != hipSuccess) {
    // release what the cache holds on this device and try once more
    std::vector<void*> drop;
    {
      std::lock_guard<std::mutex> lk(bc.mu);
      for (auto& kv : bc.free)
        if (kv.first.first == dev) {
          drop.insert(drop.end(), kv.second.begin(), kv.second.end());
          kv.second.clear();
        }
      bc.held[dev] = 0;
    }
    (void)hipDeviceSynchronize();
    for (void* q : drop) (void)hipFree(q);
    e = hipMalloc(&p, c);
  }
  HipCheck(e, "hipMalloc");
  *cap = c;
  return p;
}

void CachedFree(void* p, size_t cap, int dev) {
  if (!p) return;
  BlockCache& bc = Cache();
  {
    std::lock_guard<std::mutex> lk(bc.mu);
    if (bc.held[dev] + cap <= kMaxCached) {
      bc.free[{dev, cap}].push_back(p);
      bc.held[dev] += cap;
      return;
    }
  }
  // hipFree acts on the block's own device whichever is current
  (void)hipFree(p);
}

void HipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

void Report(const char* api, const std::exception& e) {
  g_last_error = std::string(api) + ": " + e.what();
  std::fprintf(stderr, "hvd-gpu: %s\n", g_last_error.c_str());
}

hvd_field_geometry ToAbi(const FieldGeometry& g) {
  hvd_field_geometry r{};
  r.field_width = g.field_width;
  r.field_height = g.field_height;
  r.active_video_start = g.active_video_start;
  r.active_video_end = g.active_video_end;
  r.colour_burst_start = g.colour_burst_start;
  r.colour_burst_end = g.colour_burst_end;
  r.first_active_field_line = g.first_active_field_line;
  r.last_active_field_line = g.last_active_field_line;
  r.sample_rate = g.sample_rate;
  r.subcarrier_hz = g.subcarrier_hz;
  r.standard = static_cast<int>(g.standard);
  return r;
}

hvd_solver_params ToAbi(const HvdConfig& cfg, bool is_pal) {
  hvd_solver_params p{};
  p.lambda_c = cfg.lambda_c;
  p.chroma_aniso = cfg.chroma_aniso;
  p.charbonnier_eps = cfg.charbonnier_eps;
  p.chroma_eps = cfg.chroma_eps;
  p.structure_coupling = cfg.structure_coupling;
  p.diag_prior = cfg.diag_prior;
  p.cg_iterations = cfg.cg_iterations;
  p.irls_outer = cfg.irls_outer;
  p.cg_tol = cfg.cg_tol;
  p.fast = cfg.fast ? 1 : 0;
  p.is_pal = is_pal ? 1 : 0;
  p.lambda_y = 1.0;
  return p;
}

hvd_frame_params ToAbi(const FrameParams& fp) {
  hvd_frame_params r{};
  r.standard = static_cast<int>(fp.standard);
  r.frame_width = fp.frame_width;
  r.frame_height = fp.frame_height;
  r.field1_lines = fp.field1_lines;
  r.active_video_start = fp.active_video_start;
  r.active_video_end = fp.active_video_end;
  r.colour_burst_start = fp.colour_burst_start;
  r.colour_burst_end = fp.colour_burst_end;
  r.first_active_frame_line = fp.first_active_frame_line;
  r.last_active_frame_line = fp.last_active_frame_line;
  r.black_level = fp.black_level;
  r.white_level = fp.white_level;
  r.blanking_level = fp.blanking_level;
  r.chroma_dc = fp.chroma_dc;
  r.sample_rate = fp.sample_rate;
  r.subcarrier_hz = fp.subcarrier_hz;
  for (int i = 0; i < 8; ++i) r.extra_sample_lines[i] = fp.extra_sample_lines[i];
  r.extra_sample_line_count = fp.extra_sample_line_count;
  return r;
}

hvd_bridge_params BridgeOf(const HvdConfig& cfg) {
  hvd_bridge_params b{};
  b.chroma_gain = cfg.chroma_gain;
  b.acc = cfg.acc ? 1 : 0;
  b.field_order = cfg.field_order;
  b.monochrome = cfg.monochrome ? 1 : 0;
  b.chroma_phase_deg = cfg.chroma_phase_deg;
  b.temporal_strength = cfg.temporal_strength;
  return b;
}

hvd_window_params WindowOf(const HvdConfig& cfg) {
  hvd_window_params w{};
  w.temporal_strength = cfg.temporal_strength;
  w.temporal_eps = cfg.temporal_eps;
  w.nr_anchor = cfg.nr_anchor;
  w.nr_eps = cfg.nr_eps;
  w.coherence_gate = cfg.coherence_gate;
  w.odd_gate_floor = cfg.odd_gate_floor;
  w.passes = cfg.passes;
  w.nr_radius = cfg.nr_radius;
  w.mc_tile = cfg.mc_tile;
  w.mc_search = cfg.mc_search;
  w.bidirectional = cfg.bidirectional ? 1 : 0;
  w.extended_temporal = cfg.extended_temporal ? 1 : 0;
  w.trajectory_fit = cfg.trajectory_fit ? 1 : 0;
  w.psi_init = cfg.psi_init ? 1 : 0;
  w.output_fidelity = cfg.output_fidelity ? 1 : 0;
  w.parallel_across_fields = cfg.parallel_across_fields ? 1 : 0;
  return w;
}

bool WantsTemporal(const HvdConfig& cfg, int nfields, bool pal) {
  const int stride = pal ? 4 : 2;
  return cfg.cg_iterations > 0 &&
         (cfg.temporal_strength > 0.0F || (cfg.temporal_strength == 0.0F && nfields > stride));
}

}  // namespace gpu
}  // namespace hvd
