// abi.hip -- host implementation of the C ABI declared in include/hvd_gpu.h.
//
// A context owns: one grow-only device workspace carved per call, a ring of
// pinned-host/device descriptor slots (each guarded by an event so a slot is
// never rewritten while kernels of an earlier call still read it), and the
// FFT root tables per transform length.  Every entry point validates its
// arguments on the host before queueing any work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "hvd_gpu.h"
#include "hvd_kernels.h"
#include "window.h"

using namespace hvdk;

// Scratch of hvd_gpu_variational_refine_noalloc: a fixed device array that
// every GPU the library runs on holds until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: host code only
// ever uses the address hipGetSymbolAddress returns for the current device
// (kept per device by scratch_address(), looked up once at context creation).
static __device__ __align__(256) char scratch[HVD_GPU_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kSlots = 8;
constexpr double kPi = 3.14159265358979323846;
// The tail of the no-allocation scratch that holds a call's uploaded
// descriptors; the rest is that call's workspace.
constexpr size_t kNoAllocDescBytes = size_t(1) << 20;
constexpr int kNoAllocMaxPlanes = 256;

// Device address of `scratch` on each device, set once per device.
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_addr;

// With `device` current: the address of this device's copy of `scratch`.
hipError_t scratch_address(int device, char** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidSymbol;
    it = g_scratch_addr.emplace(device, static_cast<char*>(p)).first;
  }
  *out = it->second;
  return hipSuccess;
}

struct Slot {
  void* host = nullptr;
  void* dev = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool armed = false;
};

}  // namespace

// Launch shape of the solver: lines per strip and pixels per solve group
// (0: the built-in policy / default).
struct SolverTuning {
  int rows = 0;
  size_t group_px = 0;
};

struct hvd_gpu_ctx {
  int device = 0;
  std::string err;
  char* ws = nullptr;
  size_t ws_cap = 0;
  // this device's `scratch` (hipGetSymbolAddress at creation), and whether a
  // no-allocation call is running: then ws/ws_cap point into the scratch and
  // nothing may allocate device or pinned memory
  char* scratch = nullptr;
  bool no_alloc = false;
  // device memory of the 3-D window driver (window.hip), grow-only and
  // apart from ws, which the component calls it makes may reallocate
  char* win = nullptr;
  size_t win_cap = 0;
  // the 3-D frame window's field planes (lines x width per field: S, carrier,
  // Y, chi and the carrier tables), grow-only: they must outlive the window
  // driver's component calls for the same reason
  char* fwin = nullptr;
  size_t fwin_cap = 0;
  char* sel = nullptr;  // the selective window's inits, score scratch and crops
  size_t sel_cap = 0;
  Slot na_slot;  // descriptor region of a no-allocation call (no event)
  Slot slots[kSlots];
  int next_slot = 0;
  std::map<int, std::vector<float2>> roots;
  std::map<int, std::vector<float2>> rader_tabs;  // per prime radix
  int max_dyn_lds = 0;
  // profiling (diagnostics): event pairs around every CG loop and pinned
  // copies of the per-field solver state, read by hvd_gpu_cg_timing().
  bool profiling = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> cg_events;
  std::vector<bool> cg_events_owned;  // false: the events belong to a cached graph
  std::vector<long long> cg_iters;  // CG iterations launched per loop
  std::vector<std::pair<void*, int>> cg_states;  // pinned FieldCG copies, count
  // solver launch shape: fixed by hvd_gpu_set_solver_tuning, else measured per
  // field shape by hvd_gpu_autotune_solver, else the built-in policy
  SolverTuning fixed;
  std::map<std::array<int, 3>, std::map<int, SolverTuning>> tuned;  // (lines, width, pal) -> fields -> tuning
  // HIP graphs of a forward solve group's launches (hvd_gpu_set_graphs),
  // keyed by the launches they bake in (kernels, launch shapes, argument
  // bytes); dropped when the workspace they point into is reallocated.  A
  // graph built for profiling also holds event-record nodes around each
  // stage's CG iterations, with its own events.
  struct CachedGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    std::vector<long long> iters;
  };
  bool graphs = true;
  std::map<std::string, CachedGraph> cg_graphs;
  // events of dropped graphs, destroyed with the profiling record that may
  // still refer to them
  std::vector<hipEvent_t> retired;
};

namespace {

struct Fail {
  int code;
};

void set_err(hvd_gpu_ctx* c, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->err = buf;
}

#define HIPCHK(ctx, expr)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      set_err(ctx, "%s failed: %s", #expr, hipGetErrorString(e_));                            \
      throw Fail{HVD_GPU_ERR_HIP};                                                            \
    }                                                                                         \
  } while (0)

#define REQUIRE(ctx, cond, ...)            \
  do {                                     \
    if (!(cond)) {                         \
      set_err(ctx, __VA_ARGS__);           \
      throw Fail{HVD_GPU_ERR_INVALID};     \
    }                                      \
  } while (0)

#define LAUNCH_CHECK(ctx) HIPCHK(ctx, hipGetLastError())

template <typename F>
int guarded(hvd_gpu_ctx* ctx, F&& fn) {
  if (!ctx) return HVD_GPU_ERR_INVALID;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) prev = ctx->device;
  if (hipSetDevice(ctx->device) != hipSuccess) {
    set_err(ctx, "hipSetDevice(%d) failed", ctx->device);
    return HVD_GPU_ERR_HIP;
  }
  int rc = HVD_GPU_OK;
  try {
    fn();
    ctx->err.clear();
  } catch (const Fail& f) {
    rc = f.code;
  } catch (const std::bad_alloc&) {
    set_err(ctx, "host allocation failed");
    rc = HVD_GPU_ERR_NOMEM;
  }
  if (prev != ctx->device) (void)hipSetDevice(prev);
  return rc;
}

inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// ---- workspace ------------------------------------------------------------
struct Carve {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  }
};

// Destroys the cached graphs (their arguments point into the workspace); the
// caller has synchronised the device.  Their profiling events live on until
// the profiling record is cleared.
void drop_graphs(hvd_gpu_ctx* ctx) {
  for (auto& kv : ctx->cg_graphs) {
    (void)hipGraphExecDestroy(kv.second.exec);
    for (auto& e : kv.second.events) {
      ctx->retired.push_back(e.first);
      ctx->retired.push_back(e.second);
    }
  }
  ctx->cg_graphs.clear();
}

char* ensure_ws(hvd_gpu_ctx* ctx, size_t bytes, hipStream_t st) {
  if (bytes <= ctx->ws_cap) return ctx->ws;
  if (ctx->no_alloc) {
    set_err(ctx, "the call needs %zu bytes of device scratch; the no-allocation entry point holds at most %zu", bytes,
            ctx->ws_cap);
    throw Fail{HVD_GPU_ERR_INVALID};
  }
  HIPCHK(ctx, hipStreamSynchronize(st));
  HIPCHK(ctx, hipDeviceSynchronize());
  drop_graphs(ctx);
  if (ctx->ws) HIPCHK(ctx, hipFree(ctx->ws));
  ctx->ws = nullptr;
  ctx->ws_cap = 0;
  const size_t want = std::max(bytes + bytes / 4, size_t(64) << 20);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    set_err(ctx, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    throw Fail{HVD_GPU_ERR_NOMEM};
  }
  ctx->ws = static_cast<char*>(p);
  ctx->ws_cap = want;
  return ctx->ws;
}

// ---- descriptor staging ----------------------------------------------------
struct Stage {
  std::vector<char> buf;
  template <typename T>
  size_t put(const T* p, size_t n) {
    const size_t o = align256(buf.size());
    buf.resize(o + sizeof(T) * n);
    if (n) std::memcpy(buf.data() + o, p, sizeof(T) * n);
    return o;
  }
  template <typename T>
  size_t put(const std::vector<T>& v) {
    return put(v.data(), v.size());
  }
};

// Uploads the staged bytes through the next slot; returns the device base.
// `extra` more device bytes follow the staged ones, from
// align256(stg.buf.size()), for the launches that read the slot (the slot is
// reused only after they finish).
char* upload(hvd_gpu_ctx* ctx, const Stage& stg, hipStream_t st, Slot** used, size_t extra = 0) {
  if (ctx->no_alloc) {
    // the descriptor region at the scratch's tail; calls on one stream run in
    // order, so a later upload never overwrites what an earlier launch reads
    REQUIRE(ctx, align256(stg.buf.size()) + extra <= ctx->na_slot.cap,
            "descriptors exceed the no-allocation region (%zu bytes)", ctx->na_slot.cap);
    if (!stg.buf.empty())
      HIPCHK(ctx, hipMemcpyAsync(ctx->na_slot.dev, stg.buf.data(), stg.buf.size(), hipMemcpyHostToDevice, st));
    *used = &ctx->na_slot;
    return static_cast<char*>(ctx->na_slot.dev);
  }
  Slot& sl = ctx->slots[ctx->next_slot];
  ctx->next_slot = (ctx->next_slot + 1) % kSlots;
  if (sl.armed) HIPCHK(ctx, hipEventSynchronize(sl.ev));
  const size_t need = std::max<size_t>(align256(stg.buf.size()) + extra, 1);
  if (need > sl.cap) {
    if (sl.host) HIPCHK(ctx, hipHostFree(sl.host));
    if (sl.dev) HIPCHK(ctx, hipFree(sl.dev));
    sl.host = sl.dev = nullptr;
    const size_t cap = std::max(need * 2, size_t(1) << 16);
    HIPCHK(ctx, hipHostMalloc(&sl.host, cap, hipHostMallocDefault));
    HIPCHK(ctx, hipMalloc(&sl.dev, cap));
    sl.cap = cap;
  }
  std::memcpy(sl.host, stg.buf.data(), stg.buf.size());
  HIPCHK(ctx, hipMemcpyAsync(sl.dev, sl.host, stg.buf.size(), hipMemcpyHostToDevice, st));
  *used = &sl;
  return static_cast<char*>(sl.dev);
}

// As upload, but into a caller-chosen device region (a fixed place in the
// workspace, so launches that read the descriptors keep the same arguments
// from call to call); the slot only lends its pinned host buffer.
char* upload_into(hvd_gpu_ctx* ctx, const Stage& stg, char* dst, hipStream_t st, Slot** used) {
  if (ctx->no_alloc) {
    // straight from the (pageable) staging buffer: the copy is taken from it
    // before hipMemcpyAsync returns
    if (!stg.buf.empty()) HIPCHK(ctx, hipMemcpyAsync(dst, stg.buf.data(), stg.buf.size(), hipMemcpyHostToDevice, st));
    *used = &ctx->na_slot;
    return dst;
  }
  Slot& sl = ctx->slots[ctx->next_slot];
  ctx->next_slot = (ctx->next_slot + 1) % kSlots;
  if (sl.armed) HIPCHK(ctx, hipEventSynchronize(sl.ev));
  const size_t need = std::max<size_t>(stg.buf.size(), 1);
  if (need > sl.cap) {
    if (sl.host) HIPCHK(ctx, hipHostFree(sl.host));
    if (sl.dev) HIPCHK(ctx, hipFree(sl.dev));
    sl.host = sl.dev = nullptr;
    const size_t cap = std::max(need * 2, size_t(1) << 16);
    HIPCHK(ctx, hipHostMalloc(&sl.host, cap, hipHostMallocDefault));
    HIPCHK(ctx, hipMalloc(&sl.dev, cap));
    sl.cap = cap;
  }
  std::memcpy(sl.host, stg.buf.data(), stg.buf.size());
  HIPCHK(ctx, hipMemcpyAsync(dst, sl.host, stg.buf.size(), hipMemcpyHostToDevice, st));
  *used = &sl;
  return dst;
}

void arm(hvd_gpu_ctx* ctx, Slot* sl, hipStream_t st) {
  if (!sl->ev) return;  // the no-allocation region is guarded by stream order
  HIPCHK(ctx, hipEventRecord(sl->ev, st));
  sl->armed = true;
}

// ---- FFT plans --------------------------------------------------------------
std::vector<int> factorize(int n) {
  std::vector<int> f;
  for (int r : {8, 4, 2, 3, 5, 7})
    while (n > 1 && n % r == 0) {
      f.push_back(r);
      n /= r;
    }
  for (int p = 11; n > 1; p += 2)
    while (n % p == 0) {
      f.push_back(p);
      n /= p;
    }
  return f;
}

const std::vector<float2>& root_table(hvd_gpu_ctx* ctx, int n) {
  auto it = ctx->roots.find(n);
  if (it != ctx->roots.end()) return it->second;
  std::vector<float2> r(n);
  for (int j = 0; j < n; ++j) {
    const double a = -2.0 * kPi * (double)j / (double)n;
    r[j] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  return ctx->roots.emplace(n, std::move(r)).first->second;
}

// Smallest generator of the multiplicative group mod the prime q.
int primitive_root(int q) {
  std::vector<int> pf;
  int m = q - 1;
  for (int d = 2; d * d <= m; ++d)
    if (m % d == 0) {
      pf.push_back(d);
      while (m % d == 0) m /= d;
    }
  if (m > 1) pf.push_back(m);
  for (int g = 2; g < q; ++g) {
    bool ok = true;
    for (int f : pf) {
      long long r = 1, b = g;
      for (int e = (q - 1) / f; e; e >>= 1, b = b * b % q)
        if (e & 1) r = r * b % q;
      if (r == 1) ok = false;
    }
    if (ok) return g;
  }
  return 1;
}

// Rader tables of prime R (L = R - 1): the L roots of L, bhat for the forward
// and the inverse direction (bhat = DFT_L(b) / L with b_l = w^{g^-l},
// w = e^{-+2 pi i / R}), and the permutation {g^l, g^-l} mod R stored as int
// pairs in float2 slots; all in FP64, rounded once.
const std::vector<float2>& rader_table(hvd_gpu_ctx* ctx, int R) {
  auto it = ctx->rader_tabs.find(R);
  if (it != ctx->rader_tabs.end()) return it->second;
  const int L = R - 1, g = primitive_root(R);
  std::vector<int> gp(L), gm(L);
  long long x = 1;
  for (int l = 0; l < L; ++l, x = x * g % R) gp[l] = (int)x;
  for (int l = 0; l < L; ++l) gm[l] = gp[(L - l) % L];  // g^-l = g^(L - l)
  std::vector<float2> t(4 * (size_t)L);
  for (int j = 0; j < L; ++j) {
    const double a = -2.0 * kPi * (double)j / (double)L;
    t[j] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  for (int dir = 0; dir < 2; ++dir) {
    const double sgn = dir == 0 ? -1.0 : 1.0;
    for (int k = 0; k < L; ++k) {
      double re = 0.0, im = 0.0;
      for (int l = 0; l < L; ++l) {
        // b_l e^{-2 pi i k l / L}, phases reduced exactly before scaling
        const double a = sgn * 2.0 * kPi * (double)gm[l] / (double)R -
                         2.0 * kPi * (double)((long long)k * l % L) / (double)L;
        re += std::cos(a);
        im += std::sin(a);
      }
      t[(size_t)(1 + dir) * L + k] = make_float2((float)(re / L), (float)(im / L));
    }
  }
  for (int l = 0; l < L; ++l) {
    int2 pr{gp[l], gm[l]};
    std::memcpy(&t[(size_t)3 * L + l], &pr, sizeof(pr));
  }
  return ctx->rader_tabs.emplace(R, std::move(t)).first->second;
}

// FFT plans of one call, keyed by (length, interleaved sequences per block),
// with their root and Rader tables appended to one buffer.  A prime stage
// R >= kRaderMin runs by Rader's algorithm when its tables fit the block's
// LDS, else as the direct pair form.
struct PlanSet {
  std::vector<FftPlan> plans;
  std::vector<float2> roots;
  std::map<std::pair<int, int>, int> index;
  int get(hvd_gpu_ctx* ctx, int n, int ncols) {
    auto it = index.find({n, ncols});
    if (it != index.end()) return it->second;
    FftPlan p{};
    p.n = n;
    const std::vector<int> f = factorize(n);
    REQUIRE(ctx, (int)f.size() <= kMaxRadix, "FFT length %d has too many factors", n);
    p.nst = (int)f.size();
    for (size_t i = 0; i < f.size(); ++i) {
      p.radix[i] = f[i];
      p.rader_of[i] = -1;
    }
    p.root_off = (int)roots.size();
    const std::vector<float2>& r = root_table(ctx, n);
    roots.insert(roots.end(), r.begin(), r.end());
    FftPlan pr = p;  // with Rader stages
    for (size_t i = 0; i < f.size(); ++i) {
      const int R = f[i];
      if (R < kRaderMin || pr.nrader == kMaxRader) continue;
      const std::vector<int> fi = factorize(R - 1);
      if ((int)fi.size() > kMaxInnerRadix) continue;
      FftRader& rd = pr.rader[pr.nrader];
      rd.L = R - 1;
      rd.nst = (int)fi.size();
      for (size_t j = 0; j < fi.size(); ++j) rd.radix[j] = fi[j];
      pr.rader_of[i] = (signed char)pr.nrader++;
      pr.rader_fixed = std::max(pr.rader_fixed, 3 * rd.L);
      pr.rader_per_col = std::max(pr.rader_per_col, 2 * (n / R));
    }
    if (pr.nrader > 0 && sizeof(float2) * (size_t)fft_lds_slots(pr, ncols) <= (size_t)ctx->max_dyn_lds) {
      for (int k = 0; k < pr.nrader; ++k) {
        pr.rader[k].tab = (int)roots.size();
        const std::vector<float2>& t = rader_table(ctx, pr.rader[k].L + 1);
        roots.insert(roots.end(), t.begin(), t.end());
      }
      p = pr;
    }
    plans.push_back(p);
    const int id = (int)plans.size() - 1;
    index[{n, ncols}] = id;
    return id;
  }
};

// ---- geometry helpers --------------------------------------------------------
double phase_per_sample(const hvd_field_geometry& g) {
  if (g.subcarrier_hz <= 0.0) return 0.5 * kPi;
  return 2.0 * kPi * g.subcarrier_hz / g.sample_rate;
}
double line_advance(const hvd_field_geometry& g) {
  if (g.subcarrier_hz <= 0.0) {
    if (g.standard == 1) return 2.0 * kPi * 0.7516;
    if (g.standard == 2) return 0.5 * kPi;
    return kPi;
  }
  const double cyc = g.subcarrier_hz * (double)g.field_width / g.sample_rate;
  return 2.0 * kPi * (cyc - std::floor(cyc));
}
bool uses_vswitch(const hvd_field_geometry& g) { return g.standard == 1 || g.standard == 2; }
int last_line(const hvd_field_geometry& g) {
  return g.last_active_field_line != 0 ? g.last_active_field_line : g.field_height;
}

void check_geom(hvd_gpu_ctx* ctx, const hvd_field_geometry& g, int i) {
  REQUIRE(ctx, g.field_width > 0 && g.field_height > 0, "field %d: empty geometry", i);
  REQUIRE(ctx, g.field_height <= kMaxLines, "field %d: %d lines exceeds %d", i, g.field_height, kMaxLines);
  REQUIRE(ctx, g.colour_burst_start >= 0 && g.colour_burst_end <= g.field_width, "field %d: burst window outside the line", i);
  REQUIRE(ctx, g.colour_burst_end - g.colour_burst_start <= 67, "field %d: burst window longer than 67 samples", i);
  REQUIRE(ctx, g.standard >= 0 && g.standard <= 2, "field %d: unknown standard %d", i, g.standard);
  REQUIRE(ctx, !(g.subcarrier_hz > 0.0) || g.sample_rate > 0.0, "field %d: custom subcarrier needs a sample rate", i);
}

void check_active(hvd_gpu_ctx* ctx, const hvd_field_geometry& g, int i) {
  const int lal = last_line(g);
  REQUIRE(ctx, g.first_active_field_line >= 0 && lal <= g.field_height && lal - g.first_active_field_line >= 1,
          "field %d: bad active line range", i);
  REQUIRE(ctx, g.active_video_start >= 0 && g.active_video_end <= g.field_width &&
                   g.active_video_end - g.active_video_start >= 1,
          "field %d: bad active sample range", i);
}

LockinGeom lockin_geom(const hvd_field_geometry& g) {
  LockinGeom L{};
  L.fw = g.field_width;
  L.fh = g.field_height;
  L.x0 = g.colour_burst_start;
  L.x1 = g.colour_burst_end;
  L.std_phase = g.subcarrier_hz <= 0.0;
  L.vswitch = uses_vswitch(g);
  L.w0 = phase_per_sample(g);
  L.line_advance = line_advance(g);
  L.nominal_burst = g.standard == 1 ? 21.43f : 20.0f;
  return L;
}

// FieldGeometryFromParams (frame_bridge.cpp:43-58).
hvd_field_geometry frame_field_geometry(const hvd_frame_params& fp) {
  hvd_field_geometry g{};
  g.field_width = fp.frame_width;
  g.field_height = fp.field1_lines;
  g.active_video_start = fp.active_video_start;
  g.active_video_end = fp.active_video_end;
  g.colour_burst_start = fp.colour_burst_start;
  g.colour_burst_end = fp.colour_burst_end;
  g.first_active_field_line = fp.first_active_frame_line / 2;
  g.last_active_field_line = fp.last_active_frame_line != 0 ? fp.last_active_frame_line / 2 : 0;
  g.sample_rate = fp.sample_rate;
  g.subcarrier_hz = fp.subcarrier_hz;
  g.standard = fp.standard;
  return g;
}

void check_frame(hvd_gpu_ctx* ctx, const hvd_frame_params& fp) {
  const int fw = fp.frame_width, fh = fp.frame_height, f1 = fp.field1_lines;
  REQUIRE(ctx, fw > 0 && fh > 0 && f1 > 0 && f1 <= fh && fh - f1 <= f1, "bad frame geometry %d x %d (field 1: %d)",
          fw, fh, f1);
  REQUIRE(ctx, fp.extra_sample_line_count >= 0 && fp.extra_sample_line_count <= 8, "bad extra-sample table");
  for (int i = 0; i < fp.extra_sample_line_count; ++i)
    REQUIRE(ctx, fp.extra_sample_lines[i] >= 0 && fp.extra_sample_lines[i] < fh, "extra-sample line %d outside the frame",
            fp.extra_sample_lines[i]);
  REQUIRE(ctx, fp.white_level != fp.blanking_level, "white level equals blanking");
}

BridgeFrameParams bridge_params(const hvd_frame_params& fp) {
  BridgeFrameParams bp{};
  bp.standard = fp.standard;
  bp.frame_width = fp.frame_width;
  bp.frame_height = fp.frame_height;
  bp.field1_lines = fp.field1_lines;
  bp.active_video_start = fp.active_video_start;
  bp.active_video_end = fp.active_video_end;
  bp.colour_burst_start = fp.colour_burst_start;
  bp.colour_burst_end = fp.colour_burst_end;
  bp.first_active_frame_line = fp.first_active_frame_line;
  bp.last_active_frame_line = fp.last_active_frame_line;
  bp.black_level = fp.black_level;
  bp.white_level = fp.white_level;
  bp.blanking_level = fp.blanking_level;
  bp.chroma_dc = fp.chroma_dc;
  bp.sample_rate = fp.sample_rate;
  bp.subcarrier_hz = fp.subcarrier_hz;
  for (int i = 0; i < 8; ++i) bp.extra_lines[i] = fp.extra_sample_lines[i];
  bp.extra_count = fp.extra_sample_line_count;
  return bp;
}

// ---- launch sequences -------------------------------------------------------------

void run_lockin(hvd_gpu_ctx* ctx, const std::vector<LockinGeom>& G, const std::vector<const float*>& fields,
                const std::vector<float*>& theta, const std::vector<int8_t*>& vsw, hipStream_t st) {
  const int n = (int)G.size();
  std::vector<int> lp(n + 1, 0);  // line prefix over the batch
  for (int i = 0; i < n; ++i) lp[i + 1] = lp[i] + G[i].fh;
  Stage stg;
  const size_t oG = stg.put(G), oF = stg.put(fields), oT = stg.put(theta), oV = stg.put(vsw), oL = stg.put(lp);
  Slot* sl;
  // the per-line demodulation (zr, zi, |z|) in the slot's scratch tail
  char* base = upload(ctx, stg, st, &sl, sizeof(double) * 3 * (size_t)std::max(lp[n], 1));
  double* z = (double*)(base + align256(stg.buf.size()));
  const int* dlp = (const int*)(base + oL);
  if (lp[n] > 0) {
    hipLaunchKernelGGL(k_lockin_demod, dim3((lp[n] + 3) / 4), dim3(256), 0, st, (const LockinGeom*)(base + oG),
                       (const float* const*)(base + oF), dlp, n, z);
    LAUNCH_CHECK(ctx);
  }
  hipLaunchKernelGGL(k_lockin, dim3(n), dim3(256), 0, st, (const LockinGeom*)(base + oG), (const double*)z, dlp,
                     (float* const*)(base + oT), (int8_t* const*)(base + oV));
  LAUNCH_CHECK(ctx);
  arm(ctx, sl, st);
}

// Tables (and optionally the active crop) for a batch.
void run_prepare(hvd_gpu_ctx* ctx, const std::vector<PrepDesc>& P, bool crop, hipStream_t st) {
  if (P.empty()) return;
  Stage stg;
  const size_t oP = stg.put(P);
  Slot* sl;
  char* base = upload(ctx, stg, st, &sl);
  int mx = 1, mlines = 1, mwidth = 1;
  for (const PrepDesc& p : P) {
    mx = std::max(mx, std::max(p.lines, p.width));
    mlines = std::max(mlines, p.lines);
    mwidth = std::max(mwidth, p.width);
  }
  hipLaunchKernelGGL(k_prepare, dim3((mx + 255) / 256, (unsigned)P.size()), dim3(256), 0, st,
                     (const PrepDesc*)(base + oP));
  LAUNCH_CHECK(ctx);
  if (crop) {
    hipLaunchKernelGGL(k_crop, dim3((mwidth + 255) / 256, mlines, (unsigned)P.size()), dim3(256), 0, st,
                       (const PrepDesc*)(base + oP));
    LAUNCH_CHECK(ctx);
  }
  arm(ctx, sl, st);
}

void run_init(hvd_gpu_ctx* ctx, std::vector<InitField>& F, PlanSet& ps, hipStream_t st) {
  const int nf = (int)F.size();
  // per FFT pass and transform class: the fields, the running count of
  // blocks, the LDS a block needs
  struct PassList {
    std::vector<int> fields, prefix{0};
    size_t lds = 0, o_fields = 0, o_prefix = 0;
  };
  PassList L[4][3];
  std::vector<int> tiles(nf + 1, 0);
  for (int i = 0; i < nf; ++i) {
    const InitField& f = F[i];
    const FftPlan& pr = ps.plans[f.row_plan];
    const FftPlan& pc = ps.plans[f.col_plan];
    const int rc = fft_init_class(pr, true), cc = fft_init_class(pc, false);
    const int rows_fwd = (f.hp + kRowGroup - 1) / kRowGroup, cols = (f.w + kColGroup - 1) / kColGroup,
              rows_inv = (f.h + kRowGroup - 1) / kRowGroup;
    const size_t lds_r = sizeof(float2) * (size_t)fft_lds_slots(pr, kRowGroup),
                 lds_c = sizeof(float2) * (size_t)fft_lds_slots(pc, kColGroup),
                 lds_ci = lds_c + sizeof(float) * (size_t)f.hp;  // + the Gaussian's row factors
    REQUIRE(ctx, lds_r <= (size_t)ctx->max_dyn_lds && lds_ci <= (size_t)ctx->max_dyn_lds,
            "FFT sizes %d x %d exceed the LDS budget (%zu / %zu bytes per block)", f.hp, f.w, lds_ci, lds_r);
    const int cls[4] = {rc, cc, cc, rc}, blocks[4] = {rows_fwd, cols, cols, rows_inv};
    const size_t lds[4] = {lds_r, lds_c, lds_ci, lds_r};
    for (int p = 0; p < 4; ++p) {
      PassList& pl = L[p][cls[p]];
      pl.fields.push_back(i);
      pl.prefix.push_back(pl.prefix.back() + blocks[p]);
      pl.lds = std::max(pl.lds, lds[p]);
    }
    tiles[i + 1] = tiles[i] + ((f.w + kBlendW - 1) / kBlendW) * ((f.h + kBlendH - 1) / kBlendH);
  }
  Stage stg;
  const size_t oF = stg.put(F), oP = stg.put(ps.plans), oR = stg.put(ps.roots), o4 = stg.put(tiles);
  for (auto& pass : L)
    for (PassList& pl : pass)
      if (!pl.fields.empty()) {
        pl.o_fields = stg.put(pl.fields);
        pl.o_prefix = stg.put(pl.prefix);
      }
  Slot* sl;
  char* base = upload(ctx, stg, st, &sl);
  const InitField* dF = (const InitField*)(base + oF);
  const FftPlan* dP = (const FftPlan*)(base + oP);
  const float2* dR = (const float2*)(base + oR);
  for (int p = 0; p < 4; ++p)
    for (int c = 0; c < 3; ++c) {
      const PassList& pl = L[p][c];
      if (pl.fields.empty()) continue;
      launch_fft_pass(p, c, pl.prefix.back(), pl.lds, st, dF, dP, dR, (const int*)(base + pl.o_prefix),
                      (const int*)(base + pl.o_fields), (int)pl.fields.size());
      LAUNCH_CHECK(ctx);
    }
  hipLaunchKernelGGL(k_holo_blend, dim3(tiles[nf]), dim3(256), 0, st, dF, (const int*)(base + o4), nf);
  LAUNCH_CHECK(ctx);
  arm(ctx, sl, st);
}

struct RefineOut {
  int* stage_iters = nullptr;
  float* aniso = nullptr;
};

constexpr int kMinStripRows = 2;

// Lines per solver strip: the tallest strips (the fewest re-read halo rows)
// that still give a launch enough waves to keep every CU busy; `forced` > 0
// (a tuning) wins.
int strip_rows(const std::vector<DevField>& F, int forced) {
  if (forced > 0) return std::min(std::max(forced, kMinStripRows), kMaxStripRows);
  // measured on NTSC fields in 16-field solve groups: 16-line strips are
  // fastest once a launch has ~2 waves per SIMD; below that, shorter strips
  auto waves = [&](int rows) {
    long n = 0;
    for (const DevField& f : F) n += solver_strips(f.h, f.w, rows);
    return n;
  };
  for (int rows : {16, 8, 4})
    if (waves(rows) >= 2048) return rows;
  return kMinStripRows;
}

// Strip decomposition of a batch for the solver kernels; fills each field's
// tiles_x and tile0.
struct StripSetup {
  std::vector<int> tp;
  int nb = 0, rows = 0, spf = 0;
};

StripSetup strip_setup(std::vector<DevField>& F, int rows) {
  StripSetup S;
  const int nf = (int)F.size();
  S.rows = strip_rows(F, rows);
  S.tp.assign(nf + 1, 0);
  for (int i = 0; i < nf; ++i) {
    F[i].tiles_x = (F[i].w + kStripCols - 1) / kStripCols;
    F[i].tile0 = S.tp[i];
    S.tp[i + 1] = S.tp[i] + solver_strips(F[i].h, F[i].w, S.rows);
  }
  S.nb = S.tp[nf];
  S.spf = S.tp[1];
  for (int i = 1; i < nf; ++i)
    if (F[i].h != F[0].h || F[i].w != F[0].w) S.spf = 0;
  return S;
}

// A recorded sequence of kernel launches (and, when profiling, event marks
// around each CG loop): launched eagerly, or built into a HIP graph of
// kernel and event-record nodes (no memset nodes) that is cached under the
// bytes of the whole sequence -- kernels, launch shapes and argument values --
// and replayed.  Every argument of the solver's launches points into the
// workspace or is a value, so an equal sequence is an equal graph.
struct LaunchList {
  struct Item {
    const void* func;  // null: an event mark
    dim3 grid, block;
    unsigned shmem;
    std::vector<size_t> args;  // offsets of the argument copies in `bytes`
    int mark;                  // event marks: 2 * loop (start) or 2 * loop + 1 (end)
  };
  std::vector<Item> items;
  std::vector<char> bytes;
  std::vector<long long> loop_iters;  // CG iterations inside each marked loop
  std::string key;                    // the sequence as bytes

  template <typename... A>
  void kernel(const void* f, dim3 grid, dim3 block, unsigned shmem, A... a) {
    Item it{f, grid, block, shmem, {}, -1};
    append_key(&f, sizeof f);
    const unsigned dims[7] = {grid.x, grid.y, grid.z, block.x, block.y, block.z, shmem};
    append_key(dims, sizeof dims);
    (put(it, a), ...);
    items.push_back(std::move(it));
  }
  // marks the start (end = false) or end of CG loop `loop`
  void mark(int loop, bool end) {
    items.push_back(Item{nullptr, dim3(), dim3(), 0, {}, 2 * loop + (end ? 1 : 0)});
    const int m = items.back().mark;
    append_key(&m, sizeof m);
  }
  template <typename T>
  void put(Item& it, const T& v) {
    const size_t off = (bytes.size() + 15) & ~size_t(15);
    bytes.resize(off + sizeof(T));
    std::memcpy(bytes.data() + off, &v, sizeof(T));
    it.args.push_back(off);
    append_key(&v, sizeof(T));
  }
  void append_key(const void* p, size_t n) { key.append(static_cast<const char*>(p), n); }
  std::vector<void*> argv(const Item& it) {
    std::vector<void*> v;
    for (size_t off : it.args) v.push_back(bytes.data() + off);
    return v;
  }
};

// Records one forward or adjoint CG loop (variational.cpp:519-549): the
// stage-start pass, n_inner fused iterations and the stage's close, the
// iterations and close marked as profiling loop `loop`.
void record_cg_loop(LaunchList& L, const DevField* dF, FieldCG* dstate, const StripGrid& sg, double* partials,
                    int nf, int n_inner, bool adjoint, int stage, int loop) {
  L.kernel((const void*)k_cg_start, dim3(sg.nb), dim3(64), 0, dF, (const FieldCG*)dstate, sg, partials, (int)adjoint,
           stage);
  L.mark(loop, false);
  for (int it = 0; it < n_inner; ++it)
    L.kernel(adjoint ? (const void*)k_cg_iter_adj : (const void*)k_cg_iter, dim3(sg.nb), dim3(64), 0, dF, dstate, sg,
             partials, it);
  L.kernel((const void*)k_cg_finish, dim3(nf), dim3(64), 0, dstate, sg, (const double*)partials, n_inner, stage);
  L.mark(loop, true);
  L.loop_iters.push_back(n_inner);
}

// Runs a recorded sequence on `st`: as a cached graph when `graph`, else
// launch by launch.  When profiling, the marked loops' event pairs and
// iteration counts go to the profiling record.
void run_launches(hvd_gpu_ctx* ctx, LaunchList& L, bool graph, hipStream_t st) {
  const bool prof = ctx->profiling;
  const size_t nloops = L.loop_iters.size();
  if (graph) {
    const std::string key = (prof ? "P" : "N") + L.key;
    auto found = ctx->cg_graphs.find(key);
    if (found == ctx->cg_graphs.end()) {
      if (ctx->cg_graphs.size() >= 64) {  // bounded cache
        HIPCHK(ctx, hipDeviceSynchronize());
        drop_graphs(ctx);
      }
      hvd_gpu_ctx::CachedGraph cg;
      hipGraph_t g = nullptr;
      HIPCHK(ctx, hipGraphCreate(&g, 0));
      struct Free {
        hipGraph_t g;
        ~Free() { (void)hipGraphDestroy(g); }
      } guard{g};
      struct FreeEvents {
        std::vector<std::pair<hipEvent_t, hipEvent_t>>* evs;
        bool keep = false;
        ~FreeEvents() {
          if (keep) return;
          for (auto& e : *evs) {
            if (e.first) (void)hipEventDestroy(e.first);
            if (e.second) (void)hipEventDestroy(e.second);
          }
        }
      } evguard{&cg.events};
      if (prof)
        for (size_t i = 0; i < nloops; ++i) {
          cg.events.emplace_back(nullptr, nullptr);
          HIPCHK(ctx, hipEventCreate(&cg.events.back().first));
          HIPCHK(ctx, hipEventCreate(&cg.events.back().second));
        }
      cg.iters = L.loop_iters;
      hipGraphNode_t prev = nullptr;
      for (const LaunchList::Item& it : L.items) {
        hipGraphNode_t node = nullptr;
        if (it.func) {
          std::vector<void*> av = L.argv(it);
          hipKernelNodeParams kp{};
          kp.func = const_cast<void*>(it.func);
          kp.gridDim = it.grid;
          kp.blockDim = it.block;
          kp.sharedMemBytes = it.shmem;
          kp.kernelParams = av.data();
          kp.extra = nullptr;
          HIPCHK(ctx, hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
        } else if (prof) {
          const auto& e = cg.events[it.mark / 2];
          HIPCHK(ctx, hipGraphAddEventRecordNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0,
                                                 (it.mark & 1) ? e.second : e.first));
        } else {
          continue;
        }
        prev = node;
      }
      HIPCHK(ctx, hipGraphInstantiate(&cg.exec, g, nullptr, nullptr, 0));
      evguard.keep = true;
      found = ctx->cg_graphs.emplace(key, std::move(cg)).first;
    }
    HIPCHK(ctx, hipGraphLaunch(found->second.exec, st));
    if (prof)
      for (size_t i = 0; i < nloops; ++i) {
        // a cached graph's events hold its latest replay; every replay of
        // one launch sequence does the same work
        ctx->cg_events.push_back(found->second.events[i]);
        ctx->cg_events_owned.push_back(false);
        ctx->cg_iters.push_back(found->second.iters[i]);
      }
    return;
  }
  std::vector<std::pair<hipEvent_t, hipEvent_t>> evs(nloops, {nullptr, nullptr});
  for (const LaunchList::Item& it : L.items) {
    if (it.func) {
      std::vector<void*> av = L.argv(it);
      HIPCHK(ctx, hipLaunchKernel(it.func, it.grid, it.block, av.data(), it.shmem, st));
    } else if (prof) {
      hipEvent_t& e = (it.mark & 1) ? evs[it.mark / 2].second : evs[it.mark / 2].first;
      HIPCHK(ctx, hipEventCreate(&e));
      HIPCHK(ctx, hipEventRecord(e, st));
    }
  }
  if (prof)
    for (size_t i = 0; i < nloops; ++i) {
      ctx->cg_events.push_back(evs[i]);
      ctx->cg_events_owned.push_back(true);
      ctx->cg_iters.push_back(L.loop_iters[i]);
    }
}

// A field's standard for the AUTO aniso grouping: its own (DevField::pal,
// set where one solve mixes standards) or the solve's.
int field_pal(const DevField& f, int is_pal) { return f.pal ? f.pal - 1 : is_pal; }

// Per-field CG state of a solve with these parameters (floors: hvd-core's).
FieldCG field_cg_init(const DevField& f, const hvd_solver_params& p, int n_inner, double tol) {
  FieldCG s{};
  s.npix = (double)f.h * (double)f.w;
  s.tol2 = tol * tol;
  s.hh_floor = 1e-12;
  s.gg_floor = 1e-10;
  s.lam_y = (float)(p.lambda_y > 0.0 ? p.lambda_y : 1.0);
  s.eps = (float)p.charbonnier_eps;
  s.eps_c = (float)p.chroma_eps;
  s.k = (float)p.structure_coupling;
  s.aniso_forced = (float)p.chroma_aniso;
  s.diag = (float)p.diag_prior;
  s.is_pal = field_pal(f, p.is_pal);
  s.n_inner = n_inner;
  return s;
}

// Static LDS of k_aniso_select (per-wave histograms, the scan, two scalars).
constexpr int kSelStaticLds = (16 * 256 + 256 + 2) * 4;

// ResolveChromaAniso's p98 selection for the fields of sample_prefix (host
// copy; one block per field).  A field's samples are staged in LDS when the
// largest field's set fits.
void record_aniso_select(const hvd_gpu_ctx* ctx, LaunchList& L, FieldCG* dstate, const int* dsp, const float* gx,
                         const float* gy, float lambda_c, const std::vector<int>& sample_prefix) {
  const int nf = (int)sample_prefix.size() - 1;
  int max_n = 0;
  for (int i = 0; i < nf; ++i) max_n = std::max(max_n, sample_prefix[i + 1] - sample_prefix[i]);
  const int lds_floats = (size_t)max_n * 4 + kSelStaticLds <= (size_t)ctx->max_dyn_lds ? max_n : 0;
  L.kernel((const void*)k_aniso_select, dim3(nf), dim3(kSelThreads), (unsigned)(sizeof(float) * lds_floats), dstate,
           dsp, gx, gy, lambda_c, lds_floats);
}

// Bytes run_refine stages for n fields (its descriptor region).
size_t refine_desc_bytes(int n) {
  return align256(sizeof(DevField) * n) + align256(sizeof(int) * (n + 1)) + align256(sizeof(FieldCG) * n) +
         align256(sizeof(int) * (n + 1)) + align256(sizeof(int) * (n + 1));
}

// Strip layout of the per-stage weights pass and the output pass: strips of
// at most 8 lines (no taller than the CG's), shorter still while a launch has
// fewer than 4096 waves.
struct AuxStrips {
  std::vector<int> tp;
  int nb = 0, rows = 0, spf = 0;
};

AuxStrips aux_strips(const std::vector<DevField>& F, int cg_rows) {
  AuxStrips A;
  const int nf = (int)F.size();
  auto strips = [&](int rows) {
    long n = 0;
    for (const DevField& f : F) n += solver_strips(f.h, f.w, rows);
    return n;
  };
  A.rows = std::min(cg_rows, 8);
  while (A.rows > kMinStripRows && strips(A.rows) < 4096) A.rows /= 2;
  A.rows = std::max(A.rows, kMinStripRows);
  A.tp.assign(nf + 1, 0);
  for (int i = 0; i < nf; ++i) A.tp[i + 1] = A.tp[i] + solver_strips(F[i].h, F[i].w, A.rows);
  A.nb = A.tp[nf];
  A.spf = A.tp[1];
  for (int i = 1; i < nf; ++i)
    if (F[i].h != F[0].h || F[i].w != F[0].w) A.spf = 0;
  return A;
}

// One solve group.  desc: a fixed workspace region of refine_desc_bytes(n)
// for the group's descriptors (the CG iterations then replay as a graph when
// the context allows it), or null (descriptors in an upload slot, eager).
void run_refine(hvd_gpu_ctx* ctx, std::vector<DevField>& F, const hvd_solver_params& p, double* partials,
                float* gx, float* gy, const std::vector<int>& sample_prefix, FieldCG* dstate, hipStream_t st,
                RefineOut out, int rows_tuned, char* desc) {
  const int nf = (int)F.size();
  for (int i = 0; i < nf; ++i) REQUIRE(ctx, F[i].s != nullptr, "field %d: no composite plane", i);
  const StripSetup SS = strip_setup(F, rows_tuned);
  const std::vector<int>& tp = SS.tp;
  const int nb = SS.nb, rows = SS.rows, spf = SS.spf;
  const int outer = std::max(1, p.irls_outer);
  const int cg_total = p.fast ? p.cg_iterations * 2 / 3 : p.cg_iterations;
  const int n_inner = std::max(1, cg_total / outer);
  const float tolf = p.cg_tol > 0.0 ? (float)p.cg_tol : (p.fast ? 0.10f : 0.02f);
  std::vector<FieldCG> init(nf);
  for (int i = 0; i < nf; ++i) init[i] = field_cg_init(F[i], p, n_inner, (double)tolf);
  // the once-per-stage weights pass and the output pass walk shorter strips
  // (more waves; they publish no partials, so their strip layout is free)
  const AuxStrips AX = aux_strips(F, rows);
  Stage stg;
  const size_t oF = stg.put(F), oT = stg.put(tp), oI = stg.put(init), oS = stg.put(sample_prefix),
               oA = stg.put(AX.tp);
  Slot* sl;
  REQUIRE(ctx, !desc || stg.buf.size() <= refine_desc_bytes(nf), "internal: descriptor region too small");
  char* base = desc ? upload_into(ctx, stg, desc, st, &sl) : upload(ctx, stg, st, &sl);
  const DevField* dF = (const DevField*)(base + oF);
  const int* dtp = (const int*)(base + oT);
  const int* dsp = (const int*)(base + oS);
  const StripGrid sg{dtp, nf, nb, rows, spf};
  const StripGrid sga{(const int*)(base + oA), nf, AX.nb, AX.rows, AX.spf};
  // the group's launches, replayed as one graph when the descriptors sit in
  // the fixed workspace region
  LaunchList L;
  L.kernel((const void*)k_refine_begin, dim3((nf + 63) / 64), dim3(64), 0, dstate, (const FieldCG*)(base + oI), nf);
  if (!(p.chroma_aniso > 0.0) && sample_prefix[nf] > 0) {
    const int blocks = std::min(4096, (sample_prefix[nf] + 255) / 256);
    L.kernel((const void*)k_aniso_samples, dim3(blocks), dim3(256), 0, dF, (const FieldCG*)dstate, dsp, nf, gx, gy);
  }
  record_aniso_select(ctx, L, dstate, dsp, gx, gy, (float)p.lambda_c, sample_prefix);
  // each stage: weights + stage-start gradient (one launch), then the CG
  for (int o = 0; o < outer; ++o) record_cg_loop(L, dF, dstate, sg, partials, nf, n_inner, false, o, o);
  L.kernel((const void*)k_refine_finalize, dim3(sga.nb), dim3(64), 0, dF, (const FieldCG*)dstate, sga);
  // no graphs in a no-allocation call: instantiating one allocates
  run_launches(ctx, L, desc != nullptr && ctx->graphs && !ctx->no_alloc, st);
  arm(ctx, sl, st);
  if (ctx->profiling) {
    void* h = nullptr;
    HIPCHK(ctx, hipHostMalloc(&h, sizeof(FieldCG) * nf, hipHostMallocDefault));
    ctx->cg_states.emplace_back(h, nf);
    HIPCHK(ctx, hipMemcpyAsync(h, dstate, sizeof(FieldCG) * nf, hipMemcpyDeviceToHost, st));
  }
  if (out.stage_iters || out.aniso) {
    std::vector<FieldCG> hs(nf);
    HIPCHK(ctx, hipMemcpyAsync(hs.data(), dstate, sizeof(FieldCG) * nf, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    for (int i = 0; i < nf; ++i) {
      if (out.stage_iters)
        for (int o = 0; o < outer; ++o) out.stage_iters[i * outer + o] = o < 16 ? hs[i].iters[o] : -1;
      if (out.aniso) out.aniso[i] = hs[i].aniso;
    }
  }
}

int aniso_samples(int h, int w, int is_pal) {
  const int group = is_pal ? 4 : 2;
  const int hh = h / group;
  const int nt = hh / 2, nx = (w - 1 + 2) / 3;
  return (w >= 2 && nt > 0) ? nt * nx : 0;
}

// Solve groups: consecutive fields whose solver working set (about 52 B per
// pixel: chi x2, d x2, weights, composite) mostly stays in the 256 MB
// Infinity Cache, so a group's CG iterations re-read their planes from the
// cache rather than HBM.  A batch is cut into the fewest groups of at most
// about `max_px` pixels, balanced in size; groups run one after another on
// the stream and share one workspace region.
constexpr size_t kSolveGroupPx = 4000000;  // default; 21-22 NTSC fields measured best

std::vector<std::pair<int, int>> solve_groups(const std::vector<DevField>& F, size_t max_px) {
  std::vector<std::pair<int, int>> g;
  const int n = (int)F.size();
  if (n == 0) return g;
  if (max_px == 0) max_px = kSolveGroupPx;
  double total = 0.0;
  for (const DevField& f : F) total += (double)f.h * f.w;
  const int ng = std::max(1, std::min(n, (int)std::ceil(total / (double)max_px)));
  const double target = total / ng;
  int b = 0;
  double acc = 0.0;
  for (int i = 0; i < n; ++i) {
    const double np = (double)F[i].h * F[i].w;
    // close the group before field i when it holds enough and i would
    // overshoot the target by more than half of itself
    if (i > b && (int)g.size() < ng - 1 && acc + 0.5 * np > target) {
      g.emplace_back(b, i);
      b = i;
      acc = 0.0;
    }
    acc += np;
  }
  g.emplace_back(b, n);
  return g;
}

// The launch shape for a batch: the fixed tuning, else the autotuned entry for
// the shape of its largest fields (the fields of that width within two lines
// of that height and of its standard, as the two fields of a frame are; in a
// mixed batch, the standard whose fields are larger): all of it for the same
// count of such fields, else the strip height of the entry with the nearest
// count (its group size belongs to that count; the default applies); else the
// built-in policy.
SolverTuning tuning_for(const hvd_gpu_ctx* ctx, const std::vector<DevField>& F, int is_pal) {
  SolverTuning t = ctx->fixed;
  if ((t.rows > 0 && t.group_px > 0) || F.empty() || ctx->tuned.empty()) return t;
  size_t i0 = 0;
  for (size_t i = 1; i < F.size(); ++i)
    if ((long)F[i].h * F[i].w > (long)F[i0].h * F[i0].w) i0 = i;
  const DevField& f0 = F[i0];
  const int pal0 = field_pal(f0, is_pal) ? 1 : 0;
  int n = 0;
  for (const DevField& f : F)
    if (f.w == f0.w && std::abs(f.h - f0.h) <= 2 && (field_pal(f, is_pal) ? 1 : 0) == pal0) ++n;
  auto it = ctx->tuned.end();
  for (int dh : {0, -1, 1, -2, 2}) {
    it = ctx->tuned.find({f0.h + dh, f0.w, pal0});
    if (it != ctx->tuned.end()) break;
  }
  if (it == ctx->tuned.end() || it->second.empty()) return t;
  const auto exact = it->second.find(n);
  const SolverTuning* best = exact != it->second.end() ? &exact->second : nullptr;
  double bd = 1e300;
  if (!best)
    for (const auto& kv : it->second) {
      const double d = std::fabs(std::log((double)kv.first / n));
      if (d < bd) {
        bd = d;
        best = &kv.second;
      }
    }
  if (t.rows <= 0) t.rows = best->rows;
  if (t.group_px == 0 && exact != it->second.end()) t.group_px = best->group_px;
  return t;
}

// One solve group's workspace, carved after `cv`.  Partials are sized for the
// shortest strips so any strip height the launch picks fits.
struct GroupLayout {
  std::vector<size_t> chi1, d, w;
  std::vector<int> sp;  // aniso sample prefix
  size_t part = 0, gx = 0, gy = 0, st = 0, desc = 0, end = 0;
};

GroupLayout group_layout(Carve cv, const DevField* F, int n, int is_pal) {
  GroupLayout L;
  L.chi1.resize(n);
  L.d.resize(n);
  L.w.resize(n);
  L.sp.assign(n + 1, 0);
  size_t strips = 0;
  for (int i = 0; i < n; ++i) {
    const size_t np = (size_t)F[i].h * F[i].w;
    L.chi1[i] = cv.take(sizeof(float2) * np);
    L.d[i] = cv.take(sizeof(float2) * np * 2);
    L.w[i] = cv.take(sizeof(float4) * np);
    strips += (size_t)solver_strips(F[i].h, F[i].w, kMinStripRows);
    L.sp[i + 1] = L.sp[i] + aniso_samples(F[i].h, F[i].w, field_pal(F[i], is_pal));
  }
  L.part = cv.take(sizeof(double) * 2 * kCgPartials * strips);  // two launch slots
  L.gx = cv.take(sizeof(float) * (L.sp[n] + 1));
  L.gy = cv.take(sizeof(float) * (L.sp[n] + 1));
  L.st = cv.take(sizeof(FieldCG) * n);
  L.desc = cv.take(refine_desc_bytes(n));
  L.end = cv.off;
  return L;
}

// Workspace of the per-pixel solve (pixsolve.hip) for fields F, field i with
// nn[i] neighbours, carved after `cv`.
struct PxLayout {
  std::vector<size_t> chi1, d, g, w, wd, part, y1, dyj, gyj;
  std::vector<std::vector<size_t>> nb;
  std::vector<int> cp;  // chunk prefix
  std::vector<int> sp;  // aniso sample prefix
  size_t px = 0, st = 0, gx = 0, gy = 0, desc = 0, end = 0;
};

size_t px_desc_bytes(int n) {
  return align256(sizeof(PxField) * n) + align256(sizeof(DevField) * n) + align256(sizeof(int) * (n + 1)) +
         align256(sizeof(FieldCG) * n) + align256(sizeof(int) * (n + 1));
}

PxLayout px_layout(Carve cv, const std::vector<DevField>& F, const std::vector<int>& nn, bool diag, int is_pal,
                   bool joint = false) {
  const int n = (int)F.size();
  PxLayout L;
  for (auto* v : {&L.chi1, &L.d, &L.g, &L.w, &L.wd, &L.part, &L.y1, &L.dyj, &L.gyj}) v->resize(n);
  L.nb.resize(n);
  L.cp.assign(n + 1, 0);
  L.sp.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const size_t np = (size_t)F[i].h * F[i].w;
    const int nc = (int)((np + kPxChunk - 1) / kPxChunk);
    L.cp[i + 1] = L.cp[i] + nc;
    L.sp[i + 1] = L.sp[i] + aniso_samples(F[i].h, F[i].w, field_pal(F[i], is_pal));
    L.chi1[i] = cv.take(sizeof(float2) * np);
    L.d[i] = cv.take(sizeof(float2) * np * 2);
    L.g[i] = cv.take(sizeof(float2) * np);
    L.w[i] = cv.take(sizeof(float4) * np);
    L.wd[i] = diag ? cv.take(sizeof(float2) * np) : 0;
    for (int j = 0; j < (nn.empty() ? 0 : nn[i]); ++j) L.nb[i].push_back(cv.take(sizeof(float4) * np));
    L.part[i] = cv.take(sizeof(double) * 4 * nc);
    if (joint) {
      L.y1[i] = cv.take(sizeof(float) * np);
      L.dyj[i] = cv.take(sizeof(float) * np * 2);
      L.gyj[i] = cv.take(sizeof(float) * np);
    }
  }
  L.px = cv.take(sizeof(PxState) * n);
  L.st = cv.take(sizeof(FieldCG) * n);
  L.gx = cv.take(sizeof(float) * (L.sp[n] + 1));
  L.gy = cv.take(sizeof(float) * (L.sp[n] + 1));
  L.desc = cv.take(px_desc_bytes(n));
  L.end = cv.off;
  return L;
}

// End offset of the solver workspace refine_planes carves after `cv`.
size_t refine_ws_end(const hvd_gpu_ctx* ctx, Carve cv, const std::vector<DevField>& F, int is_pal,
                     bool diag = false) {
  // the oriented prior runs the per-pixel solve
  if (diag) return std::max(cv.off, px_layout(cv, F, {}, true, is_pal).end);
  size_t end = cv.off;
  for (const auto& gr : solve_groups(F, tuning_for(ctx, F, is_pal).group_px))
    end = std::max(end, group_layout(cv, F.data() + gr.first, gr.second - gr.first, is_pal).end);
  return end;
}

// Refine workspace for planes already described by DevField (chi[0], s, y and
// the carrier set); fills the remaining pointers from the carve.
void check_params(hvd_gpu_ctx* ctx, const hvd_solver_params* p) {
  REQUIRE(ctx, p != nullptr, "null solver parameters");
  REQUIRE(ctx, p->diag_prior >= 0.0 && std::isfinite(p->diag_prior), "diag_prior must be finite and >= 0");
  REQUIRE(ctx, p->charbonnier_eps > 0.0 && p->chroma_eps > 0.0, "Charbonnier scales must be positive");
  REQUIRE(ctx, p->cg_iterations >= 0, "negative cg_iterations");
}

void check_plane(hvd_gpu_ctx* ctx, const hvd_gpu_plane& pl, int i) {
  REQUIRE(ctx, i < 65535, "more than 65535 planes in one call");
  REQUIRE(ctx, pl.lines >= 1 && pl.width >= 1, "plane %d: empty", i);
  REQUIRE(ctx, pl.s && pl.chi, "plane %d: null s or chi", i);
  REQUIRE(ctx, pl.carrier || pl.theta, "plane %d: needs a carrier or theta", i);
  REQUIRE(ctx, pl.width <= 2048 && pl.lines + 32 <= 1024, "plane %d: %d x %d exceeds the FFT limits (1024 - 32 lines, 2048 samples)",
          i, pl.lines, pl.width);
}

PrepDesc plane_tables(const hvd_gpu_plane& pl, float2* rowph, int8_t* vsw, float2* colph) {
  PrepDesc d{};
  d.theta_field = pl.theta;
  d.vswitch_field = pl.vswitch;
  d.rowph = rowph;
  d.vsw_active = pl.vswitch ? vsw : nullptr;
  d.colph = colph;
  d.fal = 0;
  d.a0 = pl.active_video_start;
  d.lines = pl.lines;
  d.width = pl.width;
  d.std_phase = pl.phase_per_sample == 0.5 * kPi;
  d.w0 = pl.phase_per_sample;
  return d;
}

}  // namespace

namespace {
void clear_profile(hvd_gpu_ctx* ctx) {
  for (size_t i = 0; i < ctx->cg_events.size(); ++i)
    if (ctx->cg_events_owned[i]) {
      (void)hipEventDestroy(ctx->cg_events[i].first);
      (void)hipEventDestroy(ctx->cg_events[i].second);
    }
  for (auto& p : ctx->cg_states) (void)hipHostFree(p.first);
  ctx->cg_events.clear();
  ctx->cg_events_owned.clear();
  for (hipEvent_t e : ctx->retired) (void)hipEventDestroy(e);
  ctx->retired.clear();
  ctx->cg_iters.clear();
  ctx->cg_states.clear();
}
}  // namespace

// ============================================================================
extern "C" {

int hvd_gpu_abi_version(void) { return 103; }

int hvd_gpu_create(int device, hvd_gpu_ctx** out) {
  if (!out) return HVD_GPU_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return HVD_GPU_ERR_INVALID;
  hvd_gpu_ctx* ctx = new hvd_gpu_ctx;
  ctx->device = device;
  const int rc = guarded(ctx, [&] {
    for (Slot& s : ctx->slots) HIPCHK(ctx, hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    int lds = 0;
    HIPCHK(ctx, hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    ctx->max_dyn_lds = lds;
    for (const void* k : {(const void*)k_fft2_rows, (const void*)k_fft2_cols, (const void*)k_me_conf_tiles})
      HIPCHK(ctx, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
    REQUIRE(ctx, (size_t)lds >= kConfBandMaxLds, "the device has %d bytes of LDS per block (needs %zu)", lds,
            kConfBandMaxLds);
    HIPCHK(ctx, fft_allow_lds(lds));
    HIPCHK(ctx, hipFuncSetAttribute((const void*)k_aniso_select, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    std::max(0, lds - kSelStaticLds)));
    HIPCHK(ctx, scratch_address(device, &ctx->scratch));
  });
  if (rc != HVD_GPU_OK) {
    hvd_gpu_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return HVD_GPU_OK;
}

void hvd_gpu_destroy(hvd_gpu_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipDeviceSynchronize();
  for (Slot& s : ctx->slots) {
    if (s.ev) (void)hipEventDestroy(s.ev);
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
  }
  drop_graphs(ctx);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->win) (void)hipFree(ctx->win);
  if (ctx->fwin) (void)hipFree(ctx->fwin);
  if (ctx->sel) (void)hipFree(ctx->sel);
  clear_profile(ctx);
  delete ctx;
}

const char* hvd_gpu_last_error(const hvd_gpu_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hvd_gpu_synchronize(hvd_gpu_ctx* ctx, void* stream) {
  return guarded(ctx, [&] { HIPCHK(ctx, hipStreamSynchronize(S(stream))); });
}

int hvd_gpu_set_profiling(hvd_gpu_ctx* ctx, int enable) {
  return guarded(ctx, [&] {
    HIPCHK(ctx, hipDeviceSynchronize());
    clear_profile(ctx);
    ctx->profiling = enable != 0;
  });
}

int hvd_gpu_cg_timing(hvd_gpu_ctx* ctx, double* ms, long long* iterations, double* pixel_iterations) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, ms && iterations && pixel_iterations, "null output");
    HIPCHK(ctx, hipDeviceSynchronize());
    double tot = 0.0, pxit = 0.0;
    long long nl = 0;
    for (size_t i = 0; i < ctx->cg_events.size(); ++i) {
      float t = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&t, ctx->cg_events[i].first, ctx->cg_events[i].second));
      tot += t;
      nl += ctx->cg_iters[i];
    }
    for (auto& p : ctx->cg_states) {
      const FieldCG* st = static_cast<const FieldCG*>(p.first);
      for (int f = 0; f < p.second; ++f) {
        long long it = 0;
        for (int o = 0; o < 16; ++o) it += st[f].iters[o];
        pxit += st[f].npix * (double)it;
      }
    }
    *ms = tot;
    *iterations = nl;
    *pixel_iterations = pxit;
    clear_profile(ctx);
  });
}

int hvd_gpu_sample_to_ire_u16(hvd_gpu_ctx* ctx, const uint16_t* raw, size_t n, float blanking, float white,
                              float* out, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, raw && out, "null buffer");
    REQUIRE(ctx, white != blanking, "white level equals blanking");
    if (n == 0) return;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_u16_to_ire, dim3(blocks), dim3(256), 0, S(stream), raw, n, blanking,
                       (white - blanking) / 100.0f, out);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_burst_lockin_phase(hvd_gpu_ctx* ctx, const float* const* fields, int n, const hvd_field_geometry* geom,
                               float* const* theta, int8_t* const* vswitch, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (fields && geom && theta)), "null argument");
    if (n == 0) return;
    std::vector<LockinGeom> G(n);
    std::vector<const float*> F(n);
    std::vector<float*> T(n);
    std::vector<int8_t*> V(n);
    for (int i = 0; i < n; ++i) {
      check_geom(ctx, geom[i], i);
      REQUIRE(ctx, fields[i] && theta[i], "field %d: null pointer", i);
      G[i] = lockin_geom(geom[i]);
      F[i] = fields[i];
      T[i] = theta[i];
      V[i] = vswitch ? vswitch[i] : nullptr;
      REQUIRE(ctx, !G[i].vswitch || V[i], "field %d: PAL-family field needs a vswitch buffer", i);
    }
    run_lockin(ctx, G, F, T, V, S(stream));
  });
}

int hvd_gpu_burst_amplitude_ire(hvd_gpu_ctx* ctx, const float* const* fields, int n, const hvd_field_geometry* geom,
                                float* amp, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (fields && geom && amp)), "null argument");
    if (n == 0) return;
    std::vector<LockinGeom> G(n);
    std::vector<const float*> F(n);
    for (int i = 0; i < n; ++i) {
      check_geom(ctx, geom[i], i);
      REQUIRE(ctx, fields[i], "field %d: null pointer", i);
      G[i] = lockin_geom(geom[i]);
      F[i] = fields[i];
    }
    Stage stg;
    const size_t oG = stg.put(G), oF = stg.put(F);
    Slot* sl;
    char* base = upload(ctx, stg, S(stream), &sl);
    hipLaunchKernelGGL(k_burst_amp, dim3(n), dim3(256), 0, S(stream), (const LockinGeom*)(base + oG),
                       (const float* const*)(base + oF), amp);
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, S(stream));
  });
}

int hvd_gpu_holographic_init(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, int symmetry_variant,
                             void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || planes), "null argument");
    if (n == 0) return;
    hipStream_t st = S(stream);
    Carve cv;
    std::vector<size_t> o_row(n), o_col(n), o_vsw(n), o_spec(n), o_var(n);
    const int nvar = symmetry_variant ? 3 : 2;
    for (int i = 0; i < n; ++i) {
      const hvd_gpu_plane& pl = planes[i];
      check_plane(ctx, pl, i);
      const int pad = pl.lines > 1 ? std::min(16, pl.lines - 1) : 0;
      const size_t np = (size_t)pl.lines * pl.width;
      o_row[i] = cv.take(sizeof(float2) * pl.lines);
      o_vsw[i] = cv.take(pl.lines);
      o_col[i] = cv.take(sizeof(float2) * pl.width);
      o_spec[i] = cv.take(sizeof(float2) * (size_t)(pl.lines + 2 * pad) * pl.width);
      o_var[i] = cv.take(sizeof(float2) * np * nvar);
    }
    char* ws = ensure_ws(ctx, cv.off, st);
    PlanSet ps;
    std::vector<PrepDesc> prep;
    std::vector<InitField> F(n);
    for (int i = 0; i < n; ++i) {
      const hvd_gpu_plane& pl = planes[i];
      InitField f{};
      f.h = pl.lines;
      f.w = pl.width;
      f.pad = pl.lines > 1 ? std::min(16, pl.lines - 1) : 0;
      f.hp = f.h + 2 * f.pad;
      f.s = pl.s;
      f.carrier = reinterpret_cast<const float2*>(pl.carrier);
      if (!pl.carrier) {
        prep.push_back(plane_tables(pl, (float2*)(ws + o_row[i]), (int8_t*)(ws + o_vsw[i]), (float2*)(ws + o_col[i])));
        f.rowph = (const float2*)(ws + o_row[i]);
        f.colph = (const float2*)(ws + o_col[i]);
        f.vsw = pl.vswitch ? (const int8_t*)(ws + o_vsw[i]) : nullptr;
      }
      f.spec = (float2*)(ws + o_spec[i]);
      for (int v = 0; v < nvar; ++v) f.var[v] = (float2*)(ws + o_var[i]) + (size_t)v * f.h * f.w;
      f.chi0 = reinterpret_cast<float2*>(pl.chi);
      f.y0 = pl.y;
      f.row_plan = ps.get(ctx, f.w, kRowGroup);
      f.col_plan = ps.get(ctx, f.hp, kColGroup);
      f.nvar = nvar;
      f.sample_rate = (float)pl.sample_rate;
      F[i] = f;
    }
    run_prepare(ctx, prep, false, st);
    run_init(ctx, F, ps, st);
  });
}

}  // extern "C"

namespace {

// Shared by the plane-level refine and the field driver: carve the solver
// workspace after `cv` and run.
// One field's extra planes in the joint solve (VariationalRefineJoint).
struct PxJointField {
  float* y;  // Y0 in, Y out
  const float* a_luma;
  const float2* a_chroma;
  const float* a_weight;  // null: no anchor
};

void refine_pixel(hvd_gpu_ctx* ctx, const std::vector<DevField>& F,
                  const std::vector<std::vector<hvd_gpu_neighbor>>& N, const hvd_solver_params& p, float nu,
                  float eps_t, hipStream_t st, Carve cv, RefineOut out,
                  const std::vector<PxJointField>* joint = nullptr, float nu_a = 0.f);

void refine_planes(hvd_gpu_ctx* ctx, std::vector<DevField>& F, const hvd_solver_params& p, hipStream_t st,
                   Carve cv, RefineOut out) {
  // the caller sized the workspace (its tables point into it already)
  REQUIRE(ctx, refine_ws_end(ctx, cv, F, p.is_pal, p.diag_prior > 0.0) <= ctx->ws_cap,
          "internal: solver workspace not sized");
  if (p.diag_prior > 0.0) {
    refine_pixel(ctx, F, {}, p, 0.f, 0.f, st, cv, out);
    return;
  }
  char* ws = ctx->ws;
  const int outer = std::max(1, p.irls_outer);
  const SolverTuning tune = tuning_for(ctx, F, p.is_pal);
  for (const auto& gr : solve_groups(F, tune.group_px)) {
    const int b = gr.first, n = gr.second - gr.first;
    const GroupLayout L = group_layout(cv, F.data() + b, n, p.is_pal);
    std::vector<DevField> G(F.begin() + b, F.begin() + gr.second);
    for (int i = 0; i < n; ++i) {
      G[i].chi[1] = (float2*)(ws + L.chi1[i]);
      G[i].d[0] = (float2*)(ws + L.d[i]);
      G[i].d[1] = G[i].d[0] + (size_t)G[i].h * G[i].w;
      G[i].wts = (float4*)(ws + L.w[i]);
    }
    RefineOut o;
    o.stage_iters = out.stage_iters ? out.stage_iters + (size_t)b * outer : nullptr;
    o.aniso = out.aniso ? out.aniso + b : nullptr;
    run_refine(ctx, G, p, (double*)(ws + L.part), (float*)(ws + L.gx), (float*)(ws + L.gy), L.sp,
               (FieldCG*)(ws + L.st), st, o, tune.rows, ws + L.desc);
  }
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// The solve's adjoint (adjoint.hip): stages K..1, each an adjoint CG with the
// stage's weights (recomputed from its starting chi), then the local and
// transposed per-pixel passes.  All fields advance together.
void run_adjoint(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, const hvd_solver_params& p,
                 const float* const* stages, const float* const* gchi, const float* const* gy, int adj_iters,
                 double adj_tol, float* const* gs, float* const* gchi0, double* glam_host, hipStream_t st) {
  const int K = std::max(1, p.irls_outer);
  Carve cv;
  std::vector<size_t> o_row(n), o_col(n), o_vsw(n), o_chi1(n), o_d(n), o_w(n), o_a(n), o_sc(n);
  std::vector<int> sp(n + 1, 0);
  size_t strips = 0, maxpx = 0;
  for (int i = 0; i < n; ++i) {
    check_plane(ctx, planes[i], i);
    REQUIRE(ctx, stages[i] && gchi[i] && gs[i] && gchi0[i], "plane %d: null buffer", i);
    const size_t np = (size_t)planes[i].lines * planes[i].width;
    maxpx = std::max(maxpx, np);
    o_row[i] = cv.take(sizeof(float2) * planes[i].lines);
    o_vsw[i] = cv.take(planes[i].lines);
    o_col[i] = cv.take(sizeof(float2) * planes[i].width);
    o_a[i] = cv.take(sizeof(float2) * np);
    o_chi1[i] = cv.take(sizeof(float2) * np);
    o_d[i] = cv.take(sizeof(float2) * np * 2);
    o_w[i] = cv.take(sizeof(float4) * np);
    o_sc[i] = cv.take(sizeof(float) * 6 * np);
    strips += (size_t)solver_strips(planes[i].lines, planes[i].width, kMinStripRows);
    sp[i + 1] = sp[i] + aniso_samples(planes[i].lines, planes[i].width, p.is_pal);
  }
  const int nbx = (int)((maxpx + 255) / 256);
  const size_t o_part = cv.take(sizeof(double) * 2 * kCgPartials * strips);  // two launch slots
  const size_t o_apart = cv.take(sizeof(double) * 2 * (size_t)nbx * n);
  const size_t o_glam = cv.take(sizeof(double) * 2 * n);
  const size_t o_gx = cv.take(sizeof(float) * (sp[n] + 1));
  const size_t o_gy = cv.take(sizeof(float) * (sp[n] + 1));
  const size_t o_st = cv.take(sizeof(FieldCG) * n);
  char* ws = ensure_ws(ctx, cv.off, st);
  // carrier tables (factored planes) as for the forward
  std::vector<PrepDesc> prep;
  std::vector<DevField> Fw(n), Fa(n);
  std::vector<AdjField> A(n);
  for (int i = 0; i < n; ++i) {
    const hvd_gpu_plane& pl = planes[i];
    const size_t np = (size_t)pl.lines * pl.width;
    DevField f{};
    f.h = pl.lines;
    f.w = pl.width;
    f.s = pl.s;
    f.carrier = reinterpret_cast<const float2*>(pl.carrier);
    if (!pl.carrier) {
      prep.push_back(plane_tables(pl, (float2*)(ws + o_row[i]), (int8_t*)(ws + o_vsw[i]), (float2*)(ws + o_col[i])));
      f.rowph = (const float2*)(ws + o_row[i]);
      f.colph = (const float2*)(ws + o_col[i]);
      f.vsw = pl.vswitch ? (const int8_t*)(ws + o_vsw[i]) : nullptr;
    }
    f.wts = (float4*)(ws + o_w[i]);
    f.d[0] = (float2*)(ws + o_d[i]);
    f.d[1] = f.d[0] + np;
    Fw[i] = f;  // weights pass: chi[0] set per stage
    DevField a = f;
    a.s = nullptr;  // S = 0: the gradient is the Hessian applied to chi
    a.chi[0] = (float2*)(ws + o_a[i]);
    a.chi[1] = (float2*)(ws + o_chi1[i]);
    a.rhs = reinterpret_cast<const float2*>(gchi0[i]);  // gbar lives in the gchi0 output
    a.y = nullptr;
    Fa[i] = a;
    AdjField& q = A[i];
    q = AdjField{};
    q.h = f.h;
    q.w = f.w;
    q.s = pl.s;
    q.rowph = f.rowph;
    q.colph = f.colph;
    q.vsw = f.vsw;
    q.carrier = f.carrier;
    q.abar = a.chi[0];
    q.wts = f.wts;
    q.gbar = reinterpret_cast<float2*>(gchi0[i]);
    q.sbar = gs[i];
    q.scratch = (float*)(ws + o_sc[i]);
    q.gchi = reinterpret_cast<const float2*>(gchi[i]);
    q.gy = gy ? gy[i] : nullptr;
  }
  run_prepare(ctx, prep, false, st);
  const StripSetup SS = strip_setup(Fw, tuning_for(ctx, Fw, p.is_pal).rows);
  for (int i = 0; i < n; ++i) {
    Fa[i].tiles_x = Fw[i].tiles_x;
    Fa[i].tile0 = Fw[i].tile0;
  }
  std::vector<FieldCG> init(n);
  for (int i = 0; i < n; ++i) {
    init[i] = field_cg_init(Fw[i], p, adj_iters, adj_tol);
    init[i].hh_floor = 0.0;  // the adjoint's right-hand side may be tiny: relative exit only
    init[i].gg_floor = 0.0;
  }
  FieldCG* dstate = (FieldCG*)(ws + o_st);
  double* partials = (double*)(ws + o_part);
  double* apart = (double*)(ws + o_apart);
  double* glam = (double*)(ws + o_glam);
  // per stage, the weights table reads that stage's starting chi
  std::vector<std::vector<DevField>> FwK(K, Fw);
  std::vector<std::vector<AdjField>> AK(K, A);
  for (int k = 0; k < K; ++k)
    for (int i = 0; i < n; ++i) {
      const size_t np = (size_t)planes[i].lines * planes[i].width;
      const float2* chip = reinterpret_cast<const float2*>(stages[i]) + (size_t)k * np;
      // both ping-pong slots: the weights kernel reads chi[cur], and the
      // adjoint CG leaves cur at either value
      FwK[k][i].chi[0] = const_cast<float2*>(chip);
      FwK[k][i].chi[1] = const_cast<float2*>(chip);
      AK[k][i].chip = chip;
      AK[k][i].chik = k + 1 < K ? reinterpret_cast<const float2*>(stages[i]) + (size_t)(k + 1) * np
                                : reinterpret_cast<const float2*>(planes[i].chi);
    }
  Stage stg;
  const size_t oT = stg.put(SS.tp), oI = stg.put(init), oS = stg.put(sp), oFa = stg.put(Fa);
  std::vector<size_t> oFw(K), oA(K);
  for (int k = 0; k < K; ++k) {
    oFw[k] = stg.put(FwK[k]);
    oA[k] = stg.put(AK[k]);
  }
  Slot* sl;
  char* base = upload(ctx, stg, st, &sl);
  const int* dtp = (const int*)(base + oT);
  const StripGrid sg{dtp, n, SS.nb, SS.rows, SS.spf};
  const DevField* dFa = (const DevField*)(base + oFa);
  hipLaunchKernelGGL(k_refine_begin, dim3((n + 63) / 64), dim3(64), 0, st, dstate, (const FieldCG*)(base + oI), n);
  LAUNCH_CHECK(ctx);
  // the forward's aniso, from chi0 (stage 1's start)
  if (!(p.chroma_aniso > 0.0) && sp[n] > 0) {
    const int blocks = std::min(4096, (sp[n] + 255) / 256);
    hipLaunchKernelGGL(k_aniso_samples, dim3(blocks), dim3(256), 0, st, (const DevField*)(base + oFw[0]),
                       (const FieldCG*)dstate, (const int*)(base + oS), n, (float*)(ws + o_gx), (float*)(ws + o_gy));
    LAUNCH_CHECK(ctx);
  }
  {
    LaunchList L;
    record_aniso_select(ctx, L, dstate, (const int*)(base + oS), (float*)(ws + o_gx), (float*)(ws + o_gy),
                        (float)p.lambda_c, sp);
    run_launches(ctx, L, false, st);
  }
  hipLaunchKernelGGL(k_zero_d, dim3((2 * n + 255) / 256), dim3(256), 0, st, glam, 2 * n);
  LAUNCH_CHECK(ctx);
  const dim3 pgrid(nbx, n);
  hipLaunchKernelGGL(k_adj_seed, pgrid, dim3(256), 0, st, (const AdjField*)(base + oA[K - 1]));
  LAUNCH_CHECK(ctx);
  for (int k = K - 1; k >= 0; --k) {
    hipLaunchKernelGGL(k_irls_weights, dim3(SS.nb), dim3(64), 0, st, (const DevField*)(base + oFw[k]), dstate, sg, 0);
    LAUNCH_CHECK(ctx);
    for (int i = 0; i < n; ++i) {
      const size_t np = (size_t)planes[i].lines * planes[i].width;
      hipLaunchKernelGGL(k_zero2, dim3((unsigned)std::min<size_t>((np + 255) / 256, 4096)), dim3(256), 0, st,
                         Fa[i].chi[0], np);
      LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_zero2, dim3((unsigned)std::min<size_t>((np + 255) / 256, 4096)), dim3(256), 0, st,
                         Fa[i].chi[1], np);
      LAUNCH_CHECK(ctx);
    }
    {
      LaunchList L;
      record_cg_loop(L, dFa, dstate, sg, partials, n, adj_iters, true, -1, 0);
      run_launches(ctx, L, false, st);
    }
    hipLaunchKernelGGL(k_refine_finalize, dim3(SS.nb), dim3(64), 0, st, dFa, (const FieldCG*)dstate, sg);
    LAUNCH_CHECK(ctx);
    const AdjField* dA = (const AdjField*)(base + oA[k]);
    hipLaunchKernelGGL(k_adj_local, pgrid, dim3(256), 0, st, dA, (const FieldCG*)dstate, nbx, apart);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_adj_reduce, dim3(n), dim3(256), 0, st, (const FieldCG*)dstate, nbx, (const double*)apart,
                       glam);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_adj_transpose, pgrid, dim3(256), 0, st, dA);
    LAUNCH_CHECK(ctx);
  }
  arm(ctx, sl, st);
  HIPCHK(ctx, hipMemcpyAsync(glam_host, glam, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
}

// The refine entry points: optional per-stage chi saving for the adjoint.
void refine_entry(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, const hvd_solver_params* p,
                  int* stage_iters, float* aniso, float* const* stages, void* stream) {
  {
    REQUIRE(ctx, n >= 0 && (n == 0 || planes), "null argument");
    check_params(ctx, p);
    if (n == 0) return;
    hipStream_t st = S(stream);
    Carve cv;
    std::vector<size_t> o_row(n), o_col(n), o_vsw(n);
    for (int i = 0; i < n; ++i) {
      check_plane(ctx, planes[i], i);
      o_row[i] = cv.take(sizeof(float2) * planes[i].lines);
      o_vsw[i] = cv.take(planes[i].lines);
      o_col[i] = cv.take(sizeof(float2) * planes[i].width);
    }
    std::vector<DevField> F(n);
    for (int i = 0; i < n; ++i) {
      DevField f{};
      f.h = planes[i].lines;
      f.w = planes[i].width;
      f.s = planes[i].s;
      f.carrier = reinterpret_cast<const float2*>(planes[i].carrier);
      f.chi[0] = reinterpret_cast<float2*>(planes[i].chi);
      f.y = planes[i].y;
      if (stages) {
        REQUIRE(ctx, stages[i], "plane %d: null stage buffer", i);
        f.save = reinterpret_cast<float2*>(stages[i]);
      }
      F[i] = f;
    }
    // size the whole call (tables + solver) before any pointer is taken
    ensure_ws(ctx, refine_ws_end(ctx, cv, F, p->is_pal, p->diag_prior > 0.0), st);
    char* ws = ctx->ws;
    std::vector<PrepDesc> prep;
    for (int i = 0; i < n; ++i) {
      if (!planes[i].carrier) {
        prep.push_back(plane_tables(planes[i], (float2*)(ws + o_row[i]), (int8_t*)(ws + o_vsw[i]),
                                    (float2*)(ws + o_col[i])));
        F[i].rowph = (const float2*)(ws + o_row[i]);
        F[i].colph = (const float2*)(ws + o_col[i]);
        F[i].vsw = planes[i].vswitch ? (const int8_t*)(ws + o_vsw[i]) : nullptr;
      }
    }
    run_prepare(ctx, prep, false, st);
    refine_planes(ctx, F, *p, st, cv, RefineOut{stage_iters, aniso});
  }
}

// The per-pixel IRLS + CG solve of fields F (chi[0], s, y and the carrier
// set; variational.cpp:198-556) with the oriented prior when diag_prior > 0
// and field i's neighbours N[i] (N empty: none) at strength nu, robust scale
// eps_t.  All fields advance together, one wave per chunk of kPxChunk pixels.
void refine_pixel(hvd_gpu_ctx* ctx, const std::vector<DevField>& F,
                  const std::vector<std::vector<hvd_gpu_neighbor>>& N, const hvd_solver_params& p, float nu,
                  float eps_t, hipStream_t st, Carve cv, RefineOut out, const std::vector<PxJointField>* joint,
                  float nu_a) {
  const int nf = (int)F.size();
  if (nf == 0) return;
  std::vector<int> nn(nf, 0);
  if (!N.empty())
    for (int i = 0; i < nf; ++i) nn[i] = (int)N[i].size();
  const bool diag = p.diag_prior > 0.0;
  const bool jt = joint != nullptr;
  const PxLayout L = px_layout(cv, F, nn, diag, p.is_pal, jt);
  REQUIRE(ctx, L.end <= ctx->ws_cap, "internal: pixel-solve workspace not sized");
  char* ws = ctx->ws;
  const int outer = std::max(1, p.irls_outer);
  const int cg_total = p.fast ? p.cg_iterations * 2 / 3 : p.cg_iterations;
  const int n_inner = std::max(1, cg_total / outer);
  const float tolf = p.cg_tol > 0.0 ? (float)p.cg_tol : (p.fast ? 0.10f : 0.02f);
  FieldCG* dstate = (FieldCG*)(ws + L.st);
  PxState* dpx = (PxState*)(ws + L.px);
  std::vector<PxField> P(nf);
  std::vector<DevField> D(F);
  std::vector<FieldCG> init(nf);
  for (int i = 0; i < nf; ++i) {
    const size_t np = (size_t)F[i].h * F[i].w;
    DevField& f = D[i];
    f.chi[1] = (float2*)(ws + L.chi1[i]);
    f.d[0] = (float2*)(ws + L.d[i]);
    f.d[1] = f.d[0] + np;
    f.wts = (float4*)(ws + L.w[i]);
    PxField& q = P[i];
    q.f = f;
    q.g = (float2*)(ws + L.g[i]);
    q.wts = f.wts;
    q.wd = diag ? (float2*)(ws + L.wd[i]) : nullptr;
    q.nn = nn[i];
    for (int j = 0; j < nn[i]; ++j) {
      q.nb[j] = (float4*)(ws + L.nb[i][j]);
      q.conf[j] = N[i][j].confidence;
      q.sw[j] = N[i][j].composite;
      q.cw[j] = reinterpret_cast<const float2*>(N[i][j].carrier);
    }
    q.nu = nn[i] > 0 ? nu : 0.f;
    q.et2 = eps_t * eps_t;
    if (jt) {
      const PxJointField& J = (*joint)[i];
      q.joint = 1;
      q.yv[0] = J.y;
      q.yv[1] = (float*)(ws + L.y1[i]);
      q.dyv[0] = (float*)(ws + L.dyj[i]);
      q.dyv[1] = q.dyv[0] + np;
      q.gy = (float*)(ws + L.gyj[i]);
      q.a_luma = J.a_luma;
      q.a_chroma = J.a_chroma;
      q.a_weight = J.a_weight;
      q.nu_a = J.a_weight ? nu_a : 0.f;
    }
    q.nchunks = L.cp[i + 1] - L.cp[i];
    q.part = (double*)(ws + L.part[i]);
    q.px = dpx + i;
    q.st = dstate + i;
    init[i] = field_cg_init(f, p, n_inner, (double)tolf);
  }
  Stage stg;
  const size_t oP = stg.put(P), oF = stg.put(D), oC = stg.put(L.cp), oI = stg.put(init), oS = stg.put(L.sp);
  REQUIRE(ctx, stg.buf.size() <= px_desc_bytes(nf), "internal: descriptor region too small");
  Slot* sl;
  char* base = upload_into(ctx, stg, ws + L.desc, st, &sl);
  const PxField* dP = (const PxField*)(base + oP);
  const int* dcp = (const int*)(base + oC);
  const int* dsp = (const int*)(base + oS);
  float* gx = (float*)(ws + L.gx);
  float* gy = (float*)(ws + L.gy);
  const dim3 grid(L.cp[nf]), blk(64 * kPxWaves);
  LaunchList Lq;
  Lq.kernel((const void*)k_refine_begin, dim3((nf + 63) / 64), dim3(64), 0, dstate, (const FieldCG*)(base + oI), nf);
  if (!(p.chroma_aniso > 0.0) && L.sp[nf] > 0) {
    const int blocks = std::min(4096, (L.sp[nf] + 255) / 256);
    Lq.kernel((const void*)k_aniso_samples, dim3(blocks), dim3(256), 0, (const DevField*)(base + oF),
              (const FieldCG*)dstate, dsp, nf, gx, gy);
  }
  record_aniso_select(ctx, Lq, dstate, dsp, gx, gy, (float)p.lambda_c, L.sp);
  bool any_nb = false;
  for (int i = 0; i < nf; ++i) any_nb = any_nb || nn[i] > 0;
  if (any_nb) Lq.kernel((const void*)k_px_prep, grid, blk, 0, dP, dcp, nf);
  int cur = 0;
  const void* kw = jt ? (const void*)k_pj_weights : (const void*)k_px_weights;
  const void* kg = jt ? (const void*)k_pj_grad0 : (const void*)k_px_grad0;
  const void* kd = jt ? (const void*)k_pj_dir : (const void*)k_px_dir;
  const void* ks = jt ? (const void*)k_pj_step : (const void*)k_px_step;
  for (int o = 0; o < outer; ++o) {
    Lq.kernel(kw, grid, blk, 0, dP, dcp, nf, cur, o);
    Lq.kernel(kg, grid, blk, 0, dP, dcp, nf, cur);
    Lq.mark(o, false);
    for (int it = 0; it < n_inner; ++it) {
      Lq.kernel(kd, grid, blk, 0, dP, dcp, nf, it, o);
      Lq.kernel(ks, grid, blk, 0, dP, dcp, nf, it, cur, o);
    }
    Lq.mark(o, true);
    Lq.loop_iters.push_back(n_inner);
    cur ^= n_inner & 1;
  }
  Lq.kernel(jt ? (const void*)k_pj_finalize : (const void*)k_px_finalize, grid, blk, 0, dP, dcp, nf, cur);
  run_launches(ctx, Lq, ctx->graphs && !ctx->no_alloc, st);
  arm(ctx, sl, st);
  if (out.stage_iters || out.aniso) {
    std::vector<PxState> hp(nf);
    std::vector<FieldCG> hs(nf);
    HIPCHK(ctx, hipMemcpyAsync(hp.data(), dpx, sizeof(PxState) * nf, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipMemcpyAsync(hs.data(), dstate, sizeof(FieldCG) * nf, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    for (int i = 0; i < nf; ++i) {
      if (out.stage_iters)
        for (int o = 0; o < outer; ++o) out.stage_iters[i * outer + o] = o < 16 ? hp[i].iters[o] : -1;
      if (out.aniso) out.aniso[i] = hs[i].aniso;
    }
  }
}

// VariationalRefine with temporal neighbours for one field (the per-pixel
// solve of a batch of one).
void refine_temporal(hvd_gpu_ctx* ctx, const hvd_gpu_plane& pl, const hvd_gpu_neighbor* nbrs, int nn,
                     const hvd_solver_params& p, float nu, float eps_t, int* stage_iters, hipStream_t st) {
  Carve cv;
  const size_t o_row = cv.take(sizeof(float2) * pl.lines), o_vsw = cv.take(pl.lines),
               o_col = cv.take(sizeof(float2) * pl.width);
  DevField f{};
  f.h = pl.lines;
  f.w = pl.width;
  f.s = pl.s;
  f.carrier = reinterpret_cast<const float2*>(pl.carrier);
  f.chi[0] = reinterpret_cast<float2*>(pl.chi);
  f.y = pl.y;
  const std::vector<DevField> F{f};
  const std::vector<std::vector<hvd_gpu_neighbor>> N{std::vector<hvd_gpu_neighbor>(nbrs, nbrs + nn)};
  ensure_ws(ctx, px_layout(cv, F, {nn}, p.diag_prior > 0.0, p.is_pal).end, st);
  char* ws = ctx->ws;
  std::vector<DevField> G(F);
  if (!pl.carrier) {
    run_prepare(ctx, {plane_tables(pl, (float2*)(ws + o_row), (int8_t*)(ws + o_vsw), (float2*)(ws + o_col))}, false,
                st);
    G[0].rowph = (const float2*)(ws + o_row);
    G[0].colph = (const float2*)(ws + o_col);
    G[0].vsw = pl.vswitch ? (const int8_t*)(ws + o_vsw) : nullptr;
  }
  refine_pixel(ctx, G, N, p, nu, eps_t, st, cv, RefineOut{stage_iters, nullptr});
}

}  // namespace

namespace {

// The autotuner: solves `n` synthetic fields of one shape with every
// candidate launch shape (strip lines x fields per solve group) and keeps the
// one with the least CG-loop time per pixel-iteration, timed with the HIP
// events of the profiling record.  The fields' data are smooth patterns with
// a modulated chroma term; the solve's cost does not depend on them because
// cg_tol 0 runs every iteration.
struct TuneResult {
  SolverTuning t;
  double ns = 0.0;
};

TuneResult autotune(hvd_gpu_ctx* ctx, int h, int w, bool pal, int n, int iterations) {
  const size_t np = (size_t)h * w;
  std::vector<float> hs(np), hchi(2 * np), hth(h);
  std::vector<int8_t> hvs(h);
  for (int y = 0; y < h; ++y) {
    hth[y] = (float)(0.5 * kPi * y);
    hvs[y] = (y & 1) ? -1 : 1;
    for (int x = 0; x < w; ++x) {
      const size_t i = (size_t)y * w + x;
      hs[i] = (float)(30.0 + 20.0 * std::sin(0.21 * x + 0.37 * y) + 8.0 * std::sin(0.5 * kPi * x + hth[y]));
      hchi[2 * i] = (float)(4.0 * std::sin(0.013 * x));
      hchi[2 * i + 1] = (float)(3.0 * std::cos(0.017 * y));
    }
  }
  // device copies: one composite / theta / V-switch table shared by every
  // field (read only), a pristine chi0 and each field's own chi
  const size_t b_s = sizeof(float) * np, b_c = sizeof(float2) * np;
  char* buf = nullptr;
  HIPCHK(ctx, hipMalloc(&buf, b_s + sizeof(float) * h + h + b_c * (n + 1) + 256));
  struct Free {
    char* p;
    ~Free() { (void)hipFree(p); }
  } guard{buf};
  float* ds = (float*)buf;
  float* dth = (float*)(buf + b_s);
  int8_t* dvs = (int8_t*)(buf + b_s + sizeof(float) * h);
  float* dchi0 = (float*)(buf + ((b_s + sizeof(float) * h + h + 255) / 256) * 256);
  HIPCHK(ctx, hipMemcpy(ds, hs.data(), b_s, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(dth, hth.data(), sizeof(float) * h, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(dvs, hvs.data(), h, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(dchi0, hchi.data(), b_c, hipMemcpyHostToDevice));
  std::vector<hvd_gpu_plane> planes(n);
  for (int i = 0; i < n; ++i) {
    hvd_gpu_plane& pl = planes[i];
    pl = hvd_gpu_plane{};
    pl.lines = h;
    pl.width = w;
    pl.s = ds;
    pl.theta = dth;
    pl.vswitch = pal ? dvs : nullptr;
    pl.active_video_start = 0;
    pl.phase_per_sample = 0.5 * kPi;
    pl.sample_rate = 4.0 * 3579545.0;
    pl.chi = dchi0 + 2 * np * (size_t)(i + 1);
    pl.y = nullptr;
  }
  hvd_solver_params p{};
  p.lambda_c = 1.0;
  p.chroma_aniso = 0.7;
  p.charbonnier_eps = 0.5;
  p.chroma_eps = 1.0;
  p.structure_coupling = 0.25;
  p.cg_iterations = iterations;
  p.irls_outer = 1;
  p.cg_tol = 1e-30;
  p.fast = 0;
  p.is_pal = pal ? 1 : 0;
  p.lambda_y = 1.0;
  // candidates: fields per group from ceil(n / G)
  std::vector<int> per;
  for (int G = 1; G <= n; ++G) {
    const int k = (n + G - 1) / G;
    const double px = (double)k * np;
    if ((px >= 2.5e5 && px <= 1.2e7) || k == n)
      if (per.empty() || per.back() != k) per.push_back(k);
  }
  // fine-pass strip heights: every one from 2 to 32 (how evenly a height
  // divides the field and how many waves it leaves both matter)
  std::vector<int> rows_c;
  for (int r = kMinStripRows; r <= kMaxStripRows; ++r) rows_c.push_back(r);
  // the user's fixed tuning and profiling state are restored on every exit
  struct Restore {
    hvd_gpu_ctx* c;
    SolverTuning fixed;
    bool prof;
    ~Restore() {
      c->fixed = fixed;
      clear_profile(c);
      c->profiling = prof;
    }
  } restore{ctx, ctx->fixed, ctx->profiling};
  HIPCHK(ctx, hipDeviceSynchronize());
  clear_profile(ctx);
  ctx->profiling = true;
  auto time_one = [&](const SolverTuning& t) {
    ctx->fixed = t;
    for (int i = 0; i < n; ++i) HIPCHK(ctx, hipMemcpyAsync(planes[i].chi, dchi0, b_c, hipMemcpyDeviceToDevice, nullptr));
    refine_entry(ctx, planes.data(), n, &p, nullptr, nullptr, nullptr, nullptr);
    HIPCHK(ctx, hipDeviceSynchronize());
    double ms = 0.0;
    for (auto& e : ctx->cg_events) {
      float t_ms = 0.f;
      HIPCHK(ctx, hipEventElapsedTime(&t_ms, e.first, e.second));
      ms += t_ms;
    }
    clear_profile(ctx);
    return 1e6 * ms / ((double)np * n * iterations);  // ns per pixel-iteration
  };
  std::vector<std::pair<double, SolverTuning>> res;
  auto by_time = [](const auto& a, const auto& b) { return a.first < b.first; };
  time_one(SolverTuning{16, (size_t)n * np});  // warm the workspace and code
  // coarse pass: every group size at four strip heights
  for (int k : per)
    for (int r : {4, 8, 16, 32}) {
      SolverTuning t{r, (size_t)k * np};
      res.emplace_back(time_one(t), t);
    }
  std::sort(res.begin(), res.end(), by_time);
  // fine pass: every strip height at the best group size and its neighbours
  const int kbest = (int)(res[0].second.group_px / np);
  const size_t at = std::find(per.begin(), per.end(), kbest) - per.begin();
  for (size_t q = at > 0 ? at - 1 : 0; q <= std::min(at + 1, per.size() - 1); ++q)
    for (int r : rows_c) {
      if (r == 4 || r == 8 || r == 16 || r == 32) continue;
      SolverTuning t{r, (size_t)per[q] * np};
      res.emplace_back(time_one(t), t);
    }
  std::sort(res.begin(), res.end(), by_time);
  // re-time the leaders and keep the best of three
  const size_t lead = std::min<size_t>(res.size(), 4);
  for (size_t i = 0; i < lead; ++i)
    for (int rep = 0; rep < 2; ++rep) res[i].first = std::min(res[i].first, time_one(res[i].second));
  std::sort(res.begin(), res.begin() + lead, by_time);
  TuneResult out;
  out.t = res[0].second;
  out.ns = res[0].first;
  ctx->tuned[{h, w, pal ? 1 : 0}][n] = out.t;
  return out;
}

}  // namespace

extern "C" {

int hvd_gpu_variational_refine(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, const hvd_solver_params* p,
                               int* stage_iters, float* aniso, void* stream) {
  return guarded(ctx, [&] { refine_entry(ctx, planes, n, p, stage_iters, aniso, nullptr, stream); });
}

int hvd_gpu_variational_refine_noalloc(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                       const hvd_solver_params* p, int* stage_iters, float* aniso, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, ctx->scratch != nullptr, "no scratch on device %d", ctx->device);
    REQUIRE(ctx, n <= kNoAllocMaxPlanes, "%d planes: the no-allocation entry point takes at most %d", n,
            kNoAllocMaxPlanes);
    REQUIRE(ctx, !ctx->profiling, "profiling allocates events: switch it off for the no-allocation entry point");
    // the workspace and the upload region become the scratch for this call
    struct Swap {
      hvd_gpu_ctx* c;
      char* ws;
      size_t cap;
      ~Swap() {
        c->ws = ws;
        c->ws_cap = cap;
        c->no_alloc = false;
      }
    } swap{ctx, ctx->ws, ctx->ws_cap};
    ctx->ws = ctx->scratch;
    ctx->ws_cap = HVD_GPU_NOALLOC_SCRATCH_BYTES - kNoAllocDescBytes;
    ctx->na_slot.dev = ctx->scratch + ctx->ws_cap;
    ctx->na_slot.cap = kNoAllocDescBytes;
    ctx->no_alloc = true;
    refine_entry(ctx, planes, n, p, stage_iters, aniso, nullptr, stream);
  });
}

int hvd_gpu_variational_refine_temporal(hvd_gpu_ctx* ctx, const hvd_gpu_plane* plane, const hvd_gpu_neighbor* nbrs,
                                        int nn, const hvd_solver_params* p, float temporal_strength,
                                        float temporal_eps, int* stage_iters, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, plane != nullptr, "null plane");
    REQUIRE(ctx, nn >= 0 && nn <= kMaxNeighbors, "neighbour count %d outside 0..%d", nn, kMaxNeighbors);
    REQUIRE(ctx, nn == 0 || nbrs, "null neighbour array");
    // no neighbours or no strength: the weights stay zero and every temporal
    // term vanishes (variational.cpp:221, :452): the 2-D solve
    if (nn == 0 || !(temporal_strength > 0.0f)) {
      refine_entry(ctx, plane, 1, p, stage_iters, nullptr, nullptr, stream);
      return;
    }
    check_params(ctx, p);
    check_plane(ctx, *plane, 0);
    REQUIRE(ctx, temporal_eps > 0.0f, "temporal_eps must be positive (resolve it first, as sequence.cpp:445-446 does)");
    for (int j = 0; j < nn; ++j)
      REQUIRE(ctx, nbrs[j].composite && nbrs[j].carrier && nbrs[j].confidence, "neighbour %d: null plane", j);
    refine_temporal(ctx, *plane, nbrs, nn, *p, temporal_strength, temporal_eps, stage_iters, S(stream));
  });
}

}  // extern "C"

namespace {

// The batched per-pixel solves: VariationalRefine with neighbours or
// VariationalRefineJoint (anchors non-null or joint set), n planes, plane i
// with nn[i] neighbours from nbrs (back to back).
void pixel_batch_entry(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, const hvd_gpu_neighbor* nbrs,
                       const int* nn, const hvd_gpu_anchor* anchors, bool joint, const hvd_solver_params* p,
                       float nu, float eps_t, float nu_a, int* stage_iters, void* stream) {
  REQUIRE(ctx, n >= 0 && (n == 0 || planes), "null argument");
  check_params(ctx, p);
  if (n == 0) return;
  REQUIRE(ctx, n <= 65535, "%d planes in one call (at most 65535)", n);
  std::vector<std::vector<hvd_gpu_neighbor>> N(n);
  int off = 0;
  bool any = false;
  for (int i = 0; i < n; ++i) {
    check_plane(ctx, planes[i], i);
    const int k = nn ? nn[i] : 0;
    REQUIRE(ctx, k >= 0 && k <= kMaxNeighbors, "plane %d: neighbour count %d outside 0..%d", i, k, kMaxNeighbors);
    REQUIRE(ctx, k == 0 || nbrs, "null neighbour array");
    for (int j = 0; j < k; ++j) {
      const hvd_gpu_neighbor& b = nbrs[off + j];
      REQUIRE(ctx, b.composite && b.carrier && b.confidence, "plane %d neighbour %d: null plane", i, j);
      N[i].push_back(b);
    }
    off += k;
    any = any || k > 0;
    if (joint) REQUIRE(ctx, planes[i].y, "plane %d: the joint solve needs Y0 in planes[i].y", i);
    if (anchors && anchors[i].weight)
      REQUIRE(ctx, anchors[i].luma && anchors[i].chroma, "plane %d: incomplete anchor", i);
  }
  if (any && nu > 0.f)
    REQUIRE(ctx, eps_t > 0.f, "temporal_eps must be positive (resolve it first, as sequence.cpp:445-446 does)");
  hipStream_t st = S(stream);
  Carve cv;
  std::vector<size_t> o_row(n), o_vsw(n), o_col(n);
  std::vector<DevField> F(n);
  for (int i = 0; i < n; ++i) {
    o_row[i] = cv.take(sizeof(float2) * planes[i].lines);
    o_vsw[i] = cv.take(planes[i].lines);
    o_col[i] = cv.take(sizeof(float2) * planes[i].width);
    DevField f{};
    f.h = planes[i].lines;
    f.w = planes[i].width;
    f.s = planes[i].s;
    f.carrier = reinterpret_cast<const float2*>(planes[i].carrier);
    f.chi[0] = reinterpret_cast<float2*>(planes[i].chi);
    f.y = joint ? nullptr : planes[i].y;
    F[i] = f;
  }
  std::vector<int> nnv(n);
  for (int i = 0; i < n; ++i) nnv[i] = (int)N[i].size();
  ensure_ws(ctx, px_layout(cv, F, nnv, p->diag_prior > 0.0, p->is_pal, joint).end, st);
  char* ws = ctx->ws;
  std::vector<PrepDesc> prep;
  for (int i = 0; i < n; ++i)
    if (!planes[i].carrier) {
      prep.push_back(plane_tables(planes[i], (float2*)(ws + o_row[i]), (int8_t*)(ws + o_vsw[i]),
                                  (float2*)(ws + o_col[i])));
      F[i].rowph = (const float2*)(ws + o_row[i]);
      F[i].colph = (const float2*)(ws + o_col[i]);
      F[i].vsw = planes[i].vswitch ? (const int8_t*)(ws + o_vsw[i]) : nullptr;
    }
  run_prepare(ctx, prep, false, st);
  std::vector<PxJointField> J;
  if (joint) {
    J.resize(n);
    for (int i = 0; i < n; ++i) {
      J[i].y = planes[i].y;
      if (anchors && anchors[i].weight) {
        J[i].a_luma = anchors[i].luma;
        J[i].a_chroma = reinterpret_cast<const float2*>(anchors[i].chroma);
        J[i].a_weight = anchors[i].weight;
      }
    }
  }
  refine_pixel(ctx, F, N, *p, nu, eps_t, st, cv, RefineOut{stage_iters, nullptr}, joint ? &J : nullptr, nu_a);
}

}  // namespace

extern "C" {

int hvd_gpu_variational_refine_temporal_batch(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                              const hvd_gpu_neighbor* nbrs, const int* nn,
                                              const hvd_solver_params* p, float temporal_strength,
                                              float temporal_eps, int* stage_iters, void* stream) {
  return guarded(ctx, [&] {
    pixel_batch_entry(ctx, planes, n, nbrs, nn, nullptr, false, p, temporal_strength, temporal_eps, 0.f,
                      stage_iters, stream);
  });
}

int hvd_gpu_variational_refine_joint(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                     const hvd_gpu_neighbor* nbrs, const int* nn, const hvd_gpu_anchor* anchors,
                                     const hvd_solver_params* p, float temporal_strength, float temporal_eps,
                                     float nr_anchor, int* stage_iters, void* stream) {
  return guarded(ctx, [&] {
    pixel_batch_entry(ctx, planes, n, nbrs, nn, anchors, true, p, temporal_strength, temporal_eps, nr_anchor,
                      stage_iters, stream);
  });
}

int hvd_gpu_variational_refine_stages(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                      const hvd_solver_params* p, float* const* stages, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n == 0 || stages, "null stage array");
    REQUIRE(ctx, !p || p->diag_prior == 0.0, "the stage-saving solve (the adjoint's forward) takes diag_prior 0");
    refine_entry(ctx, planes, n, p, nullptr, nullptr, stages, stream);
  });
}

int hvd_gpu_refine_adjoint(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, const hvd_solver_params* p,
                           const float* const* stages, const float* const* gchi, const float* const* gy,
                           int adj_iterations, double adj_tol, float* const* gs, float* const* gchi0, double* glam,
                           void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (planes && stages && gchi && gs && gchi0 && glam)), "null argument");
    check_params(ctx, p);
    REQUIRE(ctx, p->diag_prior == 0.0, "the adjoint takes diag_prior 0");
    REQUIRE(ctx, adj_iterations >= 1, "adj_iterations must be positive");
    REQUIRE(ctx, adj_tol >= 0.0, "negative adj_tol");
    if (n == 0) return;
    run_adjoint(ctx, planes, n, *p, stages, gchi, gy, adj_iterations, adj_tol, gs, gchi0, glam, S(stream));
  });
}

}  // extern "C"

namespace {

// Workspace layout of the field driver, carved after `cv`; `end` covers the
// solver's carve too, so one ensure_ws(end) sizes the whole call.
// The order of a decode batch's solve: PAL-family fields first, then NTSC,
// each in batch order, so the solve groups (runs of consecutive fields) mix
// the standards at one boundary at most.  Each field keeps its own buffers.
void solve_order(std::vector<DevField>& F) {
  std::stable_partition(F.begin(), F.end(), [](const DevField& f) { return f.pal == 2; });
}

struct DecodeLayout {
  std::vector<size_t> ire, th, vs, s, row, vsa, col, spec, var, tha;
  Carve solver;  // where refine_planes starts carving
  size_t end = 0;
};

DecodeLayout decode_layout(const hvd_gpu_ctx* ctx, int n, const hvd_field_geometry* geom, bool from_raw, Carve cv,
                           bool diag) {
  DecodeLayout L;
  for (auto* v : {&L.ire, &L.th, &L.vs, &L.s, &L.row, &L.vsa, &L.col, &L.spec, &L.var, &L.tha}) v->resize(n);
  std::vector<DevField> shapes(n);
  for (int i = 0; i < n; ++i) {
    const hvd_field_geometry& g = geom[i];
    const int lines = last_line(g) - g.first_active_field_line, width = g.active_video_end - g.active_video_start;
    const int pad = lines > 1 ? std::min(16, lines - 1) : 0;
    const size_t np = (size_t)lines * width;
    L.ire[i] = from_raw ? cv.take(sizeof(float) * (size_t)g.field_width * g.field_height) : 0;
    L.th[i] = cv.take(sizeof(float) * g.field_height);
    L.vs[i] = cv.take(g.field_height);
    L.s[i] = cv.take(sizeof(float) * np);
    L.row[i] = cv.take(sizeof(float2) * lines);
    L.vsa[i] = cv.take(lines);
    L.col[i] = cv.take(sizeof(float2) * width);
    L.spec[i] = cv.take(sizeof(float2) * (size_t)(lines + 2 * pad) * width);
    L.var[i] = cv.take(sizeof(float2) * np * 2);
    L.tha[i] = cv.take(sizeof(float) * lines);
    shapes[i].h = lines;
    shapes[i].w = width;
    shapes[i].pal = 1 + (uses_vswitch(g) ? 1 : 0);
  }
  L.solver = cv;
  // one solve over the batch (see decode_fields_run), from L.solver
  solve_order(shapes);
  L.end = refine_ws_end(ctx, cv, shapes, 0, diag);
  return L;
}

void check_fields(hvd_gpu_ctx* ctx, int n, const hvd_field_geometry* geom) {
  REQUIRE(ctx, n <= 65535, "%d fields in one call (at most 65535)", n);
  for (int i = 0; i < n; ++i) {
    const hvd_field_geometry& g = geom[i];
    check_geom(ctx, g, i);
    check_active(ctx, g, i);
    const int lines = last_line(g) - g.first_active_field_line, width = g.active_video_end - g.active_video_start;
    REQUIRE(ctx, width <= 2048 && lines + 32 <= 1024, "field %d: active area exceeds the FFT limits", i);
  }
}

// The field driver proper.  The workspace must already hold L.end bytes.
// dec_out (optional) receives, per field, the pointers the frame weave reads.
// carriers (optional, lines x width per field): the fields' carrier planes
// are materialised there with hvd-core's float MakeCarrier / MakeCarrierPal
// arithmetic (ntsc_geometry.cpp:29-47, :75-93) and the inits read them;
// else the inits and the solve use the exact factored carrier tables.
// solve = false stops after the inits.
void decode_fields_run(hvd_gpu_ctx* ctx, const uint16_t* const* raw, const float* const* field_ire, int n,
                       const hvd_field_geometry* geom, float blanking, float white, const hvd_solver_params* p,
                       float chroma_phase_deg, float* const* y_out, float* const* chi_out, float* const* theta_out,
                       hipStream_t st, const DecodeLayout& L, std::vector<BridgeField>* dec_out,
                       bool solve = true, float2* const* carriers = nullptr) {
  char* ws = ctx->ws;
  REQUIRE(ctx, ctx->ws_cap >= L.end, "internal: workspace not sized");
  // (1) codes -> IRE, every field in one launch
  std::vector<const float*> fields(n);
  if (raw) {
    std::vector<IreDesc> D(n);
    size_t most = 0;
    for (int i = 0; i < n; ++i) {
      const hvd_field_geometry& g = geom[i];
      D[i] = IreDesc{raw[i], (float*)(ws + L.ire[i]), (size_t)g.field_width * g.field_height};
      most = std::max(most, D[i].n);
      fields[i] = D[i].out;
    }
    Stage stg;
    const size_t oD = stg.put(D);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    hipLaunchKernelGGL(k_u16_to_ire_batch, dim3((unsigned)std::min<size_t>((most + 255) / 256, 4096), (unsigned)n),
                       dim3(256), 0, st, (const IreDesc*)(base + oD), blanking, (white - blanking) / 100.0f);
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, st);
  } else {
    for (int i = 0; i < n; ++i) fields[i] = field_ire[i];
  }
  // (2) lock-in
  std::vector<LockinGeom> G(n);
  std::vector<float*> T(n);
  std::vector<int8_t*> V(n);
  for (int i = 0; i < n; ++i) {
    G[i] = lockin_geom(geom[i]);
    T[i] = (float*)(ws + L.th[i]);
    V[i] = (int8_t*)(ws + L.vs[i]);
  }
  run_lockin(ctx, G, fields, T, V, st);
  // (3) active crop + carrier tables (PrepareFieldObs + ApplyChromaPhase)
  std::vector<PrepDesc> prep(n);
  for (int i = 0; i < n; ++i) {
    const hvd_field_geometry& g = geom[i];
    PrepDesc d{};
    d.field = fields[i];
    d.theta_field = T[i];
    d.vswitch_field = uses_vswitch(g) ? V[i] : nullptr;
    d.s = (float*)(ws + L.s[i]);
    d.rowph = (float2*)(ws + L.row[i]);
    d.vsw_active = uses_vswitch(g) ? (int8_t*)(ws + L.vsa[i]) : nullptr;
    d.colph = (float2*)(ws + L.col[i]);
    d.fw = g.field_width;
    d.fal = g.first_active_field_line;
    d.a0 = g.active_video_start;
    d.lines = last_line(g) - g.first_active_field_line;
    d.width = g.active_video_end - g.active_video_start;
    d.std_phase = g.subcarrier_hz <= 0.0;
    d.signed_by_parity = uses_vswitch(g);
    d.phase_offset = chroma_phase_deg == 0.0f ? 0.0f : -chroma_phase_deg * 3.14159265358979323846f / 180.0f;
    d.w0 = phase_per_sample(g);
    d.th_active = carriers ? (float*)(ws + L.tha[i]) : nullptr;
    prep[i] = d;
  }
  run_prepare(ctx, prep, true, st);
  if (carriers)
    for (int i = 0; i < n; ++i) {
      hipLaunchKernelGGL(k_make_carrier, dim3((prep[i].width + 255) / 256, prep[i].lines), dim3(256), 0, st,
                         (const float*)prep[i].th_active, (const int8_t*)prep[i].vsw_active, prep[i].width,
                         prep[i].a0, (float)prep[i].w0, carriers[i]);
      LAUNCH_CHECK(ctx);
    }
  if (theta_out)
    for (int i = 0; i < n; ++i)
      if (theta_out[i])
        HIPCHK(ctx, hipMemcpyAsync(theta_out[i], T[i], sizeof(float) * geom[i].field_height, hipMemcpyDeviceToDevice,
                                   st));
  // (4) holographic init into the caller's chi / y
  PlanSet ps;
  std::vector<InitField> IF(n);
  for (int i = 0; i < n; ++i) {
    InitField f{};
    f.h = prep[i].lines;
    f.w = prep[i].width;
    f.pad = f.h > 1 ? std::min(16, f.h - 1) : 0;
    f.hp = f.h + 2 * f.pad;
    f.s = prep[i].s;
    f.rowph = prep[i].rowph;
    f.vsw = prep[i].vsw_active;
    f.colph = prep[i].colph;
    f.carrier = carriers ? carriers[i] : nullptr;
    f.spec = (float2*)(ws + L.spec[i]);
    f.var[0] = (float2*)(ws + L.var[i]);
    f.var[1] = f.var[0] + (size_t)f.h * f.w;
    f.chi0 = reinterpret_cast<float2*>(chi_out[i]);
    f.y0 = y_out[i];
    f.row_plan = ps.get(ctx, f.w, kRowGroup);
    f.col_plan = ps.get(ctx, f.hp, kColGroup);
    f.nvar = 2;
    f.sample_rate = (float)geom[i].sample_rate;
    IF[i] = f;
  }
  run_init(ctx, IF, ps, st);
  if (dec_out) {
    dec_out->resize(n);
    for (int i = 0; i < n; ++i)
      (*dec_out)[i] = BridgeField{IF[i].h, IF[i].w, y_out[i], reinterpret_cast<const float2*>(chi_out[i]),
                                  IF[i].rowph, IF[i].colph, IF[i].vsw, IF[i].carrier};
  }
  if (!solve) return;
  // (5) solve (sequence.cpp:1046: always, with is_pal from the geometry)
  std::vector<DevField> F(n);
  bool any_pal = false, any_ntsc = false;
  for (int i = 0; i < n; ++i) {
    DevField f{};
    f.h = IF[i].h;
    f.w = IF[i].w;
    f.s = IF[i].s;
    f.rowph = IF[i].rowph;
    f.vsw = IF[i].vsw;
    f.colph = IF[i].colph;
    f.chi[0] = IF[i].chi0;
    f.y = y_out[i];
    f.pal = 1 + (uses_vswitch(geom[i]) ? 1 : 0);  // the AUTO aniso grouping per field
    F[i] = f;
    (uses_vswitch(geom[i]) ? any_pal : any_ntsc) = true;
  }
  // one solve over the whole batch, NTSC and PAL fields side by side (in
  // solve_order, as decode_layout sized it)
  solve_order(F);
  hvd_solver_params pp = *p;
  pp.is_pal = any_pal && !any_ntsc ? 1 : 0;
  Carve c = L.solver;
  refine_planes(ctx, F, pp, st, c, RefineOut{});
}

int decode_fields(hvd_gpu_ctx* ctx, const uint16_t* const* raw, const float* const* field_ire, int n,
                  const hvd_field_geometry* geom, float blanking, float white, const hvd_solver_params* p,
                  float chroma_phase_deg, float* const* y_out, float* const* chi_out, float* const* theta_out,
                  void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (geom && y_out && chi_out && (raw || field_ire))), "null argument");
    check_params(ctx, p);
    if (raw) REQUIRE(ctx, white != blanking, "white level equals blanking");
    if (n == 0) return;
    check_fields(ctx, n, geom);
    for (int i = 0; i < n; ++i)
      REQUIRE(ctx, (raw ? (const void*)raw[i] : (const void*)field_ire[i]) && y_out[i] && chi_out[i],
              "field %d: null pointer", i);
    hipStream_t st = S(stream);
    const DecodeLayout L = decode_layout(ctx, n, geom, raw != nullptr, Carve{}, p->diag_prior > 0.0);
    ensure_ws(ctx, L.end, st);
    decode_fields_run(ctx, raw, field_ire, n, geom, blanking, white, p, chroma_phase_deg, y_out, chi_out, theta_out,
                      st, L, nullptr);
  });
}

// HvdConfig's defaults of the window driver's knobs (hvd_config.h:336-418).
hvd_window_params default_window_params() {
  hvd_window_params w{};
  w.temporal_strength = 0.0;
  w.temporal_eps = 0.0;
  w.nr_anchor = 1.0;
  w.nr_eps = 0.0;
  w.coherence_gate = 0.6;
  w.odd_gate_floor = 0.35;
  w.passes = 2;
  w.nr_radius = 2;
  w.mc_tile = 32;
  w.mc_search = 16;
  w.bidirectional = 1;
  w.extended_temporal = 1;
  w.trajectory_fit = 1;
  w.psi_init = 0;
  w.output_fidelity = 1;
  w.parallel_across_fields = 1;
  w.selective_3d = 0;
  w.selective_max_area = 0.45f;
  return w;
}

char* ensure_fwin(hvd_gpu_ctx* ctx, size_t bytes, hipStream_t st) {
  if (bytes <= ctx->fwin_cap) return ctx->fwin;
  HIPCHK(ctx, hipStreamSynchronize(st));
  HIPCHK(ctx, hipDeviceSynchronize());
  if (ctx->fwin) HIPCHK(ctx, hipFree(ctx->fwin));
  ctx->fwin = nullptr;
  ctx->fwin_cap = 0;
  const size_t want = bytes + bytes / 4;
  void* q = nullptr;
  if (hipMalloc(&q, want) != hipSuccess) {
    set_err(ctx, "hipMalloc(%zu) for the frame window failed", want);
    throw Fail{HVD_GPU_ERR_NOMEM};
  }
  ctx->fwin = static_cast<char*>(q);
  ctx->fwin_cap = want;
  return ctx->fwin;
}

// DecodeFrameSequenceWindow (frame_bridge.cpp:722-947) on the device.  The
// 2-D window runs the field driver's batch; a window that may engage 3-D
// (sequence.cpp:526-652: a positive strength, or the adaptive 0 with more
// fields than the ambiguity stride) or asks for the selective window runs
// the window driver on the fields' holographic inits, with the parities of
// the measured field order.
void frame_window(hvd_gpu_ctx* ctx, const int16_t* const* frames, int nframes, int core_begin, int core_end,
                  const hvd_frame_params* fp, const hvd_solver_params* p, const hvd_bridge_params* b,
                  const hvd_window_params* wp_in, const hvd_yc_frame* out, float* acc_gain, hvd_window_diag* diag,
                  hipStream_t st) {
  REQUIRE(ctx, fp && p && b && frames && out, "null argument");
  REQUIRE(ctx, nframes >= 1 && core_begin >= 0 && core_end <= nframes && core_begin < core_end,
          "bad window: %d frames, core [%d, %d)", nframes, core_begin, core_end);
  check_params(ctx, p);
  check_frame(ctx, *fp);
  const int fw = fp->frame_width, fh = fp->frame_height, f1 = fp->field1_lines;
  const hvd_field_geometry g = frame_field_geometry(*fp);
  check_geom(ctx, g, 0);
  check_active(ctx, g, 0);
  const bool pal = uses_vswitch(g);
  const int amb_stride = pal ? 4 : 2;
  const int nf = 2 * nframes, ncore = core_end - core_begin;
  const bool three_d = p->cg_iterations > 0 &&
                       (b->temporal_strength > 0.0f || (b->temporal_strength == 0.0f && nf > amb_stride));
  hvd_window_params wp = wp_in ? *wp_in : default_window_params();
  wp.temporal_strength = b->temporal_strength;
  // DecodeWindowSelective (frame_bridge.cpp:800-805)
  const bool selective = wp.selective_3d && p->cg_iterations > 0 && b->temporal_strength >= 0.0f;
  if (diag) *diag = hvd_window_diag{};
  for (int t = 0; t < nframes; ++t) REQUIRE(ctx, frames[t], "frame %d: null pointer", t);
  for (int k = 0; k < ncore; ++k)
    REQUIRE(ctx, out[k].luma && out[k].chroma && out[k].u_plane && out[k].v_plane, "output %d: null plane", k);
  const int lines = last_line(g) - g.first_active_field_line, width = g.active_video_end - g.active_video_start;
  std::vector<hvd_field_geometry> geoms(nf, g);
  check_fields(ctx, nf, geoms.data());
  Carve cv;
  const size_t o_fields = cv.take(sizeof(float) * (size_t)nf * f1 * fw);
  const size_t o_amps = cv.take(sizeof(float) * nf);
  const size_t o_dec = cv.take(sizeof(int) * 2 + sizeof(float) * 4);
  const size_t o_y = cv.take(sizeof(float) * (size_t)nf * lines * width);
  const size_t o_chi = cv.take(sizeof(float2) * (size_t)nf * lines * width);
  const DecodeLayout L = decode_layout(ctx, nf, geoms.data(), false, cv, p->diag_prior > 0.0);
  char* ws = ensure_ws(ctx, L.end, st);
  float* fields = (float*)(ws + o_fields);
  float* amps = (float*)(ws + o_amps);
  int* f1_bottom = (int*)(ws + o_dec);
  int* vote = f1_bottom + 1;
  float* gain = (float*)(ws + o_dec + 2 * sizeof(int));
  const BridgeFrameParams bp = bridge_params(*fp);
  std::vector<const int16_t*> fr(frames, frames + nframes);
  std::vector<const float*> fptr(nf);
  std::vector<float*> fout(nf);
  for (int j = 0; j < nf; ++j) fout[j] = fields + (size_t)j * f1 * fw;
  for (int j = 0; j < nf; ++j) fptr[j] = fout[j];
  std::vector<LockinGeom> LG(nf, lockin_geom(g));
  std::vector<BridgeOut> bo(ncore);
  for (int k = 0; k < ncore; ++k) bo[k] = BridgeOut{out[k].luma, out[k].chroma, out[k].u_plane, out[k].v_plane};
  Stage stg;
  const size_t oP = stg.put(&bp, 1), oF = stg.put(fr), oFP = stg.put(fptr), oG = stg.put(LG), oFO = stg.put(fout);
  Slot* sl;
  char* base = upload(ctx, stg, st, &sl);
  const BridgeFrameParams* dP = (const BridgeFrameParams*)(base + oP);
  // (1) split every frame into its two IRE fields
  hipLaunchKernelGGL(k_split_fields, dim3((fw + 255) / 256, f1, nframes), dim3(256), 0, st, dP,
                     (const int16_t* const*)(base + oF), fp->blanking_level, (float* const*)(base + oFO));
  LAUNCH_CHECK(ctx);
  // (2) burst amplitude per field, window ACC and field order
  hipLaunchKernelGGL(k_burst_amp, dim3(nf), dim3(256), 0, st, (const LockinGeom*)(base + oG),
                     (const float* const*)(base + oFP), amps);
  LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_window_decisions, dim3(1), dim3(256), 0, st, dP, (const float* const*)(base + oFP), nframes,
                     (const float*)amps, b->acc, b->chroma_gain, g.standard == 1 ? 21.43f : 20.0f, b->field_order,
                     f1_bottom, gain, vote);
  LAUNCH_CHECK(ctx);
  arm(ctx, sl, st);
  // (3) the fields' decode: Y and chi of every field, and the per-field
  // pointers the weave reads
  std::vector<BridgeField> dec;
  const int* w_f1_bottom = f1_bottom;
  const float* w_gain = gain;
  if (!three_d && !selective) {
    std::vector<float*> ys(nf), chis(nf);
    for (int j = 0; j < nf; ++j) {
      ys[j] = (float*)(ws + o_y) + (size_t)j * lines * width;
      chis[j] = (float*)(ws + o_chi) + (size_t)j * lines * width * 2;
    }
    decode_fields_run(ctx, nullptr, fptr.data(), nf, geoms.data(), 0.0f, 100.0f, p, b->chroma_phase_deg, ys.data(),
                      chis.data(), nullptr, st, L, &dec);
  } else {
    // the parities of the measured order (frame_bridge.cpp:766-774): block 1
    // of every frame is the top field unless field 1 is the bottom one
    int fb = 0;
    HIPCHK(ctx, hipMemcpyAsync(&fb, f1_bottom, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    std::vector<int> par(nf);
    for (int j = 0; j < nf; ++j) par[j] = (j % 2 == 0) == (fb == 0) ? 0 : 1;
    // the planes the driver works on live in their own pool: its component
    // calls may reallocate the workspace.  The carriers are hvd-core's float
    // MakeCarrier planes (PrepareFieldObs, sequence.cpp:229-268), which the
    // inits, the driver and the weave all read.
    const size_t np = (size_t)lines * width;
    Carve sc;
    const size_t s_s = sc.take(sizeof(float) * nf * np), s_car = sc.take(sizeof(float2) * nf * np),
                 s_y = sc.take(sizeof(float) * nf * np), s_chi = sc.take(sizeof(float2) * nf * np),
                 s_dec = sc.take(sizeof(int) * 2 + sizeof(float) * 4);
    char* fwn = ensure_fwin(ctx, sc.off, st);
    float* Ss = (float*)(fwn + s_s);
    float2* Car = (float2*)(fwn + s_car);
    std::vector<float*> ys(nf), chis(nf);
    std::vector<float2*> cars(nf);
    for (int j = 0; j < nf; ++j) {
      ys[j] = (float*)(fwn + s_y) + (size_t)j * np;
      chis[j] = (float*)(fwn + s_chi) + (size_t)j * np * 2;
      cars[j] = Car + (size_t)j * np;
    }
    std::vector<BridgeField> d0;
    decode_fields_run(ctx, nullptr, fptr.data(), nf, geoms.data(), 0.0f, 100.0f, p, b->chroma_phase_deg, ys.data(),
                      chis.data(), nullptr, st, L, &d0, /*solve=*/false, cars.data());
    for (int j = 0; j < nf; ++j)
      HIPCHK(ctx, hipMemcpyAsync(Ss + (size_t)j * np, ws + L.s[j], sizeof(float) * np, hipMemcpyDeviceToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(fwn + s_dec, ws + o_dec, sizeof(int) * 2 + sizeof(float) * 4,
                               hipMemcpyDeviceToDevice, st));
    std::vector<hvd_gpu_plane> planes(nf);
    std::vector<int> y_given(nf, 1);
    for (int j = 0; j < nf; ++j) {
      hvd_gpu_plane& q = planes[j];
      q = hvd_gpu_plane{};
      q.lines = lines;
      q.width = width;
      q.s = Ss + (size_t)j * np;
      q.carrier = reinterpret_cast<const float*>(cars[j]);
      q.active_video_start = g.active_video_start;
      q.phase_per_sample = phase_per_sample(g);
      q.sample_rate = g.sample_rate;
      q.chi = chis[j];
      q.y = ys[j];
    }
    const int rc = selective ? hvd_window::drive_selective(ctx, planes.data(), nf, par.data(), pal ? 1 : 0,
                                                           y_given.data(), p, &wp, diag, st)
                             : hvd_window::drive(ctx, planes.data(), nf, par.data(), pal ? 1 : 0, y_given.data(), p,
                                                 &wp, diag, st);
    if (rc != HVD_GPU_OK) throw Fail{rc};
    dec.resize(nf);
    for (int j = 0; j < nf; ++j)
      dec[j] = BridgeField{lines, width, ys[j], reinterpret_cast<const float2*>(chis[j]), nullptr, nullptr, nullptr,
                           cars[j]};
    w_f1_bottom = (const int*)(fwn + s_dec);
    w_gain = (const float*)(fwn + s_dec + 2 * sizeof(int));
  }
  // (4) weave + requantise the core frames
  Stage stg2;
  const size_t oD = stg2.put(dec), oF2 = stg2.put(fr), oP2 = stg2.put(&bp, 1), oO2 = stg2.put(bo);
  Slot* sl2;
  char* base2 = upload(ctx, stg2, st, &sl2);
  hipLaunchKernelGGL(k_weave_requant, dim3((fw + 255) / 256, fh, ncore), dim3(256), 0, st,
                     (const BridgeFrameParams*)(base2 + oP2), (const int16_t* const*)(base2 + oF2), core_begin,
                     (const BridgeField*)(base2 + oD), w_f1_bottom, w_gain, b->monochrome,
                     (const BridgeOut*)(base2 + oO2));
  LAUNCH_CHECK(ctx);
  arm(ctx, sl2, st);
  if (acc_gain) {
    HIPCHK(ctx, hipMemcpyAsync(acc_gain, w_gain, sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
  }
}

}  // namespace

extern "C" {

int hvd_gpu_decode_frame_window_s16(hvd_gpu_ctx* ctx, const int16_t* const* frames, int nframes, int core_begin,
                                    int core_end, const hvd_frame_params* fp, const hvd_solver_params* p,
                                    const hvd_bridge_params* b, const hvd_yc_frame* out, float* acc_gain,
                                    void* stream) {
  return guarded(ctx, [&] {
    frame_window(ctx, frames, nframes, core_begin, core_end, fp, p, b, nullptr, out, acc_gain, nullptr, S(stream));
  });
}

int hvd_gpu_decode_frame_window_ex_s16(hvd_gpu_ctx* ctx, const int16_t* const* frames, int nframes, int core_begin,
                                       int core_end, const hvd_frame_params* fp, const hvd_solver_params* p,
                                       const hvd_bridge_params* b, const hvd_window_params* wp,
                                       const hvd_yc_frame* out, float* acc_gain, hvd_window_diag* diag,
                                       void* stream) {
  return guarded(ctx, [&] {
    frame_window(ctx, frames, nframes, core_begin, core_end, fp, p, b, wp, out, acc_gain, diag, S(stream));
  });
}

int hvd_gpu_decode_fields_u16(hvd_gpu_ctx* ctx, const uint16_t* const* raw, int n, const hvd_field_geometry* geom,
                              float blanking, float white, const hvd_solver_params* p, float chroma_phase_deg,
                              float* const* y_out, float* const* chi_out, float* const* theta_out, void* stream) {
  if (!raw && n > 0) return HVD_GPU_ERR_INVALID;
  return decode_fields(ctx, raw, nullptr, n, geom, blanking, white, p, chroma_phase_deg, y_out, chi_out, theta_out,
                       stream);
}

int hvd_gpu_decode_fields_ire(hvd_gpu_ctx* ctx, const float* const* field_ire, int n, const hvd_field_geometry* geom,
                              const hvd_solver_params* p, float chroma_phase_deg, float* const* y_out,
                              float* const* chi_out, float* const* theta_out, void* stream) {
  if (!field_ire && n > 0) return HVD_GPU_ERR_INVALID;
  return decode_fields(ctx, nullptr, field_ire, n, geom, 0.0f, 100.0f, p, chroma_phase_deg, y_out, chi_out,
                       theta_out, stream);
}

int hvd_gpu_block_match_sad_u16(hvd_gpu_ctx* ctx, const uint16_t* const* ref, const uint16_t* const* cur, int n,
                                int h, int w, int bh, int bw, int sy, int sx, int16_t* const* vy, int16_t* const* vx,
                                int32_t* const* sad, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (ref && cur && vy && vx && sad)), "null argument");
    REQUIRE(ctx, h >= 1 && w >= 1, "empty plane");
    REQUIRE(ctx, bw >= 1 && bw <= 32 && bh >= 1 && bh <= 16, "block size %d x %d outside 1..32 x 1..16", bw, bh);
    REQUIRE(ctx, sx >= 0 && sx <= 32 && sy >= 0 && sy <= 16, "search range %d x %d outside 0..32 x 0..16", sx, sy);
    if (n == 0) return;
    std::vector<SadDesc> D(n);
    std::vector<int> pre(n + 1, 0);
    const int nby = (h + bh - 1) / bh, nbx = (w + bw - 1) / bw;
    for (int i = 0; i < n; ++i) {
      REQUIRE(ctx, ref[i] && cur[i] && vy[i] && vx[i] && sad[i], "pair %d: null pointer", i);
      SadDesc d{};
      d.ref = ref[i];
      d.cur = cur[i];
      d.vy = vy[i];
      d.vx = vx[i];
      d.sad = sad[i];
      d.h = h;
      d.w = w;
      d.bh = bh;
      d.bw = bw;
      d.sy = sy;
      d.sx = sx;
      d.nby = nby;
      d.nbx = nbx;
      D[i] = d;
      pre[i + 1] = pre[i] + nby * ((nbx + 3) / 4);
    }
    Stage stg;
    const size_t oD = stg.put(D), oP = stg.put(pre);
    Slot* sl;
    char* base = upload(ctx, stg, S(stream), &sl);
    hipLaunchKernelGGL(k_sad_bm, dim3(pre[n]), dim3(256), 0, S(stream), (const SadDesc*)(base + oD),
                       (const int*)(base + oP), n);
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, S(stream));
  });
}

}  // extern "C"

namespace {

// motion.cpp:286-298: 0, -1, +1, -2, +2, ... within [lo, hi].
std::vector<int> range_from_centre(int lo, int hi) {
  std::vector<int> v;
  if (lo > hi) return v;
  if (lo <= 0 && 0 <= hi) v.push_back(0);
  for (int d = 1; d <= std::max(-lo, hi); ++d) {
    if (-d >= lo && -d <= hi) v.push_back(-d);
    if (d >= lo && d <= hi) v.push_back(d);
  }
  if (v.empty())
    for (int i = lo; i <= hi; ++i) v.push_back(i);
  return v;
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

// SSD of every (pair, tile, candidate): a thread each for small tiles, a wave
// each (parallel products, the same sequential sum) from 16 x 16 up.
void launch_me_ssd(hipStream_t st, int n, int h, int w, int tile, const float* P, const int* dys, const int* dxs,
                   int ndy, int ndx, const float* bdy, const float* bdx, float* se) {
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile;
  const size_t items = (size_t)n * th * tw * ndy * ndx;
  if (tile >= 16 && tile <= kSsdTile && ndy * ndx >= 8 && tile + ndy - 1 <= kSsdWin && tile + ndx - 1 <= kSsdWin)
    hipLaunchKernelGGL(k_me_ssd_tile, dim3((unsigned)((size_t)n * th * tw)), dim3(64), 0, st, n, h, w, tile, P, dys,
                       dxs, ndy, ndx, bdy, bdx, se);
  else if (tile >= 16)
    hipLaunchKernelGGL(k_me_ssd_wave, dim3((unsigned)((items + 3) / 4)), dim3(256), 0, st, n, h, w, tile, P, dys, dxs,
                       ndy, ndx, bdy, bdx, se);
  else
    hipLaunchKernelGGL(k_me_ssd, dim3(blocks_for(items)), dim3(256), 0, st, n, h, w, tile, P, dys, dxs, ndy, ndx,
                       bdy, bdx, se);
}

}  // namespace

extern "C" {

int hvd_gpu_estimate_motion(hvd_gpu_ctx* ctx, const float* const* y_ref, const float* const* y_cur, int n, int h,
                            int w, int tile, int search, float* const* dy, float* const* dx, float* const* conf,
                            void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (y_ref && y_cur && dy && dx && conf)), "null argument");
    REQUIRE(ctx, h >= 4 && w >= 4, "planes must be at least 4 x 4 (the coarse level decimates by 4)");
    REQUIRE(ctx, tile >= 1 && tile <= 256, "tile %d outside 1..256", tile);
    REQUIRE(ctx, search >= 0 && search <= 64, "search %d outside 0..64", search);
    if (n == 0) return;
    hipStream_t st = S(stream);
    constexpr int f = 4;
    const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw;
    const int hh = h / f, ww = w / f;
    const int ct = std::max(4, tile / f);
    const int cth = (hh + ct - 1) / ct, ctw = (ww + ct - 1) / ct, nct = cth * ctw;
    const int cs = std::max(1, (search + f - 1) / f);
    const std::vector<int> rc = range_from_centre(-cs, cs), r3 = range_from_centre(-3, 3), r0 = {0};
    const size_t np = (size_t)h * w;
    const size_t nse = std::max((size_t)nct * rc.size() * rc.size(), (size_t)nt * r3.size() * r3.size());
    Carve cv;
    const size_t o_tmp = cv.take(sizeof(float) * n * 2 * np), o_P = cv.take(sizeof(float) * n * 2 * np);
    const size_t o_Pd = cv.take(sizeof(float) * n * 2 * (size_t)hh * ww);
    const size_t o_se = cv.take(sizeof(float) * n * nse);
    const size_t o_c = cv.take(sizeof(float) * n * nct * 3);
    const size_t o_v = cv.take(sizeof(float) * n * nt * 19);
    char* ws = ensure_ws(ctx, cv.off, st);
    float* tmp = (float*)(ws + o_tmp);
    float* P = (float*)(ws + o_P);
    float* Pd = (float*)(ws + o_Pd);
    float* se = (float*)(ws + o_se);
    float *cdy = (float*)(ws + o_c), *cdx = cdy + n * nct, *cbest = cdx + n * nct;
    float* v = (float*)(ws + o_v);
    auto vec = [&](int k) { return v + (size_t)k * n * nt; };
    float *bdy = vec(0), *bdx = vec(1), *mdy = vec(2), *mdx = vec(3), *best = vec(4), *tdy = vec(5), *tdx = vec(6);
    float *jy = vec(7), *jx = vec(8), *cm[4] = {vec(9), vec(10), vec(11), vec(12)}, *best0 = vec(13), *cost0 = vec(14);
    float *ody = vec(15), *odx = vec(16), *scratch = vec(17);  // scratch spans vec(17..18)
    std::vector<MotionDesc> D(n);
    for (int i = 0; i < n; ++i) {
      REQUIRE(ctx, y_ref[i] && y_cur[i] && dy[i] && dx[i] && conf[i], "pair %d: null pointer", i);
      D[i] = MotionDesc{y_ref[i], y_cur[i]};
    }
    std::vector<float*> cptr(conf, conf + n);
    Stage stg;
    const size_t oD = stg.put(D), oRc = stg.put(rc), oR3 = stg.put(r3), oR0 = stg.put(r0), oC = stg.put(cptr);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    const int* drc = (const int*)(base + oRc);
    const int* dr3 = (const int*)(base + oR3);
    const int* dr0 = (const int*)(base + oR0);
    const int nrc = (int)rc.size(), nr3 = (int)r3.size();
    // (1) BoxBlur r = 2 of both planes (:398-399)
    launch_me_blur(st, (const MotionDesc*)(base + oD), n, h, w, 2, tmp, P);
    LAUNCH_CHECK(ctx);
    // (2) 4x-decimated exhaustive coarse search (:405-414)
    hipLaunchKernelGGL(k_me_decimate, dim3(blocks_for((size_t)n * 2 * hh * ww)), dim3(256), 0, st, n, h, w, f,
                       (const float*)P, Pd);
    LAUNCH_CHECK(ctx);
    launch_me_ssd(st, n, hh, ww, ct, Pd, drc, drc, nrc, nrc, nullptr, nullptr, se);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_me_select, dim3(blocks_for((size_t)n * nct)), dim3(256), 0, st, n, nct, drc, drc, nrc, nrc,
                       (const float*)se, (const float*)nullptr, (const float*)nullptr, cdy, cdx, cbest, (float*)nullptr);
    LAUNCH_CHECK(ctx);
    // (3) upscale onto the full-res tile grid, x 4 (:416-428)
    hipLaunchKernelGGL(k_me_upscale, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n, th, tw, cth, ctw, f,
                       (const float*)cdy, (const float*)cdx, bdy, bdx);
    LAUNCH_CHECK(ctx);
    // (4) +/-3 refinement around the coarse vector (:431-435)
    launch_me_ssd(st, n, h, w, tile, P, dr3, dr3, nr3, nr3, bdy, bdx, se);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_me_select, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n, nt, dr3, dr3, nr3, nr3,
                       (const float*)se, (const float*)bdy, (const float*)bdx, mdy, mdx, best, (float*)nullptr);
    LAUNCH_CHECK(ctx);
    // (5) the four costs of the parabolic fit (:438-441)
    const float offs[4][2] = {{-1.f, 0.f}, {1.f, 0.f}, {0.f, -1.f}, {0.f, 1.f}};
    for (int k = 0; k < 4; ++k) {
      hipLaunchKernelGGL(k_me_offset, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n * nt, (const float*)mdy,
                         (const float*)mdx, offs[k][0], offs[k][1], tdy, tdx);
      LAUNCH_CHECK(ctx);
      launch_me_ssd(st, n, h, w, tile, P, dr0, dr0, 1, 1, tdy, tdx, se);
      LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_me_select, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n, nt, dr0, dr0, 1, 1,
                         (const float*)se, (const float*)tdy, (const float*)tdx, jy, jx, cm[k], (float*)nullptr);
      LAUNCH_CHECK(ctx);
    }
    // (6) explicit zero-motion evaluation (:458-460)
    launch_me_ssd(st, n, h, w, tile, P, dr0, dr0, 1, 1, nullptr, nullptr, se);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_me_select, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n, nt, dr0, dr0, 1, 1,
                       (const float*)se, (const float*)nullptr, (const float*)nullptr, jy, jx, best0, cost0);
    LAUNCH_CHECK(ctx);
    // (7) parabolic half-pel, prefer-zero margin, vector sum (:443-485)
    hipLaunchKernelGGL(k_me_finish, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n * nt, (const float*)cm[0],
                       (const float*)cm[1], (const float*)cm[2], (const float*)cm[3], (const float*)best0,
                       (const float*)cost0, mdy, mdx, best, ody, odx);
    LAUNCH_CHECK(ctx);
    // (8) confidence (:321-387, :475)
    launch_me_confidence(st, n, h, w, tile, (const float*)P, (const float*)best, (float* const*)(base + oC),
                         scratch);
    LAUNCH_CHECK(ctx);
    for (int i = 0; i < n; ++i) {
      HIPCHK(ctx, hipMemcpyAsync(dy[i], ody + (size_t)i * nt, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
      HIPCHK(ctx, hipMemcpyAsync(dx[i], odx + (size_t)i * nt, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    }
    arm(ctx, sl, st);
  });
}

namespace {

// The device tables of the trajectory entry points: m pointers each of dy,
// dx, conf, then m shifts and m offsets.
hvdk::TrajTab trajectory_table(hvd_gpu_ctx* ctx, int m, const int* offsets, const float* shifts,
                               float* const* dy, float* const* dx, const float* const* conf) {
  REQUIRE(ctx, m >= 0 && m <= kMaxNeighbors, "%d measurements (at most %d)", m, kMaxNeighbors);
  REQUIRE(ctx, m == 0 || (offsets && shifts && dy && dx && conf), "null argument");
  hvdk::TrajTab T{};
  for (int j = 0; j < m; ++j) {
    REQUIRE(ctx, dy[j] && dx[j] && conf[j], "measurement %d: null plane", j);
    T.dy[j] = dy[j];
    T.dx[j] = dx[j];
    T.conf[j] = conf[j];
    T.ho[j] = shifts[j];
    T.off[j] = offsets[j];
  }
  return T;
}

}  // namespace

int hvd_gpu_trajectory_fit(hvd_gpu_ctx* ctx, int m, const int* offsets, const float* parity_shifts,
                           const float* const* dy, const float* const* dx, const float* const* conf, int tiles_h,
                           int tiles_w, float* vy, float* vx, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, vy && vx && tiles_h >= 1 && tiles_w >= 1, "bad output grid");
    hipStream_t st = S(stream);
    const int nt = tiles_h * tiles_w;
    const hvdk::TrajTab T =
        trajectory_table(ctx, m, offsets, parity_shifts, (float* const*)dy, (float* const*)dx, conf);
    hipLaunchKernelGGL(k_traj_fit, dim3((nt + 63) / 64), dim3(64), 0, st, m, nt, T, vy, vx);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_trajectory_snap(hvd_gpu_ctx* ctx, int m, const int* offsets, const float* parity_shifts, float* const* dy,
                            float* const* dx, const float* const* conf, int tiles_h, int tiles_w, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, tiles_h >= 1 && tiles_w >= 1, "bad tile grid");
    if (m < 2) return;
    hipStream_t st = S(stream);
    const int nt = tiles_h * tiles_w;
    const hvdk::TrajTab T = trajectory_table(ctx, m, offsets, parity_shifts, dy, dx, conf);
    hipLaunchKernelGGL(k_traj_snap, dim3((nt + 63) / 64), dim3(64), 0, st, m, nt, T);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_drizzle_frame(hvd_gpu_ctx* ctx, int j0, int n, const float* const* ys, const float* const* chis,
                          const int* parities, int lines, int width, int tile, int search, int nr_radius,
                          float nr_eps, int scale, float* out_luma, float* out_chroma, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, !ctx->no_alloc, "internal: drizzle inside a no-allocation call");
    const int rc = hvd_window::drizzle(ctx, j0, n, ys, chis, parities, lines, width, tile, search, nr_radius, nr_eps,
                                       scale, out_luma, out_chroma, S(stream));
    if (rc != HVD_GPU_OK) throw Fail{rc};
  });
}

int hvd_gpu_decode_window(hvd_gpu_ctx* ctx, const hvd_gpu_plane* fields, int n, const int* parities, int is_pal,
                          const int* y_given, const hvd_solver_params* p, const hvd_window_params* wp,
                          hvd_window_diag* diag, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, !ctx->no_alloc, "internal: window inside a no-allocation call");
    if (p) check_params(ctx, p);
    if (diag) *diag = hvd_window_diag{};
    const bool selective = p && wp && wp->selective_3d && p->cg_iterations > 0 && wp->temporal_strength >= 0.0;
    const int rc = selective
                       ? hvd_window::drive_selective(ctx, fields, n, parities, is_pal, y_given, p, wp, diag, S(stream))
                       : hvd_window::drive(ctx, fields, n, parities, is_pal, y_given, p, wp, diag, S(stream));
    if (rc != HVD_GPU_OK) throw Fail{rc};
  });
}

int hvd_gpu_ambiguous_boxes(hvd_gpu_ctx* ctx, const float* y, const float* chi, int lines, int width,
                            float max_area, int* boxes, int cap, int* nboxes, float* scores, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, y && chi && nboxes && (cap == 0 || boxes), "null argument");
    REQUIRE(ctx, lines >= 1 && width >= 1, "empty plane");
    REQUIRE(ctx, !ctx->no_alloc, "internal: selective boxes inside a no-allocation call");
    hipStream_t st = S(stream);
    char* scr = hvd_window::sel_pool(ctx, hvd_window::sel_scratch_bytes(lines, width), st);
    if (!scr) throw Fail{HVD_GPU_ERR_NOMEM};
    std::vector<hvd_window::SelBox> bx;
    bool everywhere = false;
    std::vector<float> sc;
    const int rc = hvd_window::ambiguous_boxes(ctx, y, chi, lines, width, max_area, scr, st, &bx, &everywhere,
                                               scores ? &sc : nullptr);
    if (rc != HVD_GPU_OK) throw Fail{rc};
    *nboxes = everywhere ? -1 : (int)bx.size();
    for (int k = 0; k < (int)bx.size() && k < cap; ++k) {
      boxes[4 * k] = bx[k].fy0;
      boxes[4 * k + 1] = bx[k].fy1;
      boxes[4 * k + 2] = bx[k].x0;
      boxes[4 * k + 3] = bx[k].x1;
    }
    if (scores) std::copy(sc.begin(), sc.end(), scores);
  });
}

int hvd_gpu_verify_motion(hvd_gpu_ctx* ctx, const float* const* y_ref, const float* const* y_cur, int n, int h,
                          int w, const float* const* pdy, const float* const* pdx, int tile, float* const* dy,
                          float* const* dx, float* const* conf, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (y_ref && y_cur && pdy && pdx && dy && dx && conf)), "null argument");
    REQUIRE(ctx, h >= 1 && w >= 1, "empty plane");
    REQUIRE(ctx, tile >= 1 && tile <= 256, "tile %d outside 1..256", tile);
    if (n == 0) return;
    hipStream_t st = S(stream);
    const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw;
    const size_t np = (size_t)h * w;
    Carve cv;
    const size_t o_tmp = cv.take(sizeof(float) * n * 2 * np), o_P = cv.take(sizeof(float) * n * 2 * np);
    const size_t o_se = cv.take(sizeof(float) * n * nt);
    const size_t o_v = cv.take(sizeof(float) * n * nt * 11);
    char* ws = ensure_ws(ctx, cv.off, st);
    float* tmp = (float*)(ws + o_tmp);
    float* P = (float*)(ws + o_P);
    float* se = (float*)(ws + o_se);
    float* v = (float*)(ws + o_v);
    auto vec = [&](int k) { return v + (size_t)k * n * nt; };
    float *py = vec(0), *px = vec(1), *bp = vec(2), *bz = vec(3), *jy = vec(4), *jx = vec(5), *best = vec(6);
    float *ody = vec(7), *odx = vec(8);
    std::vector<MotionDesc> D(n);
    for (int i = 0; i < n; ++i) {
      REQUIRE(ctx, y_ref[i] && y_cur[i] && pdy[i] && pdx[i] && dy[i] && dx[i] && conf[i], "pair %d: null pointer", i);
      D[i] = MotionDesc{y_ref[i], y_cur[i]};
      HIPCHK(ctx, hipMemcpyAsync(py + (size_t)i * nt, pdy[i], sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
      HIPCHK(ctx, hipMemcpyAsync(px + (size_t)i * nt, pdx[i], sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    }
    std::vector<float*> cptr(conf, conf + n);
    const std::vector<int> r0 = {0};
    Stage stg;
    const size_t oD = stg.put(D), oR0 = stg.put(r0), oC = stg.put(cptr);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    const int* dr0 = (const int*)(base + oR0);
    // motion.cpp:490-525: both planes blurred r = 2; the rounded prediction
    // (k_me_ssd rounds a relative base) and the zero vector, one SSD each
    launch_me_blur(st, (const MotionDesc*)(base + oD), n, h, w, 2, tmp, P);
    LAUNCH_CHECK(ctx);
    launch_me_ssd(st, n, h, w, tile, P, dr0, dr0, 1, 1, py, px, se);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_me_select, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n, nt, dr0, dr0, 1, 1,
                       (const float*)se, (const float*)py, (const float*)px, jy, jx, bp, (float*)nullptr);
    LAUNCH_CHECK(ctx);
    launch_me_ssd(st, n, h, w, tile, P, dr0, dr0, 1, 1, nullptr, nullptr, se);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_me_select, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n, nt, dr0, dr0, 1, 1,
                       (const float*)se, (const float*)nullptr, (const float*)nullptr, jy, jx, bz, (float*)nullptr);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_verify_pick, dim3(blocks_for((size_t)n * nt)), dim3(256), 0, st, n * nt, (const float*)py,
                       (const float*)px, (const float*)bp, (const float*)bz, ody, odx, best);
    LAUNCH_CHECK(ctx);
    launch_me_confidence(st, n, h, w, tile, (const float*)P, (const float*)best, (float* const*)(base + oC),
                         vec(9) /* scratch: 2 n nt floats from vec(9) */);
    LAUNCH_CHECK(ctx);
    for (int i = 0; i < n; ++i) {
      HIPCHK(ctx, hipMemcpyAsync(dy[i], ody + (size_t)i * nt, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
      HIPCHK(ctx, hipMemcpyAsync(dx[i], odx + (size_t)i * nt, sizeof(float) * nt, hipMemcpyDeviceToDevice, st));
    }
    arm(ctx, sl, st);
  });
}

// ---- frame-bridge and component entry points of the C++ host API ----------

int hvd_gpu_split_frame_fields(hvd_gpu_ctx* ctx, const int16_t* frame, const hvd_frame_params* fp, float zero_code,
                               float* field1, float* field2, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, frame && fp && field1 && field2, "null argument");
    check_frame(ctx, *fp);
    hipStream_t st = S(stream);
    const BridgeFrameParams bp = bridge_params(*fp);
    const int16_t* fr[1] = {frame};
    float* outs[2] = {field1, field2};
    Stage stg;
    const size_t oP = stg.put(&bp, 1), oF = stg.put(fr, 1), oO = stg.put(outs, 2);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    hipLaunchKernelGGL(k_split_fields, dim3((fp->frame_width + 255) / 256, fp->field1_lines, 1), dim3(256), 0, st,
                       (const BridgeFrameParams*)(base + oP), (const int16_t* const*)(base + oF), zero_code,
                       (float* const*)(base + oO));
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, st);
  });
}

int hvd_gpu_field_parity_vote(hvd_gpu_ctx* ctx, const float* block1, const float* block2, const hvd_frame_params* fp,
                              int* vote, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, block1 && block2 && fp && vote, "null argument");
    check_frame(ctx, *fp);
    hipStream_t st = S(stream);
    Carve cv;
    const size_t o = cv.take(sizeof(int) * 4);
    char* ws = ensure_ws(ctx, cv.off, st);
    int* dec = (int*)(ws + o);
    const BridgeFrameParams bp = bridge_params(*fp);
    const float* fl[2] = {block1, block2};
    Stage stg;
    const size_t oP = stg.put(&bp, 1), oF = stg.put(fl, 2);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    hipLaunchKernelGGL(k_window_decisions, dim3(1), dim3(256), 0, st, (const BridgeFrameParams*)(base + oP),
                       (const float* const*)(base + oF), 1, (const float*)nullptr, 0, 1.0f, 20.0f, 0, dec,
                       (float*)(dec + 2), dec + 1);
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, st);
    int h[2] = {0, 0};
    HIPCHK(ctx, hipMemcpyAsync(h, dec, sizeof(h), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    *vote = h[1];
  });
}

int hvd_gpu_make_carrier(hvd_gpu_ctx* ctx, const float* theta, const int8_t* vswitch, int lines,
                         const hvd_field_geometry* g, float* carrier, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, g && (lines == 0 || (theta && carrier)), "null argument");
    const int width = g->active_video_end - g->active_video_start;
    REQUIRE(ctx, lines >= 0 && width >= 0, "bad carrier size %d x %d", lines, width);
    if (lines == 0 || width == 0) return;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_make_carrier, dim3((width + 255) / 256, lines), dim3(256), 0, st, theta, vswitch, width,
                       g->active_video_start, (float)phase_per_sample(*g), (float2*)carrier);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_tridiag_smooth(hvd_gpu_ctx* ctx, const float* d, const float* a, int n, float lam, float* out,
                           void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, n >= 0 && (n == 0 || (d && a && out)), "null argument");
    if (n == 0) return;
    hipStream_t st = S(stream);
    Carve cv;
    const size_t o = cv.take(sizeof(double) * 5 * (size_t)n);
    char* ws = ensure_ws(ctx, cv.off, st);
    hipLaunchKernelGGL(k_tridiag, dim3(1), dim3(64), 0, st, d, a, n, lam, (double*)(ws + o), out);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_difference(hvd_gpu_ctx* ctx, int op, int is_complex, const float* a, const float* weight, int h, int w,
                       float* d, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, op >= HVD_DIFF_DX && op <= HVD_DIFF_D2T, "unknown difference operator %d", op);
    REQUIRE(ctx, h >= 0 && w >= 0, "bad plane %d x %d", h, w);
    if ((size_t)h * w == 0) return;
    REQUIRE(ctx, a && d && a != d, "null or aliased plane");
    REQUIRE(ctx, !weight || op == HVD_DIFF_DX || op == HVD_DIFF_DY || op == HVD_DIFF_D1 || op == HVD_DIFF_D2,
            "a weight applies to Dx, Dy, D1 and D2 only");
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_difference, dim3(blocks_for((size_t)h * w)), dim3(256), 0, st, op, is_complex ? 2 : 1, a,
                       weight, h, w, d);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_estimate_noise_ire(hvd_gpu_ctx* ctx, const float* s, int h, int w, float* sigma, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, s && sigma, "null argument");
    REQUIRE(ctx, h >= 0 && w >= 0, "bad plane %d x %d", h, w);
    if (w <= 8 || h <= 0) {
      *sigma = 0.0f;
      return;
    }
    hipStream_t st = S(stream);
    Carve cv;
    const size_t o_d = cv.take(sizeof(float) * (size_t)h * (w - 8)), o_sig = cv.take(sizeof(float));
    char* ws = ensure_ws(ctx, cv.off, st);
    const float* tab[1] = {s};
    Stage stg;
    const size_t oT = stg.put(tab, 1);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    hipLaunchKernelGGL(k_noise_sigma, dim3(1), dim3(1024), 0, st, (const float* const*)(base + oT), h, w,
                       (float*)(ws + o_d), (float*)(ws + o_sig));
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, st);
    HIPCHK(ctx, hipMemcpyAsync(sigma, ws + o_sig, sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
  });
}

int hvd_gpu_fft2d(hvd_gpu_ctx* ctx, const float* in, int h, int w, int inverse, float* out, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, in && out, "null argument");
    REQUIRE(ctx, h >= 1 && w >= 1 && h <= 1024 && w <= 4096, "FFT plane %d x %d outside 1..1024 x 1..4096", h, w);
    hipStream_t st = S(stream);
    PlanSet ps;
    const int pr = ps.get(ctx, w, 1), pc = ps.get(ctx, h, kColGroup);
    Carve cv;
    const size_t o_roots = cv.take(sizeof(float2) * ps.roots.size());
    char* ws = ensure_ws(ctx, cv.off, st);
    float2* roots = (float2*)(ws + o_roots);
    HIPCHK(ctx, hipMemcpyAsync(roots, ps.roots.data(), sizeof(float2) * ps.roots.size(), hipMemcpyHostToDevice, st));
    const size_t lds_r = sizeof(float2) * (size_t)fft_lds_slots(ps.plans[pr], 1),
                 lds_c = sizeof(float2) * (size_t)fft_lds_slots(ps.plans[pc], kColGroup);
    REQUIRE(ctx, (int)std::max(lds_r, lds_c) <= ctx->max_dyn_lds, "FFT plane %d x %d needs more LDS than a block has",
            h, w);
    const int sign = inverse ? 1 : -1;
    hipLaunchKernelGGL(k_fft2_rows, dim3(h), dim3(256), lds_r, st, (const float2*)in, (float2*)out, h, w,
                       ps.plans[pr], (const float2*)roots, sign);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_fft2_cols, dim3((w + kColGroup - 1) / kColGroup), dim3(256), lds_c, st, (float2*)out, h, w,
                       ps.plans[pc], (const float2*)roots, sign, inverse ? (float)(1.0 / ((double)h * w)) : 1.0f);
    LAUNCH_CHECK(ctx);
    // the roots are read by the launches above before a later call reuses ws
    HIPCHK(ctx, hipStreamSynchronize(st));
  });
}

int hvd_gpu_box_blur(hvd_gpu_ctx* ctx, const float* in, int h, int w, int r, float* out, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, in && out, "null argument");
    REQUIRE(ctx, h >= 1 && w >= 1 && r >= 0 && r <= kBlurMaxR, "bad blur %d x %d, radius %d (at most %d)", h, w, r,
            kBlurMaxR);
    hipStream_t st = S(stream);
    const size_t np = (size_t)h * w;
    Carve cv;
    const size_t o_tmp = cv.take(sizeof(float) * np);
    char* ws = ensure_ws(ctx, cv.off, st);
    const MotionDesc md{in, in};
    Stage stg;
    const size_t oD = stg.put(&md, 1);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    // the column pass reads `in` whole before the row pass writes `out`, so
    // in place is fine
    launch_me_blur_planes(st, (const MotionDesc*)(base + oD), 1, h, w, r, (float*)(ws + o_tmp), out);
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, st);
  });
}

}  // extern "C"

namespace {

void check_tiles(hvd_gpu_ctx* ctx, const void* mdy, const void* mdx, int th, int tw, int tile) {
  REQUIRE(ctx, mdy && mdx, "null motion field");
  REQUIRE(ctx, th >= 1 && tw >= 1 && tile >= 1, "bad motion grid %d x %d, tile %d", th, tw, tile);
}

// VectorsPerPixel (temporal.cpp:44-113) into dy, dx (dx may be null), the
// snapped tile vectors in workspace bytes [o_tiles, o_tiles + 8 th tw).
void vectors_per_pixel(hvd_gpu_ctx* ctx, const float* mdy, const float* mdx, int th, int tw, int tile, int out_h,
                       int out_w, float* dy, float* dx, float* sy, float* sx, hipStream_t st) {
  hipLaunchKernelGGL(k_vpp_snap, dim3(blocks_for((size_t)th * tw)), dim3(256), 0, st, th, tw, mdy, mdx, sy, sx);
  LAUNCH_CHECK(ctx);
  hipLaunchKernelGGL(k_vpp_interp, dim3(blocks_for((size_t)out_h * out_w)), dim3(256), 0, st, th, tw, tile, out_h,
                     out_w, (const float*)sy, (const float*)sx, dy, dx);
  LAUNCH_CHECK(ctx);
}

// Box blur of npairs * 2 planes laid out plane after plane at `in` (radius
// r), result at `out`; `tmp` holds as many planes.
void blur_planes(hvd_gpu_ctx* ctx, const std::vector<MotionDesc>& pairs, int h, int w, int r, float* tmp, float* out,
                 hipStream_t st) {
  Stage stg;
  const size_t oD = stg.put(pairs);
  Slot* sl;
  char* base = upload(ctx, stg, st, &sl);
  const int np = (int)pairs.size();
  launch_me_blur(st, (const MotionDesc*)(base + oD), np, h, w, r, tmp, out);
  LAUNCH_CHECK(ctx);
  arm(ctx, sl, st);
}

// UpsampleConfidence[Fast] (temporal.cpp:217-249) into out; scratch from ws
// at `o` (4 planes of out_h x out_w and 2 tile planes).
void upsample_confidence(hvd_gpu_ctx* ctx, const float* conf, int th, int tw, int tile, int out_h, int out_w,
                         bool fast, float* out, char* ws, Carve& cv, hipStream_t st) {
  const size_t np = (size_t)out_h * out_w, nt = (size_t)th * tw;
  if (fast) {
    float* c2 = (float*)(ws + cv.take(sizeof(float) * nt));
    float* sy = (float*)(ws + cv.take(sizeof(float) * nt));
    float* sx = (float*)(ws + cv.take(sizeof(float) * nt));
    hipLaunchKernelGGL(k_square, dim3(blocks_for(nt)), dim3(256), 0, st, (int)nt, conf, c2);
    LAUNCH_CHECK(ctx);
    vectors_per_pixel(ctx, c2, c2, th, tw, tile, out_h, out_w, out, nullptr, sy, sx, st);
    return;
  }
  float* px = (float*)(ws + cv.take(sizeof(float) * np));
  float* tmp = (float*)(ws + cv.take(sizeof(float) * np));
  hipLaunchKernelGGL(k_conf_nearest, dim3(blocks_for(np)), dim3(256), 0, st, th, tw, tile, out_h, out_w, conf, px);
  LAUNCH_CHECK(ctx);
  Stage stg;
  const MotionDesc md{px, px};
  const size_t oD = stg.put(&md, 1);
  Slot* sl;
  char* base = upload(ctx, stg, st, &sl);
  launch_me_blur_planes(st, (const MotionDesc*)(base + oD), 1, out_h, out_w, 8, tmp, out);  // one plane, into out
  LAUNCH_CHECK(ctx);
  arm(ctx, sl, st);
}

}  // namespace

extern "C" {

int hvd_gpu_vectors_per_pixel(hvd_gpu_ctx* ctx, const float* mdy, const float* mdx, int tiles_h, int tiles_w, int tile,
                              int out_h, int out_w, float* dy, float* dx, void* stream) {
  return guarded(ctx, [&] {
    check_tiles(ctx, mdy, mdx, tiles_h, tiles_w, tile);
    REQUIRE(ctx, dy && dx && out_h >= 1 && out_w >= 1, "bad output");
    hipStream_t st = S(stream);
    Carve cv;
    const size_t nt = (size_t)tiles_h * tiles_w;
    const size_t o_y = cv.take(sizeof(float) * nt), o_x = cv.take(sizeof(float) * nt);
    char* ws = ensure_ws(ctx, cv.off, st);
    vectors_per_pixel(ctx, mdy, mdx, tiles_h, tiles_w, tile, out_h, out_w, dy, dx, (float*)(ws + o_y),
                      (float*)(ws + o_x), st);
  });
}

int hvd_gpu_warp_by_tiles(hvd_gpu_ctx* ctx, const float* a, int h, int w, int channels, const float* dy,
                          const float* dx, float* out, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, a && dy && dx && out, "null argument");
    REQUIRE(ctx, h >= 1 && w >= 1 && (channels == 1 || channels == 2), "bad warp %d x %d x %d", h, w, channels);
    REQUIRE(ctx, out != a, "the warp cannot run in place");
    const size_t n = (size_t)h * w;
    hipLaunchKernelGGL(k_warp_tiles, dim3(blocks_for(n)), dim3(256), 0, S(stream), h, w, channels, a, dy, dx, out);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_warp_bilinear(hvd_gpu_ctx* ctx, const float* a, int src_h, int src_w, const float* dy, const float* dx,
                          float row_offset, int out_h, int out_w, float* out, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, a && dy && dx && out && out != a, "null or aliased argument");
    REQUIRE(ctx, src_h >= 1 && src_w >= 1 && out_h >= 1 && out_w >= 1, "empty plane");
    const size_t n = (size_t)out_h * out_w;
    hipLaunchKernelGGL(k_warp_bilinear, dim3(blocks_for(n)), dim3(256), 0, S(stream), src_h, src_w, a, dy, dx,
                       row_offset, out_h, out_w, out);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_upsample_confidence(hvd_gpu_ctx* ctx, const float* conf, int tiles_h, int tiles_w, int tile, int out_h,
                                int out_w, int fast, float* out, void* stream) {
  return guarded(ctx, [&] {
    check_tiles(ctx, conf, conf, tiles_h, tiles_w, tile);
    REQUIRE(ctx, out && out_h >= 1 && out_w >= 1, "bad output");
    hipStream_t st = S(stream);
    const size_t np = (size_t)out_h * out_w, nt = (size_t)tiles_h * tiles_w;
    char* ws = ensure_ws(ctx, 6 * align256(sizeof(float) * 2 * np) + 3 * align256(sizeof(float) * nt), st);
    Carve cv;
    upsample_confidence(ctx, conf, tiles_h, tiles_w, tile, out_h, out_w, fast != 0, out, ws, cv, st);
  });
}

int hvd_gpu_complex_coherence(hvd_gpu_ctx* ctx, const float* z1, const float* z2, int h, int w, int r, float* out,
                              void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, z1 && z2 && out, "null argument");
    REQUIRE(ctx, h >= 1 && w >= 1 && r >= 0 && r <= kBlurMaxR, "bad coherence %d x %d, radius %d (at most %d)", h, w,
            r, kBlurMaxR);
    hipStream_t st = S(stream);
    const size_t n = (size_t)h * w;
    Carve cv;
    const size_t o_in = cv.take(sizeof(float) * 4 * n), o_tmp = cv.take(sizeof(float) * 4 * n),
                 o_bl = cv.take(sizeof(float) * 4 * n);
    char* ws = ensure_ws(ctx, cv.off, st);
    float* in = (float*)(ws + o_in);
    float* bl = (float*)(ws + o_bl);
    hipLaunchKernelGGL(k_coh_terms, dim3(blocks_for(n)), dim3(256), 0, st, n, (const float2*)z1, (const float2*)z2,
                       in, in + n, in + 2 * n, in + 3 * n);
    LAUNCH_CHECK(ctx);
    blur_planes(ctx, {MotionDesc{in, in + n}, MotionDesc{in + 2 * n, in + 3 * n}}, h, w, r, (float*)(ws + o_tmp), bl,
                st);
    hipLaunchKernelGGL(k_coh_out, dim3(blocks_for(n)), dim3(256), 0, st, n, (const float*)bl,
                       (const float*)(bl + n), (const float*)(bl + 2 * n), (const float*)(bl + 3 * n), out);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_motion_compensate_prev(hvd_gpu_ctx* ctx, const float* composite, const float* carrier, int h, int w,
                                   const float* mdy, const float* mdx, const float* mconf, int tiles_h, int tiles_w,
                                   int tile, int fast, float* out_composite, float* out_carrier, float* out_conf,
                                   void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, composite && carrier && mconf && out_composite && out_carrier && out_conf, "null argument");
    REQUIRE(ctx, out_composite != composite && out_carrier != carrier, "the warp cannot run in place");
    REQUIRE(ctx, h >= 1 && w >= 1, "empty plane");
    check_tiles(ctx, mdy, mdx, tiles_h, tiles_w, tile);
    hipStream_t st = S(stream);
    const size_t n = (size_t)h * w, nt = (size_t)tiles_h * tiles_w;
    Carve cv;
    const size_t o_dy = cv.take(sizeof(float) * n), o_dx = cv.take(sizeof(float) * n),
                 o_sy = cv.take(sizeof(float) * nt), o_sx = cv.take(sizeof(float) * nt);
    char* ws = ensure_ws(ctx, cv.off + 6 * align256(sizeof(float) * 2 * n) + 3 * align256(sizeof(float) * nt), st);
    float* dy = (float*)(ws + o_dy);
    float* dx = (float*)(ws + o_dx);
    // one per-pixel interpolation shared by both warps (temporal.cpp:266-275)
    vectors_per_pixel(ctx, mdy, mdx, tiles_h, tiles_w, tile, h, w, dy, dx, (float*)(ws + o_sy), (float*)(ws + o_sx),
                      st);
    hipLaunchKernelGGL(k_warp_tiles, dim3(blocks_for(n)), dim3(256), 0, st, h, w, 1, composite, (const float*)dy,
                       (const float*)dx, out_composite);
    LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_warp_tiles, dim3(blocks_for(n)), dim3(256), 0, st, h, w, 2, carrier, (const float*)dy,
                       (const float*)dx, out_carrier);
    LAUNCH_CHECK(ctx);
    upsample_confidence(ctx, mconf, tiles_h, tiles_w, tile, h, w, fast != 0, out_conf, ws, cv, st);
  });
}

int hvd_gpu_chroma_plane(hvd_gpu_ctx* ctx, int mode, const float* s, const float* y, const float* chi,
                         const float* carrier, size_t n, float* out, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, mode == 0 || mode == 1, "bad mode %d", mode);
    REQUIRE(ctx, out && (mode == 0 ? (s && y) : (chi && carrier)), "null argument");
    if (n == 0) return;
    hipStream_t st = S(stream);
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_chroma_plane, dim3(blocks), dim3(256), 0, st, mode, s, y, (const float2*)chi,
                       (const float2*)carrier, n, out);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_weave_frame_s16(hvd_gpu_ctx* ctx, const int16_t* src, const hvd_frame_params* fp, int rows, int width,
                            const float* y, const float* chroma, const float* chi, float gain, const hvd_yc_frame* out,
                            void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, src && fp && out && out->luma && out->chroma && out->u_plane && out->v_plane, "null argument");
    REQUIRE(ctx, rows >= 0 && width >= 0 && (rows * width == 0 || (chroma && chi)), "bad woven plane");
    check_frame(ctx, *fp);
    hipStream_t st = S(stream);
    const BridgeFrameParams bp = bridge_params(*fp);
    Stage stg;
    const size_t oP = stg.put(&bp, 1);
    Slot* sl;
    char* base = upload(ctx, stg, st, &sl);
    const BridgeOut o{out->luma, out->chroma, out->u_plane, out->v_plane};
    hipLaunchKernelGGL(k_weave_woven, dim3((fp->frame_width + 255) / 256, fp->frame_height), dim3(256), 0, st,
                       (const BridgeFrameParams*)(base + oP), src, rows, width, y, chroma, (const float2*)chi, gain,
                       o);
    LAUNCH_CHECK(ctx);
    arm(ctx, sl, st);
  });
}

}  // extern "C"

namespace {

// The woven picture's extent (hvd_chroma_decoder_stage.cpp:328-363 raster,
// :336-353 single field).
WovenDesc woven_geometry(const hvd_frame_params& fp, int full_raster, int field) {
  WovenDesc D{};
  const int fw = fp.frame_width, fh = fp.frame_height, f1 = fp.field1_lines;
  const int a0 = full_raster ? 0 : std::max(0, fp.active_video_start);
  const int a1 = full_raster ? fw : (fp.active_video_end > a0 ? fp.active_video_end : fw);
  const int y0 = full_raster ? 0 : std::max(0, fp.first_active_frame_line);
  const int y1 = full_raster ? fh : (fp.last_active_frame_line > y0 ? fp.last_active_frame_line : fh);
  D.a0 = a0;
  D.width = std::max(0, a1 - a0);
  D.y0 = y0;
  D.field = (field == 0 || field == 1) ? field : -1;
  D.f1 = f1;
  D.fw = fw;
  D.fh = fh;
  if (D.field >= 0) {
    // field lines fl with y0 <= 2 fl + field < y1, fl < that field's line count
    const int nfl = D.field == 0 ? f1 : fh - f1;
    int lo = 0;
    while (2 * lo + D.field < y0) ++lo;
    int hi = lo;
    while (hi < nfl && 2 * hi + D.field < y1) ++hi;
    D.fl0 = lo;
    D.height = std::max(0, hi - lo);
  } else {
    D.height = std::max(0, y1 - y0);
  }
  if (D.width == 0 || D.height == 0) D.width = D.height = 0;
  D.black = (double)fp.black_level;
  D.range = (double)(fp.white_level - fp.black_level);  // a float difference, as at :554
  return D;
}

}  // namespace

extern "C" {

int hvd_gpu_woven_size(const hvd_frame_params* fp, int full_raster, int field, int* width, int* height) {
  if (!fp || !width || !height || fp->frame_width <= 0 || fp->frame_height <= 0) return HVD_GPU_ERR_INVALID;
  const WovenDesc D = woven_geometry(*fp, full_raster, field);
  *width = D.width;
  *height = D.height;
  return HVD_GPU_OK;
}

int hvd_gpu_woven_picture(hvd_gpu_ctx* ctx, const hvd_yc_frame* yc, const hvd_frame_params* fp, int full_raster,
                          int field, double* y, double* u, double* v, uint8_t* rgb24, void* stream) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, yc && fp && yc->luma && yc->u_plane && yc->v_plane, "null argument");
    REQUIRE(ctx, fp->frame_width > 0 && fp->frame_height > 0, "empty frame");
    REQUIRE(ctx, fp->field1_lines >= 0 && fp->field1_lines <= fp->frame_height, "field1_lines outside the frame");
    if (rgb24)
      REQUIRE(ctx, fp->white_level > fp->black_level && fp->white_level > fp->blanking_level,
              "white level must exceed black and blanking for the RGB rendering");
    const WovenDesc D = woven_geometry(*fp, full_raster, field);
    if (D.width == 0 || D.height == 0 || !(y || u || v || rgb24)) return;
    hipLaunchKernelGGL(k_woven_rgb, dim3((D.width + 255) / 256, D.height), dim3(256), 0, S(stream), D, yc->luma,
                       yc->u_plane, yc->v_plane, y, u, v, rgb24);
    LAUNCH_CHECK(ctx);
  });
}

int hvd_gpu_set_graphs(hvd_gpu_ctx* ctx, int enable) {
  return guarded(ctx, [&] {
    HIPCHK(ctx, hipDeviceSynchronize());
    if (!enable) drop_graphs(ctx);
    ctx->graphs = enable != 0;
  });
}

int hvd_gpu_set_solver_tuning(hvd_gpu_ctx* ctx, const hvd_gpu_solver_tuning* t) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, t != nullptr, "null tuning");
    REQUIRE(ctx, t->strip_rows == 0 || (t->strip_rows >= kMinStripRows && t->strip_rows <= kMaxStripRows),
            "strip_rows %d outside %d..%d", t->strip_rows, kMinStripRows, kMaxStripRows);
    REQUIRE(ctx, t->group_pixels >= 0, "negative group_pixels");
    ctx->fixed.rows = t->strip_rows;
    ctx->fixed.group_px = (size_t)t->group_pixels;
  });
}

int hvd_gpu_autotune_solver(hvd_gpu_ctx* ctx, int lines, int width, int is_pal, int nfields, int iterations,
                            hvd_gpu_solver_tuning* best, double* ns_per_pixel_iteration) {
  return guarded(ctx, [&] {
    REQUIRE(ctx, best != nullptr, "null output");
    REQUIRE(ctx, lines >= 2 && width >= 2 && lines + 32 <= 1024 && width <= 2048, "bad field shape %d x %d", lines,
            width);
    REQUIRE(ctx, nfields >= 1 && nfields <= 4096, "nfields %d outside 1..4096", nfields);
    REQUIRE(ctx, (double)lines * width * nfields <= 2.0e8, "autotune batch too large");
    const TuneResult r = autotune(ctx, lines, width, is_pal != 0, nfields, std::max(2, std::min(iterations, 64)));
    best->strip_rows = r.t.rows;
    best->group_pixels = (long long)r.t.group_px;
    if (ns_per_pixel_iteration) *ns_per_pixel_iteration = r.ns;
  });
}

}  // extern "C"

// ---- services of the window driver (window.h) --------------------------------
namespace hvd_window {

char* pool(hvd_gpu_ctx* ctx, size_t bytes, hipStream_t st) {
  if (bytes <= ctx->win_cap) return ctx->win;
  if (hipStreamSynchronize(st) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    set_err(ctx, "synchronisation before the window allocation failed");
    return nullptr;
  }
  if (ctx->win) (void)hipFree(ctx->win);
  ctx->win = nullptr;
  ctx->win_cap = 0;
  const size_t want = bytes + bytes / 4;
  void* q = nullptr;
  if (hipMalloc(&q, want) != hipSuccess) {
    set_err(ctx, "hipMalloc(%zu) for the window failed", want);
    return nullptr;
  }
  ctx->win = static_cast<char*>(q);
  ctx->win_cap = want;
  return ctx->win;
}

char* sel_pool(hvd_gpu_ctx* ctx, size_t bytes, hipStream_t st) {
  if (bytes <= ctx->sel_cap) return ctx->sel;
  if (hipStreamSynchronize(st) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    set_err(ctx, "synchronisation before the selective window's allocation failed");
    return nullptr;
  }
  if (ctx->sel) (void)hipFree(ctx->sel);
  ctx->sel = nullptr;
  ctx->sel_cap = 0;
  void* q = nullptr;
  if (hipMalloc(&q, bytes) != hipSuccess) {
    set_err(ctx, "hipMalloc(%zu) for the selective window failed", bytes);
    return nullptr;
  }
  ctx->sel = static_cast<char*>(q);
  ctx->sel_cap = bytes;
  return ctx->sel;
}

int fail(hvd_gpu_ctx* ctx, int code, const char* msg) {
  set_err(ctx, "%s", msg);
  return code;
}

}  // namespace hvd_window
