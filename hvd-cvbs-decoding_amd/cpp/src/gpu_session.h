// gpu_session.h -- device plumbing of the C++ API over the C ABI: a lazily
// created GPU context, RAII device buffers, the hvd-core <-> ABI struct
// conversions and the failure path (empty outputs + a recorded message).
#ifndef HVD_AMD_SRC_GPU_SESSION_H_
#define HVD_AMD_SRC_GPU_SESSION_H_

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstddef>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine/hvd_config.h"
#include "engine/ntsc_geometry.h"
#include "hvd_gpu.h"

namespace hvd {
struct FrameParams;

namespace gpu {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// One GPU context, created on the calling thread's current HIP device at
// first use.  Calls through one session are serialised on the null stream.
class Session {
 public:
  Session() = default;
  ~Session();
  Session(const Session&) = delete;
  Session& operator=(const Session&) = delete;

  hvd_gpu_ctx* ctx();
  // Throws Error carrying the context's message when status != HVD_GPU_OK.
  void check(int status, const char* what);
  void sync();
  // With HVD_AUTOTUNE set to a non-zero value in the environment: the first
  // time this session solves `nfields` fields of lines x width, it runs the
  // library's autotuner for that batch first (hvd_gpu_autotune_solver).
  // Failures of the tuner are reported and otherwise ignored.
  void autotune(int lines, int width, bool pal, int nfields);
  // Page-locked host memory for staged transfers (hipHostMalloc), at least
  // `bytes`, owned by the session and grown only when a call needs more (the
  // previous contents are not kept).  Callers finish their copies through it
  // before they return.
  char* pinned(size_t bytes);

 private:
  hvd_gpu_ctx* ctx_ = nullptr;
  std::set<std::array<int, 4>> tuned_;
  char* pinned_ = nullptr;
  size_t pinned_cap_ = 0;
};

// The session free functions (BurstLockinPhase, HolographicInit, ...) use.
Session& ThreadSession();

void HipCheck(hipError_t e, const char* what);

// Records a failure for GpuLastError() and reports it on stderr.
void Report(const char* api, const std::exception& e);

// Device memory of the C++ API's buffers, cached per device: a released
// block goes back to a size-classed free list (sizes rounded up to 1/8 of
// their power of two) and the next request of that class on that device
// takes it, so a steady stream of calls of one shape allocates nothing after
// the first.  All the API's work runs on the device's null stream, which
// orders a block's reuse after the work that used it.  At most 8 GiB per
// device stays cached; beyond that a released block is freed.
void* CachedMalloc(size_t bytes, size_t* cap, int* device);
void CachedFree(void* p, size_t cap, int device);

template <typename T>
class DevBuf {
 public:
  explicit DevBuf(size_t n = 0) { resize(n); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_), dev_(o.dev_) {
    o.p_ = nullptr;
    o.n_ = 0;
    o.cap_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      cap_ = o.cap_;
      dev_ = o.dev_;
      o.p_ = nullptr;
      o.n_ = 0;
      o.cap_ = 0;
    }
    return *this;
  }

  void resize(size_t n) {
    release();
    if (n) p_ = static_cast<T*>(CachedMalloc(n * sizeof(T), &cap_, &dev_));
    n_ = n;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void upload(const T* h, size_t n) {
    if (n) HipCheck(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  void download(T* h, size_t n) const {
    if (n) HipCheck(hipMemcpy(h, p_, n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }

 private:
  void release() {
    if (p_) CachedFree(p_, cap_, dev_);
    p_ = nullptr;
    n_ = 0;
    cap_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
  size_t cap_ = 0;
  int dev_ = 0;
};

template <typename T>
DevBuf<T> Upload(const T* h, size_t n) {
  DevBuf<T> d(n);
  d.upload(h, n);
  return d;
}

hvd_field_geometry ToAbi(const FieldGeometry& g);
hvd_solver_params ToAbi(const HvdConfig& cfg, bool is_pal);
hvd_frame_params ToAbi(const FrameParams& fp);
hvd_bridge_params BridgeOf(const HvdConfig& cfg);
hvd_window_params WindowOf(const HvdConfig& cfg);

// Would hvd-core engage its 3-D temporal path for this window?
// (sequence.cpp:526-527, :652-661: strength > 0, or AUTO (0) with more
// fields than the same-parity stride, and a solve at all.)
bool WantsTemporal(const HvdConfig& cfg, int nfields, bool pal);

}  // namespace gpu
}  // namespace hvd

#endif  // HVD_AMD_SRC_GPU_SESSION_H_
