// hvd_device.h -- device-side types and helpers shared by the HIP kernels.
//
// Data layout in HBM (one active plane per field, row-major h x w):
//   s        float   composite (IRE)
//   chi      float2  chroma phasor chi = V - iU, interleaved (re, im)
//   d        float2  CG search direction (ping-ponged; the gradient is never
//                    stored: each pass recomputes it from chi)
//   wts      float4  IRLS diffusivities (wx, wy, wcx, wcy) per pixel
//   carrier  never stored: c(y, x) = rowph[y] * colph[x], with the PAL
//            V-switch folded in as c -> -conj(c) on switched lines
//            (ntsc_geometry.cpp:75-93).  An explicit plane is accepted too.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "strip_geom.h"

namespace hvdk {

constexpr int kTileW = 64;  // one wavefront per tile row (init kernels)
constexpr int kTileH = 4;
constexpr int kBlock = kTileW * kTileH;  // 256 threads
// Launch geometry of the solver kernels: strip prefix over the batch, and
// the strips per field when every field has the same size (else 0).
struct StripGrid {
  const int* tp;
  int nf, nb, rows, spf;
};

// FP64 partial sums per solver strip and CG launch (|g|^2, curv(g),
// curv(g, d), Re<g, d>).
constexpr int kCgPartials = 4;

// One field of a batch, as every per-pixel kernel sees it.
struct DevField {
  int h, w;            // active lines, active width
  int tiles_x;         // strips across: ceil(w / kStripCols)
  int tile0;           // first strip index of this field in the batch
  const float* s;      // h*w
  const float2* rowph; // h: e^{i theta_y}
  const int8_t* vsw;   // h or null (NTSC)
  const float2* colph; // w: e^{i w0 (a0 + x)}
  const float2* carrier;  // explicit h*w carrier, or null (factored form)
  float2* chi[2];      // ping-pong chroma (chi[0] is the caller's buffer)
  float2* d[2];        // ping-pong search direction
  float4* wts;         // IRLS weights
  float* y;            // output luma
  // adjoint solves: s == null means the data term's S is 0 (the gradient is
  // then the Hessian applied to chi) and rhs, when set, is subtracted from
  // the gradient, so the CG solves Hessian . chi = rhs
  const float2* rhs;
  float2* save;        // when set: chi at the start of each IRLS stage (stage-major planes)
  int pal;             // 1 + (PAL-family field: the AUTO aniso's 4-line grouping), or 0: the solve's setting
  int pad_;
};

// The scalar state of one field's CG loop after a launch (variational.cpp:
// 519-549): hh is curv(d) of the direction the next iteration moves along,
// gd = Re<g, d> of it, cur the ping-pong slot of chi, g and d(previous) that
// it reads.
struct CgDyn {
  double gg, gg0, hh, gd;
  float alpha, beta;
  int it, done, cur, pad_;
};

// Per-field solver state.  Each CG iteration launch first sums the previous
// launch's strip partials of its field (every wave forms the same sums in the
// same order) and advances the previous snapshot; the field's first strip
// publishes the new snapshot in slot it & 1, so a launch reads one slot while
// the next is written.
struct FieldCG {
  double npix;
  double tol2;
  double hh_floor;  // stop when curv(d) <= hh_floor (hvd-core: 1e-12)
  double gg_floor;  // stop when |g|^2 < gg_floor * npix (hvd-core: 1e-10)
  CgDyn dyn[2];
  float mu_h, mu_v, lam_y, eps, eps_c, k;
  float aniso_forced;   // > 0 forces chroma_aniso
  float diag;           // diag_prior: the oriented (+-45 deg) prior's share
  float mu_d;           // its weight after the renormalisation (variational.cpp:205-219)
  float aniso;          // the resolved chroma_aniso
  int is_pal;
  int n_inner;
  int cur;              // slot of chi (and g) at the start of the next stage / at the end
  int iters[16];        // CG iterations per IRLS stage (diagnostics)
  int stage;
};

// Plane accesses through global-address-space pointers: the pointers arrive
// as generic addresses in descriptor tables, and a generic (flat) access is
// counted by both the vector-memory and the LDS/scalar counters, so every wait
// on one becomes a wait on all.  Global accesses wait on the vector counter
// alone, in issue order.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const float g_cf;
typedef __attribute__((address_space(1))) const f2v g_cf2;
typedef __attribute__((address_space(1))) const f4v g_cf4;
typedef __attribute__((address_space(1))) const int8_t g_ci8;
typedef __attribute__((address_space(1))) float g_f;
typedef __attribute__((address_space(1))) f2v g_f2;
typedef __attribute__((address_space(1))) f4v g_f4;

__device__ inline float gld(const float* p, size_t i) { return ((g_cf*)p)[i]; }
__device__ inline int8_t gld(const int8_t* p, size_t i) { return ((g_ci8*)p)[i]; }
__device__ inline float2 gld(const float2* p, size_t i) {
  const f2v v = ((g_cf2*)p)[i];
  return make_float2(v.x, v.y);
}
__device__ inline float4 gld(const float4* p, size_t i) {
  const f4v v = ((g_cf4*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline void gst(float* p, size_t i, float v) { ((g_f*)p)[i] = v; }
__device__ inline void gst(float2* p, size_t i, float2 v) {
  f2v t;
  t.x = v.x;
  t.y = v.y;
  ((g_f2*)p)[i] = t;
}
__device__ inline void gst(float4* p, size_t i, float4 v) {
  f4v t;
  t.x = v.x;
  t.y = v.y;
  t.z = v.z;
  t.w = v.w;
  ((g_f4*)p)[i] = t;
}

// The same accesses at a 32-bit element offset from a wave-uniform base: the
// base stays in scalar registers and the byte offset in one VGPR (the
// global "saddr" form), instead of a 64-bit address pair per load.  Every
// plane here is far below 4 GB.
typedef __attribute__((address_space(1))) const char g_cc;
typedef __attribute__((address_space(1))) char g_c;
__device__ inline float gld32(const float* p, unsigned i) { return *(g_cf*)((g_cc*)p + (i << 2)); }
__device__ inline float2 gld32(const float2* p, unsigned i) {
  const f2v v = *(g_cf2*)((g_cc*)p + (i << 3));
  return make_float2(v.x, v.y);
}
__device__ inline float4 gld32(const float4* p, unsigned i) {
  const f4v v = *(g_cf4*)((g_cc*)p + (i << 4));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ inline void gst32(float2* p, unsigned i, float2 v) {
  f2v t;
  t.x = v.x;
  t.y = v.y;
  *(g_f2*)((g_c*)p + (i << 3)) = t;
}

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float re_mul(float2 a, float2 b) { return a.x * b.x - a.y * b.y; }

__device__ inline float2 carrier_at(const DevField& f, int y, int x) {
  if (f.carrier) return f.carrier[(size_t)y * f.w + x];
  float2 v = cmul(f.rowph[y], f.colph[x]);
  if (f.vsw && f.vsw[y] < 0) v.x = -v.x;
  return v;
}

// Deterministic block-wide FP64 sum; result valid in thread 0.
template <int NT>
__device__ inline double block_sum(double v, double* lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) lds[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) r += lds[i];
  __syncthreads();
  return r;
}

}  // namespace hvdk
