// solve.hip -- the IRLS + conjugate-gradient variational solve on gfx950.
//
// Restates hvd-core's VariationalRefine 2-D path (engine/variational.cpp:
// 198-556) as matrix-free stencil kernels over a batch of fields of uneven
// size.  Per CG iteration two kernels run:
//   k_cg_curv  d <- -g + beta d (fused, on the fly), then the FP64 partial
//              sums of curv(d) (:352-399) and Re<g, d> (:72-117);
//   k_cg_step  chi <- chi + alpha d (on the fly), g <- grad(chi) (:286-347)
//              and the FP64 partial of |g|^2;
// a one-block-per-field kernel then sums that field's strip partials in a
// fixed order and advances its scalar CG state (alpha, beta, the exits of
// :530 and :547-548), so no host round trip happens inside the solve and
// converged fields drop out by themselves (per-field masks).  chi and d are ping-ponged so a block
// never reads a neighbour's already-updated value.
#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

namespace {

// A wave (one 64-thread block) owns a strip of kStripCols output columns and
// `rows` output lines of one field.  Lane l sits on column x = x0 + l - 1, so
// lanes 0 and 63 are a one-column halo whose values reach their neighbours
// through lane shuffles; each lane walks its rows top to bottom with the rows
// above and below held in registers.  Every per-pixel value is formed by the
// same expression as at the neighbouring pixel that also needs it, so a value
// shuffled in from a neighbour lane is bit-identical to one recomputed there.
// The field descriptor and CG state are copied into registers up front (they
// are restrict-qualified, so the copies are scalar loads that no store in the
// loop can force to be reloaded).
struct StripCtx {
  int f;        // field index
  int lt;       // strip index within the field
  int x;        // this lane's column (-1 or >= w on halo lanes)
  int y0;       // first output row
  bool valid;   // column inside the plane
  bool out;     // lane produces outputs (not a halo lane, inside the plane)
};

__device__ inline int lane_id() { return threadIdx.x & 63; }

__device__ inline StripCtx strip_ctx(const DevField* F, const StripGrid& sg) {
  StripCtx c;
  const int t = xcd_remap(blockIdx.x, sg.nb);
  // equal-sized fields: the owner is a division, not a dependent search
  c.f = sg.spf > 0 ? t / sg.spf : field_of_tile(sg.tp, sg.nf, t);
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  c.lt = sg.spf > 0 ? t - c.f * sg.spf : t - sg.tp[c.f];
  const int lane = lane_id();
  c.x = (c.lt % f.tiles_x) * kStripCols + lane - 1;
  c.y0 = (c.lt / f.tiles_x) * sg.rows;
  c.valid = c.x >= 0 && c.x < f.w;
  c.out = lane >= 1 && lane <= kStripCols && c.x < f.w;
  return c;
}

// The carrier down this lane's column (ntsc_geometry.cpp:75-93).  The row
// phasors (and V-switch signs) of the strip's rows y0 - 1 ... y0 + rows + 1
// are loaded once, one per lane, and broadcast per row with a lane read; the
// column phasor is loaded once per lane.
struct Carrier {
  float2 col;
  float2 rowl;          // row phasor of row ybase + lane
  int vsl;              // its V-switch sign (< 0: switched line)
  int x, ybase, w;
  const float2* plane;  // explicit carrier plane, or null
  __device__ inline float2 at(int y) const {
    if (plane) return gld(plane, (size_t)y * w + x);
    const int l = y - ybase;  // uniform across the wave
    const float2 rp = make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(rowl.x), l)),
                                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rowl.y), l)));
    float2 v = cmul(rp, col);
    if (__builtin_amdgcn_readlane(vsl, l) < 0) v.x = -v.x;
    return v;
  }
};

__device__ inline Carrier carrier_for(const DevField& f, const StripCtx& c) {
  Carrier k;
  k.x = c.x;
  k.w = f.w;
  k.ybase = c.y0 - 1;
  k.plane = f.carrier;
  k.col = make_float2(0.f, 0.f);
  k.rowl = make_float2(0.f, 0.f);
  k.vsl = 1;
  if (!f.carrier) {
    if (c.valid) k.col = gld(f.colph, c.x);
    const int yy = k.ybase + lane_id();
    if (yy >= 0 && yy < f.h) {
      k.rowl = gld(f.rowph, yy);
      if (f.vsw) k.vsl = gld(f.vsw, yy);
    }
  }
  return k;
}

// Element of a two-entry pointer table, selected without indexing (keeps the
// descriptor copy in registers).
template <typename T>
__device__ inline T* pick(T* const (&a)[2], int i) { return i ? a[1] : a[0]; }

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Neighbour exchange: value of lane - 1 (left) / lane + 1 (right).  Lanes 0
// and 63 get their own value back and never use it.
__device__ inline float shl(float v) { return __shfl_up(v, 1, 64); }
__device__ inline float shr(float v) { return __shfl_down(v, 1, 64); }
__device__ inline float2 shl(float2 v) { return make_float2(shl(v.x), shl(v.y)); }
__device__ inline float2 shr(float2 v) { return make_float2(shr(v.x), shr(v.y)); }
__device__ inline float4 shl(float4 v) { return make_float4(shl(v.x), shl(v.y), shl(v.z), shl(v.w)); }

// chi + alpha d, written once so the centre pixel and its neighbours evaluate
// the identical expression (the value stored must equal the value used).
__device__ inline float2 axpy(float2 c, float alpha, float2 d) {
  return make_float2(__fmaf_rn(alpha, d.x, c.x), __fmaf_rn(alpha, d.y, c.y));
}
// -g + beta d (variational.cpp:541-543).
__device__ inline float2 dir_update(float2 g, float beta, float2 d) {
  return make_float2(__fmaf_rn(beta, d.x, -g.x), __fmaf_rn(beta, d.y, -g.y));
}

inline __device__ bool row_in(const DevField& f, int y, bool valid) { return valid && y >= 0 && y < f.h; }

// Loads of one row of this lane's column.
struct RawRow {
  float2 chi, d;
  float s;
};

__device__ inline RawRow raw_row(const DevField& f, const Carrier& k, const float2* chi, const float2* dd, int y,
                                 bool valid) {
  RawRow r{make_float2(0.f, 0.f), make_float2(0.f, 0.f), 0.f};
  if (row_in(f, y, valid)) {
    const size_t i = (size_t)y * f.w + k.x;
    r.chi = gld(chi, i);
    if (dd) r.d = gld(dd, i);
    r.s = f.s ? gld(f.s, i) : 0.f;
  }
  return r;
}

struct Sample {
  float yc;    // S - Re[chi c]
  float2 chi;  // chi (+ alpha d)
  float2 car;  // carrier
};

// The sample of row y from its loads; zeros outside the plane.
__device__ inline Sample make_sample(const DevField& f, const Carrier& k, const RawRow& r, bool with_d, float alpha,
                                     int y, bool valid) {
  Sample s{0.f, make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
  if (row_in(f, y, valid)) {
    const float2 c = with_d ? axpy(r.chi, alpha, r.d) : r.chi;
    s.chi = c;
    s.car = k.at(y);
    s.yc = r.s - re_mul(c, s.car);
  }
  return s;
}

__device__ inline float4 weights_row(const DevField& f, int y, int x, bool valid) {
  return row_in(f, y, valid) ? gld(f.wts, (size_t)y * f.w + x) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// g = -2 lam_y [DxT(wx Dx yc) + DyT(wy Dy yc)] conj(c)
//     + 2 [mu_h DxT(wcx Dx chi) + mu_v DyT(wcy Dy chi)]      (:286-335)
// from the 5-point neighbourhood; the has-neighbour flags apply the Neumann
// boundaries of gradients.h (a missing neighbour contributes a zero difference).
__device__ inline float2 grad_from(const FieldCG& st, const Sample& p0, const Sample& pl, const Sample& pr,
                                   const Sample& pu, const Sample& pd, float4 w0, float4 wl, float4 wu, bool hl,
                                   bool hr, bool hu, bool hd) {
  if (!hl) wl = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!hu) wu = make_float4(0.f, 0.f, 0.f, 0.f);
  // luma data term
  const float gx0 = hr ? pr.yc - p0.yc : 0.f;
  const float gxl = hl ? p0.yc - pl.yc : 0.f;
  const float gy0 = hd ? pd.yc - p0.yc : 0.f;
  const float gyu = hu ? p0.yc - pu.yc : 0.f;
  const float img = 2.f * st.lam_y * ((wl.x * gxl - w0.x * gx0) + (wu.y * gyu - w0.y * gy0));
  // chroma prior
  const float2 cx0 = hr ? make_float2(pr.chi.x - p0.chi.x, pr.chi.y - p0.chi.y) : make_float2(0.f, 0.f);
  const float2 cxl = hl ? make_float2(p0.chi.x - pl.chi.x, p0.chi.y - pl.chi.y) : make_float2(0.f, 0.f);
  const float2 cy0 = hd ? make_float2(pd.chi.x - p0.chi.x, pd.chi.y - p0.chi.y) : make_float2(0.f, 0.f);
  const float2 cyu = hu ? make_float2(p0.chi.x - pu.chi.x, p0.chi.y - pu.chi.y) : make_float2(0.f, 0.f);
  const float2 dxt = make_float2(wl.z * cxl.x - w0.z * cx0.x, wl.z * cxl.y - w0.z * cx0.y);
  const float2 dyt = make_float2(wu.w * cyu.x - w0.w * cy0.x, wu.w * cyu.y - w0.w * cy0.y);
  const float2 cp = make_float2(2.f * (st.mu_h * dxt.x + st.mu_v * dyt.x), 2.f * (st.mu_h * dxt.y + st.mu_v * dyt.y));
  const float2 c0 = p0.car;
  return make_float2(-img * c0.x + cp.x, img * c0.y + cp.y);
}

// Walks a strip computing g at every output pixel from chi (+ alpha d); calls
// emit(i, g, chi_centre) for each.  Shared by the stage-start gradient and the
// CG step.
template <typename Emit>
__device__ inline void grad_strip(const DevField& f, const FieldCG& st, const StripCtx& c, int rows,
                                  const float2* chi, const float2* dd, float alpha, Emit emit) {
  const Carrier k = carrier_for(f, c);
  const bool wd = dd != nullptr;
  Sample up = make_sample(f, k, raw_row(f, k, chi, dd, c.y0 - 1, c.valid), wd, alpha, c.y0 - 1, c.valid);
  Sample cur = make_sample(f, k, raw_row(f, k, chi, dd, c.y0, c.valid), wd, alpha, c.y0, c.valid);
  float4 wu = weights_row(f, c.y0 - 1, c.x, c.valid);
  float4 wc = weights_row(f, c.y0, c.x, c.valid);
  for (int r = 0; r < rows; ++r) {
    const int y = c.y0 + r;
    if (y >= f.h) break;  // uniform across the wave
    const float4 wn = weights_row(f, y + 1, c.x, c.valid);
    const Sample dn = make_sample(f, k, raw_row(f, k, chi, dd, y + 1, c.valid), wd, alpha, y + 1, c.valid);
    Sample pl, pr;
    pl.yc = shl(cur.yc);
    pr.yc = shr(cur.yc);
    pl.chi = shl(cur.chi);
    pr.chi = shr(cur.chi);
    const float4 wl = shl(wc);
    if (c.out) {
      float2 g = grad_from(st, cur, pl, pr, up, dn, wc, wl, wu, c.x > 0, c.x + 1 < f.w, y > 0, y + 1 < f.h);
      const size_t i = (size_t)y * f.w + c.x;
      if (f.rhs) {
        const float2 b = gld(f.rhs, i);
        g = make_float2(g.x - b.x, g.y - b.y);
      }
      emit(i, g, cur.chi);
    }
    up = cur;
    cur = dn;
    wu = wc;
    wc = wn;
  }
}

// Wave-level FP64 partial of this strip, published to its slot.  The
// per-field totals are formed by the k_reduce_* kernels that follow (a kernel
// boundary orders the two; no device-scope fence per wave, which on a
// multi-XCD part would write back the L2 each time).
__device__ inline void publish(double* partials, int tile0, int lt, int nv, int k, double v) {
  v = wave_sum(v);
  if (lane_id() == 0) partials[(size_t)(tile0 + lt) * nv + k] = v;
}

// Fixed-order sum of field fi's strip partials (component k of NV), valid in
// thread 0 of a 256-thread block.
template <int NV>
__device__ inline double field_total(const double* partials, const int* tp, int fi, int k) {
  __shared__ double lds[4];
  const int t0 = tp[fi], nblk = tp[fi + 1] - t0;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) acc += partials[(size_t)(t0 + b) * NV + k];
  return block_sum<256>(acc, lds);
}

}  // namespace

// Reset per-field solver state before a refine (replaces any memset so the
// sequence stays graph-capturable).
__global__ void k_refine_begin(FieldCG* st, const FieldCG* init, int nf) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) st[f] = init[f];
}

// ResolveChromaAniso samples (variational.cpp:150-178): group-averaged chi0,
// t = 1,3,5,..., x = 1,4,7,...; |pav(t,x) - pav(t,x-1)| and
// |pav(t,x) - pav(t-1,x)|.  One thread per sample over the whole batch.
__global__ void k_aniso_samples(const DevField* F, const FieldCG* st, const int* sample_prefix, int nf,
                                float* gx, float* gy) {
  const int n_total = sample_prefix[nf];
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n_total; idx += gridDim.x * blockDim.x) {
    const int fi = field_of_tile(sample_prefix, nf, idx);
    const DevField& f = F[fi];
    const int group = st[fi].is_pal ? 4 : 2;
    const int nx = (f.w - 1 + 2) / 3;
    const int j = idx - sample_prefix[fi];
    const int t = 1 + 2 * (j / nx);
    const int x = 1 + 3 * (j % nx);
    const float2* c = f.chi[0];
    float2 p = make_float2(0.f, 0.f), pl = p, pu = p;
    for (int k = 0; k < group; ++k) {
      const float2 a = c[(size_t)(group * t + k) * f.w + x];
      const float2 b = c[(size_t)(group * t + k) * f.w + x - 1];
      const float2 u = c[(size_t)(group * (t - 1) + k) * f.w + x];
      p.x += a.x; p.y += a.y;
      pl.x += b.x; pl.y += b.y;
      pu.x += u.x; pu.y += u.y;
    }
    const float inv = 1.f / (float)group;
    const float ex = (p.x - pl.x) * inv, ey = (p.y - pl.y) * inv;
    const float ux = (p.x - pu.x) * inv, uy = (p.y - pu.y) * inv;
    gx[idx] = sqrtf(ex * ex + ey * ey);
    gy[idx] = sqrtf(ux * ux + uy * uy);
  }
}

// Exact k-th smallest of n non-negative floats by 4 x 8-bit radix selection
// (the value std::nth_element returns at k).  One block, 256 threads.
__device__ float radix_select(const float* v, int n, int k) {
  __shared__ unsigned hist[256];
  __shared__ unsigned s_prefix, s_k;
  if (threadIdx.x == 0) { s_prefix = 0u; s_k = (unsigned)k; }
  __syncthreads();
  for (int pass = 3; pass >= 0; --pass) {
    hist[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned prefix = s_prefix;
    const int hi_shift = (pass + 1) * 8;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const unsigned u = __float_as_uint(v[i]);
      const bool match = hi_shift >= 32 ? true : ((u >> hi_shift) == prefix);
      if (match) atomicAdd(&hist[(u >> (pass * 8)) & 255u], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned kk = s_k, b = 0;
      for (; b < 256; ++b) {
        if (kk < hist[b]) break;
        kk -= hist[b];
      }
      s_k = kk;
      s_prefix = (prefix << 8) | b;
    }
    __syncthreads();
  }
  return __uint_as_float(s_prefix);
}

// p98 of both sample sets, the ratio and the NTSC / PAL map into aniso
// (variational.cpp:182-195); sets mu_h, mu_v.  One block per field.
__global__ void k_aniso_select(FieldCG* st, const int* sample_prefix, const float* gx, const float* gy,
                               float lambda_c) {
  const int fi = blockIdx.x;
  FieldCG& s = st[fi];
  float aniso = s.aniso_forced;
  if (!(aniso > 0.f)) {
    const int off = sample_prefix[fi], n = sample_prefix[fi + 1] - off;
    float px = 0.f, py = 0.f;
    if (n > 0) {
      const int k = n - 1 - n / 50;
      px = radix_select(gx + off, n, k);
      py = radix_select(gy + off, n, k);
    }
    const float r = py / fmaxf(px, 1e-6f);
    aniso = s.is_pal ? fminf(fmaxf(0.7f + 0.9f * (r - 1.3f), 0.55f), 1.f)
                     : fminf(fmaxf(0.5f + 1.1f * (r - 1.3f), 0.5f), 1.f);
  }
  if (threadIdx.x == 0) {
    s.mu_h = lambda_c * aniso;
    s.mu_v = lambda_c;
  }
}

// IRLS lagged-diffusivity weights at the current chi (variational.cpp:406-433).
__global__ void __launch_bounds__(64) k_irls_weights(const DevField* __restrict__ F,
    const FieldCG* __restrict__ st, StripGrid sg, int stage) {
  const StripCtx c = strip_ctx(F, sg);
  const int rows = sg.rows;
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  const FieldCG s = st[c.f];
  const Carrier k = carrier_for(f, c);
  const float2* chi = pick(f.chi, s.chi_cur);
  const float e2 = s.eps * s.eps, ec2 = s.eps_c * s.eps_c;
  Sample cur = make_sample(f, k, raw_row(f, k, chi, nullptr, c.y0, c.valid), false, 0.f, c.y0, c.valid);
  for (int r = 0; r < rows; ++r) {
    const int y = c.y0 + r;
    if (y >= f.h) break;
    const Sample dn = make_sample(f, k, raw_row(f, k, chi, nullptr, y + 1, c.valid), false, 0.f, y + 1, c.valid);
    const float yr = shr(cur.yc);
    const float2 cr = shr(cur.chi);
    if (c.out) {
      const bool hr = c.x + 1 < f.w, hd = y + 1 < f.h;
      const float gx = hr ? yr - cur.yc : 0.f;
      const float gy = hd ? dn.yc - cur.yc : 0.f;
      const float2 cx = hr ? make_float2(cr.x - cur.chi.x, cr.y - cur.chi.y) : make_float2(0.f, 0.f);
      const float2 cy = hd ? make_float2(dn.chi.x - cur.chi.x, dn.chi.y - cur.chi.y) : make_float2(0.f, 0.f);
      const float mx = sqrtf(cx.x * cx.x + cx.y * cx.y), my = sqrtf(cy.x * cy.x + cy.y * cy.y);
      float4 w;
      w.x = s.eps / sqrtf(gx * gx + e2);
      w.y = s.eps / sqrtf(gy * gy + e2);
      w.z = s.eps_c / sqrtf(mx * mx + s.k * gx * gx + ec2);
      w.w = s.eps_c / sqrtf(my * my + s.k * gy * gy + ec2);
      gst(f.wts, (size_t)y * f.w + c.x, w);
      if (f.save) gst(f.save + (size_t)stage * f.h * f.w, (size_t)y * f.w + c.x, cur.chi);
    }
    cur = dn;
  }
}

// Stage start (variational.cpp:519-525): g = grad(chi), d = -g, partial |g|^2.
__global__ void __launch_bounds__(64) k_grad_init(const DevField* __restrict__ F,
    const FieldCG* __restrict__ st, StripGrid sg,
                                                                 double* partials) {
  const StripCtx c = strip_ctx(F, sg);
  const int rows = sg.rows;
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  const FieldCG s = st[c.f];
  float2* dd = pick(f.d, s.d_cur);
  double gg = 0.0;
  grad_strip(f, s, c, rows, pick(f.chi, s.chi_cur), nullptr, 0.f, [&](size_t i, float2 g, float2) {
    gst(f.g, i, g);
    gst(dd, i, make_float2(-g.x, -g.y));
    gg += (double)g.x * g.x + (double)g.y * g.y;
  });
  publish(partials, F[c.f].tile0, c.lt, 1, 0, gg);
}

// gg0 = |g|^2 and the inner-loop reset (variational.cpp:524-525).
__global__ void __launch_bounds__(256) k_reduce_init(FieldCG* st, const int* tp, const double* partials) {
  const int fi = blockIdx.x;
  FieldCG& s = st[fi];
  const double gg = field_total<1>(partials, tp, fi, 0);
  if (threadIdx.x == 0) {
    s.gg = gg;
    s.gg0 = gg;
    s.it = 0;
    s.done = s.n_inner <= 0 ? 1 : 0;
  }
}

// One CG curvature pass (variational.cpp:527-531): for it > 0 first applies
// the previous iteration's d <- -g + beta d (:541-543) on the fly.
__global__ void __launch_bounds__(64) k_cg_curv(const DevField* __restrict__ F,
    const FieldCG* __restrict__ st, StripGrid sg,
                                                               double* partials, int it) {
  const StripCtx c = strip_ctx(F, sg);
  const int rows = sg.rows;
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  const FieldCG s = st[c.f];
  if (s.done) return;
  const Carrier k = carrier_for(f, c);
  const float2* dsrc = pick(f.d, s.d_cur);
  float2* ddst = pick(f.d, s.d_cur ^ 1);
  const bool upd = it > 0;
  const float beta = s.beta;
  struct Raw {
    float2 g, d;
  };
  auto raw = [&](int y) -> Raw {
    Raw r{make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    if (row_in(f, y, c.valid)) {
      const size_t i = (size_t)y * f.w + c.x;
      r.g = gld(f.g, i);
      r.d = gld(dsrc, i);
    }
    return r;
  };
  // direction, gradient and -Re[d c] of row y in this lane's column
  struct Row {
    float2 d, g;
    float dy;
  };
  auto row = [&](const Raw& q, int y) -> Row {
    Row r{make_float2(0.f, 0.f), make_float2(0.f, 0.f), 0.f};
    if (row_in(f, y, c.valid)) {
      r.g = q.g;
      r.d = upd ? dir_update(q.g, beta, q.d) : q.d;
      r.dy = -re_mul(r.d, k.at(y));
    }
    return r;
  };
  double hh = 0.0, gd = 0.0;
  Row cur = row(raw(c.y0), c.y0);
  for (int r = 0; r < rows; ++r) {
    const int y = c.y0 + r;
    if (y >= f.h) break;
    const float4 w = c.out ? gld(f.wts, (size_t)y * f.w + c.x) : make_float4(0.f, 0.f, 0.f, 0.f);
    const Row dn = row(raw(y + 1), y + 1);
    const float2 dr = shr(cur.d);
    const float dyr = shr(cur.dy);
    if (c.out) {
      const size_t i = (size_t)y * f.w + c.x;
      if (upd) gst(ddst, i, cur.d);
      const bool hr = c.x + 1 < f.w, hd = y + 1 < f.h;
      float dxdy = 0.f, dydy = 0.f;
      float2 dxc = make_float2(0.f, 0.f), dyc = dxc;
      if (hr) {
        dxdy = dyr - cur.dy;
        dxc = make_float2(dr.x - cur.d.x, dr.y - cur.d.y);
      }
      if (hd) {
        dydy = dn.dy - cur.dy;
        dyc = make_float2(dn.d.x - cur.d.x, dn.d.y - cur.d.y);
      }
      double h = (double)s.lam_y * ((double)w.x * dxdy * dxdy + (double)w.y * dydy * dydy);
      h += (double)s.mu_h * w.z * (double)(dxc.x * dxc.x + dxc.y * dxc.y);
      h += (double)s.mu_v * w.w * (double)(dyc.x * dyc.x + dyc.y * dyc.y);
      hh += h;
      gd += (double)cur.g.x * cur.d.x + (double)cur.g.y * cur.d.y;
    }
    cur = dn;
  }
  publish(partials, F[c.f].tile0, c.lt, 2, 0, hh);
  publish(partials, F[c.f].tile0, c.lt, 2, 1, gd);
}

// alpha = -1/2 <g, d> / curv(d), or stop when curv(d) <= 1e-12
// (variational.cpp:528-531); commits the d ping-pong of iteration it.
__global__ void __launch_bounds__(256) k_reduce_curv(FieldCG* st, const int* tp, const double* partials, int it) {
  const int fi = blockIdx.x;
  FieldCG& s = st[fi];
  if (s.done) return;
  const double hh = field_total<2>(partials, tp, fi, 0);
  const double gd = field_total<2>(partials, tp, fi, 1);
  if (threadIdx.x == 0) {
    if (it > 0) s.d_cur ^= 1;
    s.hh = hh;
    s.gd = gd;
    if (hh <= s.hh_floor) {
      s.done = 1;
    } else {
      s.alpha = (float)(-0.5 * gd / hh);
    }
  }
}

// One CG step (variational.cpp:532-548): chi <- chi + alpha d, g <- grad(chi),
// partial |g|^2.
__global__ void __launch_bounds__(64) k_cg_step(const DevField* __restrict__ F,
    const FieldCG* __restrict__ st, StripGrid sg,
                                                               double* partials) {
  const StripCtx c = strip_ctx(F, sg);
  const int rows = sg.rows;
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  const FieldCG s = st[c.f];
  if (s.done) return;
  float2* chi_out = pick(f.chi, s.chi_cur ^ 1);
  double gg = 0.0;
  grad_strip(f, s, c, rows, pick(f.chi, s.chi_cur), pick(f.d, s.d_cur), s.alpha,
             [&](size_t i, float2 g, float2 cc) {
    gst(f.g, i, g);
    gst(chi_out, i, cc);
    gg += (double)g.x * g.x + (double)g.y * g.y;
  });
  publish(partials, F[c.f].tile0, c.lt, 1, 0, gg);
}

// beta, the chi ping-pong, and the exits of variational.cpp:540-548.
__global__ void __launch_bounds__(256) k_reduce_step(FieldCG* st, const int* tp, const double* partials) {
  const int fi = blockIdx.x;
  FieldCG& s = st[fi];
  if (s.done) return;
  const double ggn = field_total<1>(partials, tp, fi, 0);
  if (threadIdx.x == 0) {
    s.chi_cur ^= 1;
    s.beta = (float)(ggn / (s.gg > 1e-30 ? s.gg : 1e-30));
    s.gg = ggn;
    s.it += 1;
    if (ggn < s.tol2 * s.gg0 || ggn < s.gg_floor * s.npix || s.it >= s.n_inner) s.done = 1;
  }
}

// Record the stage's iteration count (diagnostics).
__global__ void k_stage_end(FieldCG* st, int nf, int stage) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf && stage < 16) st[f].iters[stage] = st[f].it;
}

// Output (variational.cpp:552-555): Y = S - Re[chi c]; chi lands in the
// caller's buffer.
__global__ void __launch_bounds__(64) k_refine_finalize(const DevField* __restrict__ F,
    const FieldCG* __restrict__ st,
                                                                       StripGrid sg) {
  const StripCtx c = strip_ctx(F, sg);
  const int rows = sg.rows;
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  if (!c.out) return;
  const Carrier k = carrier_for(f, c);
  const int cur = st[c.f].chi_cur;
  for (int r = 0; r < rows; ++r) {
    const int y = c.y0 + r;
    if (y >= f.h) break;
    const size_t i = (size_t)y * f.w + c.x;
    const float2 chi = gld(pick(f.chi, cur), i);
    if (cur) gst(f.chi[0], i, chi);
    if (f.y) gst(f.y, i, gld(f.s, i) - re_mul(chi, k.at(y)));
  }
}

}  // namespace hvdk
