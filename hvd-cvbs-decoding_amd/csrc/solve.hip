// solve.hip -- the IRLS + conjugate-gradient variational solve on gfx950.
//
// Restates hvd-core's VariationalRefine 2-D path (engine/variational.cpp:
// 198-556) as matrix-free stencil kernels over a batch of fields of uneven
// size.  Per CG iteration two kernels run:
//   k_cg_curv  d <- -g + beta d (fused, on the fly), then the FP64 partial
//              sums of curv(d) (:352-399) and Re<g, d> (:72-117);
//   k_cg_step  chi <- chi + alpha d (on the fly), g <- grad(chi) (:286-347)
//              and the FP64 partial of |g|^2;
// a one-block-per-field kernel then sums that field's block partials in a
// fixed order and advances its scalar CG state (alpha, beta, the exits of
// :530 and :547-548), so no host round trip happens inside the solve and
// converged fields drop out by themselves (per-field masks).  chi and d are ping-ponged so a block
// never reads a neighbour's already-updated value.
#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

namespace {

struct TileCtx {
  int f;      // field index
  int lt;     // tile index within the field
  int x, y;   // pixel
  bool in;    // pixel inside the plane
  int nblk;   // blocks (tiles) of this field
};

__device__ inline TileCtx tile_ctx(const DevField* F, const int* tp, int nf, int nb) {
  TileCtx c;
  const int t = xcd_remap(blockIdx.x, nb);
  c.f = field_of_tile(tp, nf, t);
  const DevField& f = F[c.f];
  c.lt = t - tp[c.f];
  c.nblk = tp[c.f + 1] - tp[c.f];
  const int txi = c.lt % f.tiles_x, tyi = c.lt / f.tiles_x;
  c.x = txi * kTileW + (threadIdx.x & (kTileW - 1));
  c.y = tyi * kTileH + (threadIdx.x / kTileW);
  c.in = c.x < f.w && c.y < f.h;
  return c;
}

// chi + alpha d, written once so the centre pixel and its neighbours evaluate
// the identical expression (the value stored must equal the value used).
__device__ inline float2 axpy(float2 c, float alpha, float2 d) {
  return make_float2(__fmaf_rn(alpha, d.x, c.x), __fmaf_rn(alpha, d.y, c.y));
}
// -g + beta d (variational.cpp:541-543).
__device__ inline float2 dir_update(float2 g, float beta, float2 d) {
  return make_float2(__fmaf_rn(beta, d.x, -g.x), __fmaf_rn(beta, d.y, -g.y));
}

struct Sample {
  float yc;
  float2 chi;
};

__device__ inline Sample sample_at(const DevField& f, const float2* chi, const float2* dd, float alpha, int y,
                                   int x) {
  const size_t i = (size_t)y * f.w + x;
  float2 c = chi[i];
  if (dd) c = axpy(c, alpha, dd[i]);
  Sample s;
  s.chi = c;
  s.yc = f.s[i] - re_mul(c, carrier_at(f, y, x));
  return s;
}

// g = -2 lam_y [DxT(wx Dx yc) + DyT(wy Dy yc)] conj(c)
//     + 2 [mu_h DxT(wcx Dx chi) + mu_v DyT(wcy Dy chi)]      (:286-335)
// evaluated at one pixel from its 5-point neighbourhood.
__device__ inline float2 grad_at(const DevField& f, const FieldCG& st, const float2* chi, const float2* dd,
                                 float alpha, int y, int x, float2* chi_centre) {
  const Sample p0 = sample_at(f, chi, dd, alpha, y, x);
  *chi_centre = p0.chi;
  const bool hl = x > 0, hr = x + 1 < f.w, hu = y > 0, hd = y + 1 < f.h;
  Sample pl = p0, pr = p0, pu = p0, pd = p0;
  if (hl) pl = sample_at(f, chi, dd, alpha, y, x - 1);
  if (hr) pr = sample_at(f, chi, dd, alpha, y, x + 1);
  if (hu) pu = sample_at(f, chi, dd, alpha, y - 1, x);
  if (hd) pd = sample_at(f, chi, dd, alpha, y + 1, x);
  const size_t i = (size_t)y * f.w + x;
  const float4 w0 = f.wts[i];
  const float4 wl = hl ? f.wts[i - 1] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 wu = hu ? f.wts[i - f.w] : make_float4(0.f, 0.f, 0.f, 0.f);
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
  const float2 c0 = carrier_at(f, y, x);
  return make_float2(-img * c0.x + cp.x, img * c0.y + cp.y);
}

// Sum this block's per-thread values v[0..NV-1] and publish them as this
// tile's partials.  The per-field totals are formed by the k_reduce_* kernels
// that follow (a kernel boundary orders the two; no device-scope fence per
// block, which on a multi-XCD part would write back the L2 each time).
template <int NV>
__device__ inline void publish_partials(const double (&v)[NV], double* partials, int tile0, int lt) {
  __shared__ double lds[kBlock / 64];
  for (int k = 0; k < NV; ++k) {
    const double bs = block_sum<kBlock>(v[k], lds);
    if (threadIdx.x == 0) partials[(size_t)(tile0 + lt) * NV + k] = bs;
  }
}

// Fixed-order sum of field fi's tile partials (component k of NV), valid in
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
__global__ void __launch_bounds__(kBlock) k_irls_weights(const DevField* F, FieldCG* st, const int* tp, int nf,
                                                         int nb) {
  const TileCtx c = tile_ctx(F, tp, nf, nb);
  const DevField& f = F[c.f];
  const FieldCG& s = st[c.f];
  if (!c.in) return;
  const float2* chi = f.chi[s.chi_cur];
  const Sample p0 = sample_at(f, chi, nullptr, 0.f, c.y, c.x);
  const bool hr = c.x + 1 < f.w, hd = c.y + 1 < f.h;
  float gx = 0.f, gy = 0.f;
  float2 cx = make_float2(0.f, 0.f), cy = cx;
  if (hr) {
    const Sample pr = sample_at(f, chi, nullptr, 0.f, c.y, c.x + 1);
    gx = pr.yc - p0.yc;
    cx = make_float2(pr.chi.x - p0.chi.x, pr.chi.y - p0.chi.y);
  }
  if (hd) {
    const Sample pd = sample_at(f, chi, nullptr, 0.f, c.y + 1, c.x);
    gy = pd.yc - p0.yc;
    cy = make_float2(pd.chi.x - p0.chi.x, pd.chi.y - p0.chi.y);
  }
  const float e2 = s.eps * s.eps, ec2 = s.eps_c * s.eps_c;
  const float mx = sqrtf(cx.x * cx.x + cx.y * cx.y), my = sqrtf(cy.x * cy.x + cy.y * cy.y);
  float4 w;
  w.x = s.eps / sqrtf(gx * gx + e2);
  w.y = s.eps / sqrtf(gy * gy + e2);
  w.z = s.eps_c / sqrtf(mx * mx + s.k * gx * gx + ec2);
  w.w = s.eps_c / sqrtf(my * my + s.k * gy * gy + ec2);
  f.wts[(size_t)c.y * f.w + c.x] = w;
}

// Stage start (variational.cpp:519-525): g = grad(chi), d = -g, gg0 = |g|^2.
__global__ void __launch_bounds__(kBlock) k_grad_init(const DevField* F, FieldCG* st, const int* tp, int nf, int nb,
                                                      double* partials) {
  const TileCtx c = tile_ctx(F, tp, nf, nb);
  const DevField& f = F[c.f];
  FieldCG& s = st[c.f];
  double v[1] = {0.0};
  if (c.in) {
    float2 cc;
    const float2 g = grad_at(f, s, f.chi[s.chi_cur], nullptr, 0.f, c.y, c.x, &cc);
    const size_t i = (size_t)c.y * f.w + c.x;
    f.g[i] = g;
    f.d[s.d_cur][i] = make_float2(-g.x, -g.y);
    v[0] = (double)g.x * g.x + (double)g.y * g.y;
  }
  publish_partials<1>(v, partials, tp[c.f], c.lt);
}

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
__global__ void __launch_bounds__(kBlock) k_cg_curv(const DevField* F, FieldCG* st, const int* tp, int nf, int nb,
                                                    double* partials, int it) {
  const TileCtx c = tile_ctx(F, tp, nf, nb);
  const DevField& f = F[c.f];
  FieldCG& s = st[c.f];
  if (s.done) return;
  const float2* dsrc = f.d[s.d_cur];
  float2* ddst = f.d[s.d_cur ^ 1];
  const bool upd = it > 0;
  const float beta = s.beta;
  double v[2] = {0.0, 0.0};
  if (c.in) {
    const int x = c.x, y = c.y;
    const size_t i = (size_t)y * f.w + x;
    const bool hr = x + 1 < f.w, hd = y + 1 < f.h;
    const float2 g0 = f.g[i];
    float2 d0 = dsrc[i];
    if (upd) {
      d0 = dir_update(g0, beta, d0);
      ddst[i] = d0;
    }
    const float2 c0 = carrier_at(f, y, x);
    const float dy0 = -re_mul(d0, c0);
    float dxdy = 0.f, dydy = 0.f;
    float2 dxc = make_float2(0.f, 0.f), dyc = dxc;
    if (hr) {
      float2 dr = dsrc[i + 1];
      if (upd) dr = dir_update(f.g[i + 1], beta, dr);
      dxdy = -re_mul(dr, carrier_at(f, y, x + 1)) - dy0;
      dxc = make_float2(dr.x - d0.x, dr.y - d0.y);
    }
    if (hd) {
      float2 dd = dsrc[i + f.w];
      if (upd) dd = dir_update(f.g[i + f.w], beta, dd);
      dydy = -re_mul(dd, carrier_at(f, y + 1, x)) - dy0;
      dyc = make_float2(dd.x - d0.x, dd.y - d0.y);
    }
    const float4 w = f.wts[i];
    double h = (double)s.lam_y * ((double)w.x * dxdy * dxdy + (double)w.y * dydy * dydy);
    h += (double)s.mu_h * w.z * (double)(dxc.x * dxc.x + dxc.y * dxc.y);
    h += (double)s.mu_v * w.w * (double)(dyc.x * dyc.x + dyc.y * dyc.y);
    v[0] = h;
    v[1] = (double)g0.x * d0.x + (double)g0.y * d0.y;
  }
  publish_partials<2>(v, partials, tp[c.f], c.lt);
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
    if (hh <= 1e-12) {
      s.done = 1;
    } else {
      s.alpha = (float)(-0.5 * gd / hh);
    }
  }
}

// One CG step (variational.cpp:532-548): chi <- chi + alpha d, g <- grad(chi),
// beta, and the two early exits plus the iteration budget.
__global__ void __launch_bounds__(kBlock) k_cg_step(const DevField* F, FieldCG* st, const int* tp, int nf, int nb,
                                                    double* partials) {
  const TileCtx c = tile_ctx(F, tp, nf, nb);
  const DevField& f = F[c.f];
  FieldCG& s = st[c.f];
  if (s.done) return;
  const float2* chi = f.chi[s.chi_cur];
  float2* chi_out = f.chi[s.chi_cur ^ 1];
  const float2* dd = f.d[s.d_cur];
  double v[1] = {0.0};
  if (c.in) {
    float2 cc;
    const float2 g = grad_at(f, s, chi, dd, s.alpha, c.y, c.x, &cc);
    const size_t i = (size_t)c.y * f.w + c.x;
    f.g[i] = g;
    chi_out[i] = cc;
    v[0] = (double)g.x * g.x + (double)g.y * g.y;
  }
  publish_partials<1>(v, partials, tp[c.f], c.lt);
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
    if (ggn < s.tol2 * s.gg0 || ggn < 1e-10 * s.npix || s.it >= s.n_inner) s.done = 1;
  }
}

// Record the stage's iteration count (diagnostics).
__global__ void k_stage_end(FieldCG* st, int nf, int stage) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf && stage < 16) st[f].iters[stage] = st[f].it;
}

// Output (variational.cpp:552-555): Y = S - Re[chi c]; chi lands in the
// caller's buffer.
__global__ void __launch_bounds__(kBlock) k_refine_finalize(const DevField* F, const FieldCG* st, const int* tp,
                                                            int nf, int nb) {
  const TileCtx c = tile_ctx(F, tp, nf, nb);
  const DevField& f = F[c.f];
  if (!c.in) return;
  const size_t i = (size_t)c.y * f.w + c.x;
  const float2 chi = f.chi[st[c.f].chi_cur][i];
  if (st[c.f].chi_cur) f.chi[0][i] = chi;
  if (f.y) f.y[i] = f.s[i] - re_mul(chi, carrier_at(f, c.y, c.x));
}

}  // namespace hvdk
