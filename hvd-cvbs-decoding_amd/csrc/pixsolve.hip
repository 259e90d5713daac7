// pixsolve.hip -- the per-pixel IRLS + CG solve for the terms the strip
// solver does not carry: the oriented (+-45 deg) chroma prior (diag_prior > 0)
// and motion-compensated temporal neighbours (hvd-core engine/variational.cpp:
// 198-556: :205-219 mu renormalisation, :255-281 neighbour deltas, :286-347
// grad with the D1/D2 terms of gradients.h:195-277, :352-399 curv, :406-516
// weights, prior scale), for a batch of fields of uneven size, each with its
// own neighbours.
//
// Work split: one block of kPxWaves waves per chunk of kPxChunk consecutive
// pixels of one field, wave w taking the chunk's w-th quarter (lane l takes
// pixels base + l + 64 k, so every load of a step is coalesced); the
// stencil's neighbours are gathered from the L2.  Two launches per CG
// iteration:
//   k_px_dir  : beta and the tolerance exits from the previous step's |g|^2
//               partials; d = -g + beta d_prev; partials of curv(d), Re<g, d>;
//   k_px_step : alpha (or the curvature exit) from those partials; chi' =
//               chi + alpha d formed on the fly at the 3x3 stencil; chi' and
//               g = grad(chi') written; partials of |g|^2.
// A launch's sums (|g|^2, the curvature, <g, d>) go out as one partial per
// block (its waves' sums added in wave order); the next launch's blocks each
// add the field's partials in one fixed order (wave 0, batched loads, the
// result shared through LDS), so every block of a field takes the same
// alpha, beta and exits, whatever the batch.  A launch never reads state
// that the same launch writes: the stop flag is double-buffered (k_px_dir
// reads done[0] and writes done[1], k_px_step the reverse), gg0 is written by
// the first iteration's k_px_dir and read from the second on, and the |g|^2
// partials alternate between two slots by iteration parity.
#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

namespace {

__device__ inline int px_lane() { return threadIdx.x & 63; }

// The wave's field and chunk.
struct PxCtx {
  int fi, chunk;  // field, chunk within it
};

__device__ inline PxCtx px_ctx(const int* cp, int nf) {
  PxCtx c;
  const int t = blockIdx.x;
  c.fi = field_of_tile(cp, nf, t);
  c.chunk = t - cp[c.fi];
  return c;
}

__device__ inline double px_wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Pixel k of this lane: wave w of the block takes the block's w-th quarter.
__device__ inline size_t px_index(const PxCtx& c, int k) {
  return (size_t)c.chunk * kPxChunk + (size_t)(threadIdx.x >> 6) * (64 * kPxPerLane) + (size_t)k * 64 + px_lane();
}

// The block's share of one or two field sums: the waves' sums added in wave
// order into part[slot][chunk].  Called by every thread of the block, on
// uniform control flow.
__device__ void px_store(const PxField& P, int chunk, int nv, double v0, double v1, int s0, int s1) {
  __shared__ double red[kPxWaves][2];
  const int wv = threadIdx.x >> 6;
  v0 = px_wave_sum(v0);
  v1 = px_wave_sum(v1);
  if (px_lane() == 0) {
    red[wv][0] = v0;
    red[wv][1] = v1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b0 = red[0][0], b1 = red[0][1];
    for (int k = 1; k < kPxWaves; ++k) {
      b0 += red[k][0];
      b1 += red[k][1];
    }
    const int nc = P.nchunks;
    P.part[(size_t)s0 * nc + chunk] = b0;
    if (nv > 1) P.part[(size_t)s1 * nc + chunk] = b1;
  }
}

// The field totals of one or two partial slots an earlier launch stored:
// wave 0 adds the chunk partials in one fixed order (lane l adds l, l + 64,
// ..., eight loads in flight at a time, then the butterfly) and hands the
// totals to the block through LDS, so every block of the field gets the same
// bits.  Called by every thread of the block, on uniform control flow.
__device__ void px_totals(const PxField& P, int nv, int s0, int s1, double* t0, double* t1) {
  __shared__ double tot[2];
  const int nc = P.nchunks;
  if (threadIdx.x < 64) {
    for (int q = 0; q < nv; ++q) {
      const double* part = P.part + (size_t)(q == 0 ? s0 : s1) * nc;
      double acc = 0.0;
      for (int i0 = px_lane(); i0 < nc; i0 += 64 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = i0 + 64 * u < nc ? part[i0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (i0 + 64 * u < nc) acc += v[u];
      }
      acc = px_wave_sum(acc);
      if (px_lane() == 0) tot[q] = acc;
    }
  }
  __syncthreads();
  *t0 = tot[0];
  if (nv > 1) *t1 = tot[1];
}

__device__ inline float2 px_carrier(const DevField& f, int y, int x) {
#pragma clang fp contract(off)
  if (f.carrier) return f.carrier[(size_t)y * f.w + x];
  const float2 r = f.rowph[y], c = f.colph[x];
  float2 v = make_float2(r.x * c.x - r.y * c.y, r.x * c.y + r.y * c.x);
  if (f.vsw && f.vsw[y] < 0) v.x = -v.x;
  return v;
}

// chi (+ alpha d) and S - Re[chi c] of one pixel; the caller keeps (y, x)
// inside the plane.
struct PxSample {
  float yc;
  float2 chi;
  float2 car;
};

__device__ inline PxSample px_sample(const DevField& f, const float2* chi, const float2* d, float alpha, int y, int x) {
#pragma clang fp contract(off)
  const size_t i = (size_t)y * f.w + x;
  PxSample p;
  p.chi = chi[i];
  if (d) {
    const float2 di = d[i];
    p.chi.x += alpha * di.x;
    p.chi.y += alpha * di.y;
  }
  p.car = px_carrier(f, y, x);
  p.yc = f.s[i] - (p.chi.x * p.car.x - p.chi.y * p.car.y);
  return p;
}

__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cscale(float w, float2 a) { return make_float2(w * a.x, w * a.y); }
__device__ inline float cnorm(float2 a) { return a.x * a.x + a.y * a.y; }

// grad at pixel (y, x) of the fixed-weight objective (:286-347): the luma and
// axis chroma fluxes across the four edges, the diagonal fluxes across the
// four corners (D1T, D2T of gradients.h:214-277), then per neighbour
// 2 nu wt rt conj(dc).  Weights of differences that leave the plane are 0
// (k_px_weights), so those terms vanish as hvd-core's Neumann zeros do;
// samples outside the plane are clamped reads that such a weight multiplies.
__device__ inline float2 px_grad(const PxField& P, const FieldCG& s, const float2* chi, const float2* d, float a,
                                 int y, int x, float2* chi_out) {
#pragma clang fp contract(off)
  const DevField& f = P.f;
  const int xl = max(x - 1, 0), xr = min(x + 1, f.w - 1), yu = max(y - 1, 0), yd = min(y + 1, f.h - 1);
  const PxSample p0 = px_sample(f, chi, d, a, y, x), pl = px_sample(f, chi, d, a, y, xl),
                 pr = px_sample(f, chi, d, a, y, xr), pu = px_sample(f, chi, d, a, yu, x),
                 pd = px_sample(f, chi, d, a, yd, x);
  if (chi_out) *chi_out = p0.chi;
  const size_t i = (size_t)y * f.w + x;
  const float4 w0 = P.wts[i];
  const float4 wl = x > 0 ? P.wts[i - 1] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 wu = y > 0 ? P.wts[i - f.w] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float fx0 = w0.x * (pr.yc - p0.yc), fxl = wl.x * (p0.yc - pl.yc);
  const float fy0 = w0.y * (pd.yc - p0.yc), fyu = wu.y * (p0.yc - pu.yc);
  const float img = 2.f * s.lam_y * ((fxl - fx0) + (fyu - fy0));
  const float2 cx0 = cscale(w0.z, csub(pr.chi, p0.chi)), cxl = cscale(wl.z, csub(p0.chi, pl.chi));
  const float2 cy0 = cscale(w0.w, csub(pd.chi, p0.chi)), cyu = cscale(wu.w, csub(p0.chi, pu.chi));
  float2 cp = make_float2(2.f * (s.mu_h * (cxl.x - cx0.x) + s.mu_v * (cyu.x - cy0.x)),
                          2.f * (s.mu_h * (cxl.y - cx0.y) + s.mu_v * (cyu.y - cy0.y)));
  if (P.wd) {
    // a1 = wd1 (chi(y+1, x+1) - chi(y, x)), a2 = wd2 (chi(y+1, x-1) - chi(y, x));
    // D1T a1 = -a1(y, x) + a1(y-1, x-1), D2T a2 = -a2(y, x) + a2(y-1, x+1)
    const float2 wd0 = P.wd[i];
    const float2 pdr = px_sample(f, chi, d, a, yd, xr).chi, pdl = px_sample(f, chi, d, a, yd, xl).chi;
    const float2 pul = px_sample(f, chi, d, a, yu, xl).chi, pur = px_sample(f, chi, d, a, yu, xr).chi;
    const float wd1_ul = (y > 0 && x > 0) ? P.wd[i - f.w - 1].x : 0.f;
    const float wd2_ur = (y > 0 && x + 1 < f.w) ? P.wd[i - f.w + 1].y : 0.f;
    const float2 a1 = cscale(wd0.x, csub(pdr, p0.chi)), a1u = cscale(wd1_ul, csub(p0.chi, pul));
    const float2 a2 = cscale(wd0.y, csub(pdl, p0.chi)), a2u = cscale(wd2_ur, csub(p0.chi, pur));
    const float2 t1 = make_float2(-a1.x + a1u.x, -a1.y + a1u.y);
    const float2 t2 = make_float2(-a2.x + a2u.x, -a2.y + a2u.y);
    cp.x += 2.f * s.mu_d * (t1.x + t2.x);
    cp.y += 2.f * s.mu_d * (t1.y + t2.y);
  }
  float2 g;
  g.x = -img * p0.car.x + cp.x;
  g.y = img * p0.car.y + cp.y;
  for (int j = 0; j < P.nn; ++j) {
    const float4 nb = P.nb[j][i];
    const float rt = nb.x + (p0.chi.x * nb.y - p0.chi.y * nb.z);
    const float k = 2.f * P.nu * nb.w * rt;
    g.x += k * nb.y;
    g.y -= k * nb.z;
  }
  return g;
}

}  // namespace

// dc = c - c_w and S_w - S per neighbour (:255-268), fixed for the solve.
__global__ void __launch_bounds__(256) k_px_prep(const PxField* __restrict__ F, const int* __restrict__ cp, int nf) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  if (P.nn == 0) return;
  const size_t n = (size_t)P.f.h * P.f.w;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    const int y = (int)(i / P.f.w), x = (int)(i % P.f.w);
    const float2 car = px_carrier(P.f, y, x);
    const float s = P.f.s[i];
    for (int j = 0; j < P.nn; ++j) {
      const float2 cw = P.cw[j][i];
      // the joint solve keeps the neighbour's own composite and carrier
      P.nb[j][i] = P.joint ? make_float4(P.sw[j][i], cw.x, cw.y, 0.f)
                           : make_float4(P.sw[j][i] - s, car.x - cw.x, car.y - cw.y, 0.f);
    }
  }
}

// IRLS weights at chi (:406-516): the 2-D diffusivities (0 on differences that
// leave the plane), the diagonal ones when the oriented prior is on, the
// neighbours' robust weights and the prior scale 1 + nu sum_j wt_j |dc_j|^2.
// The field's first wave also opens the stage's CG state.
__global__ void __launch_bounds__(256) k_px_weights(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                   int cur, int stage) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  if (c.chunk == 0 && threadIdx.x == 0) {
    P.px->done[0] = 0;
    P.px->done[1] = 0;
    if (stage < 16) P.px->iters[stage] = s.n_inner;
  }
  const DevField& f = P.f;
  const float2* chi = f.chi[cur];
  const size_t n = (size_t)f.h * f.w;
  const float e2 = s.eps * s.eps, ec2 = s.eps_c * s.eps_c;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    const int y = (int)(i / f.w), x = (int)(i % f.w);
    const bool hr = x + 1 < f.w, hd = y + 1 < f.h, hl = x > 0;
    const PxSample p0 = px_sample(f, chi, nullptr, 0.f, y, x);
    const PxSample pr = px_sample(f, chi, nullptr, 0.f, y, hr ? x + 1 : x);
    const PxSample pd = px_sample(f, chi, nullptr, 0.f, hd ? y + 1 : y, x);
    const float gx = hr ? pr.yc - p0.yc : 0.f;
    const float gy = hd ? pd.yc - p0.yc : 0.f;
    const float2 dcx = hr ? csub(pr.chi, p0.chi) : make_float2(0.f, 0.f);
    const float2 dcy = hd ? csub(pd.chi, p0.chi) : make_float2(0.f, 0.f);
    const float mx = sqrtf(cnorm(dcx)), my = sqrtf(cnorm(dcy));
    float4 w;
    w.x = hr ? s.eps / sqrtf(gx * gx + e2) : 0.f;
    w.y = hd ? s.eps / sqrtf(gy * gy + e2) : 0.f;
    w.z = hr ? s.eps_c / sqrtf(mx * mx + s.k * gx * gx + ec2) : 0.f;
    w.w = hd ? s.eps_c / sqrtf(my * my + s.k * gy * gy + ec2) : 0.f;
    float2 wd = make_float2(0.f, 0.f);
    if (P.wd) {
      // 1 / sqrt(1 + (|D chi| / eps_c)^2) (:434-451); 0 where D1 / D2 is 0
      if (hd && hr) {
        const float a1 = sqrtf(cnorm(csub(px_sample(f, chi, nullptr, 0.f, y + 1, x + 1).chi, p0.chi))) / s.eps_c;
        wd.x = 1.f / sqrtf(1.f + a1 * a1);
      }
      if (hd && hl) {
        const float a2 = sqrtf(cnorm(csub(px_sample(f, chi, nullptr, 0.f, y + 1, x - 1).chi, p0.chi))) / s.eps_c;
        wd.y = 1.f / sqrtf(1.f + a2 * a2);
      }
    }
    if (P.nn > 0 && P.nu > 0.f) {
      float ks = 1.f;
      for (int j = 0; j < P.nn; ++j) {
        float4 nb = P.nb[j][i];
        const float rt = nb.x + (p0.chi.x * nb.y - p0.chi.y * nb.z);
        nb.w = P.conf[j][i] * P.et2 / (rt * rt + P.et2);
        P.nb[j][i] = nb;
        ks += P.nu * nb.w * (nb.y * nb.y + nb.z * nb.z);
      }
      w.x *= ks;
      w.y *= ks;
      w.z *= ks;
      w.w *= ks;
      wd.x *= ks;
      wd.y *= ks;
    }
    P.wts[i] = w;
    if (P.wd) P.wd[i] = wd;
  }
}

// The stage-start gradient g = grad(chi) and its |g|^2 partials (slot 0).
__global__ void __launch_bounds__(256) k_px_grad0(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                 int cur) {
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  const size_t n = (size_t)P.f.h * P.f.w;
  double gg = 0.0;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    const float2 g = px_grad(P, s, P.f.chi[cur], nullptr, 0.f, (int)(i / P.f.w), (int)(i % P.f.w), nullptr);
    P.g[i] = g;
    gg += (double)g.x * g.x + (double)g.y * g.y;
  }
  px_store(P, c.chunk, 1, gg, 0.0, 0, 0);
}

// First launch of CG iteration `it` (:539-547 of iteration it - 1, then
// :527-531): every wave sums the |g|^2 partials of the gradient at the current
// chi (slot it & 1) -- and, from it = 1 on, of the previous one (slot
// (it - 1) & 1) -- into beta and the tolerance exits; then d = -g + beta d_prev
// (d = -g at it = 0) into d[it & 1] and the partials of curv(d) and Re<g, d>.
__global__ void __launch_bounds__(256) k_px_dir(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                               int it, int stage) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  const bool done_in = P.px->done[0] != 0;
  bool stop = false;
  float beta = 0.f;
  if (!done_in) {
    double gg = 0.0, prev = 0.0;
    px_totals(P, it == 0 ? 1 : 2, it & 1, (it - 1) & 1, &gg, &prev);
    if (it == 0) {
      if (c.chunk == 0 && threadIdx.x == 0) P.px->gg0 = gg;
    } else {
      beta = (float)(gg / (prev > 1e-30 ? prev : 1e-30));
      const double gg0 = P.px->gg0;
      stop = gg < s.tol2 * gg0 || gg < s.gg_floor * s.npix;
    }
  }
  if (c.chunk == 0 && threadIdx.x == 0) {
    P.px->done[1] = (done_in || stop) ? 1 : 0;
    if (!done_in && stop && stage < 16) P.px->iters[stage] = it;
  }
  if (done_in || stop) return;
  const DevField& f = P.f;
  const size_t n = (size_t)f.h * f.w;
  const float2* dp = f.d[(it + 1) & 1];
  float2* dn = f.d[it & 1];
  double hh = 0.0, hd2 = 0.0, gd = 0.0;
  auto dir = [&](size_t j) {
    const float2 g = P.g[j];
    if (it == 0) return make_float2(-g.x, -g.y);
    const float2 q = dp[j];
    return make_float2(-g.x + beta * q.x, -g.y + beta * q.y);
  };
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    const int y = (int)(i / f.w), x = (int)(i % f.w);
    const bool hr = x + 1 < f.w, hdn = y + 1 < f.h;
    const float2 d0 = dir(i);
    dn[i] = d0;
    const float2 c0 = px_carrier(f, y, x);
    const float a0 = -(d0.x * c0.x - d0.y * c0.y);
    float t1 = 0.f, t2 = 0.f;
    float2 cx = make_float2(0.f, 0.f), cy = make_float2(0.f, 0.f);
    if (hr) {
      const float2 dr = dir(i + 1), cr = px_carrier(f, y, x + 1);
      t1 = -(dr.x * cr.x - dr.y * cr.y) - a0;
      cx = csub(dr, d0);
    }
    if (hdn) {
      const float2 dd = dir(i + f.w), cd = px_carrier(f, y + 1, x);
      t2 = -(dd.x * cd.x - dd.y * cd.y) - a0;
      cy = csub(dd, d0);
    }
    const float4 w = P.wts[i];
    hh += (double)(s.lam_y * w.x) * t1 * t1;
    hh += (double)(s.lam_y * w.y) * t2 * t2;
    hh += (double)s.mu_h * w.z * cnorm(cx);
    hh += (double)s.mu_v * w.w * cnorm(cy);
    if (P.wd && hdn) {
      const float2 wd = P.wd[i];
      double a = 0.0;
      if (hr) a += (double)wd.x * cnorm(csub(dir(i + f.w + 1), d0));
      if (x > 0) a += (double)wd.y * cnorm(csub(dir(i + f.w - 1), d0));
      hd2 += a;
    }
    double at = 0.0;
    for (int j = 0; j < P.nn; ++j) {
      const float4 nb = P.nb[j][i];
      const float re = d0.x * nb.y - d0.y * nb.z;
      at += (double)nb.w * re * re;
    }
    hh += (double)P.nu * at;
    const float2 g = P.g[i];
    gd += (double)g.x * d0.x + (double)g.y * d0.y;
  }
  px_store(P, c.chunk, 2, hh + (double)s.mu_d * hd2, gd, 2, 3);
}

// Second launch of CG iteration `it` (:528-538): alpha = -gd / (2 hh) from the
// partials of k_px_dir (or the curvature exit), chi' = chi + alpha d into
// chi[cur ^ ((it + 1) & 1)], g = grad(chi') and its |g|^2 partials into slot
// (it + 1) & 1.  A stopped field copies chi forward instead, so after the
// stage chi sits in chi[cur ^ (n_inner & 1)] whatever happened.
__global__ void __launch_bounds__(256) k_px_step(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                int it, int cur, int stage) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  const DevField& f = P.f;
  const size_t n = (size_t)f.h * f.w;
  const float2* chi_in = f.chi[cur ^ (it & 1)];
  float2* chi_out = f.chi[cur ^ ((it + 1) & 1)];
  bool stopped = P.px->done[1] != 0;
  double hh = 0.0, gd = 0.0;
  if (!stopped) {
    px_totals(P, 2, 2, 3, &hh, &gd);
    if (hh <= s.hh_floor) {
      stopped = true;
      if (c.chunk == 0 && threadIdx.x == 0 && stage < 16) P.px->iters[stage] = it;
    }
  }
  if (c.chunk == 0 && threadIdx.x == 0) P.px->done[0] = stopped ? 1 : 0;
  if (stopped) {
    for (int k = 0; k < kPxPerLane; ++k) {
      const size_t i = px_index(c, k);
      if (i >= n) break;
      chi_out[i] = chi_in[i];
    }
    return;
  }
  const float alpha = (float)(-0.5 * gd / hh);
  const float2* d = f.d[it & 1];
  double gg = 0.0;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    float2 cn;
    const float2 g = px_grad(P, s, chi_in, d, alpha, (int)(i / f.w), (int)(i % f.w), &cn);
    chi_out[i] = cn;
    P.g[i] = g;
    gg += (double)g.x * g.x + (double)g.y * g.y;
  }
  px_store(P, c.chunk, 1, gg, 0.0, (it + 1) & 1, 0);
}

// Y = S - Re[chi c] (:552-555) from chi[cur], and chi copied into the
// caller's buffer chi[0] when it ended in the workspace copy.
__global__ void __launch_bounds__(256) k_px_finalize(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                    int cur) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const DevField& f = P.f;
  const size_t n = (size_t)f.h * f.w;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    const float2 chi = f.chi[cur][i];
    if (cur) f.chi[0][i] = chi;
    if (f.y) {
      const int y = (int)(i / f.w), x = (int)(i % f.w);
      const float2 car = px_carrier(f, y, x);
      f.y[i] = f.s[i] - (chi.x * car.x - chi.y * car.y);
    }
  }
}

// ---------------------------------------------------------------------------
// The joint solve (variational.cpp:558-891, VariationalRefineJoint): Y is a
// free unknown beside chi.  The objective is the exact data term
// (S - Y - Re[chi c])^2, per neighbour nu wt (S_w - Y - Re[chi c_w])^2, the
// anchor nu_a w_a (|Y - Y_a|^2 + |chi - chi_a|^2), Charbonnier priors on
// grad Y and the 2-D chroma priors (the oriented ones too); the weights are
// lagged per stage and scaled by 1 + nu sum wt.  The same launch structure,
// reductions and exit bookkeeping as the chi-only kernels above.
// ---------------------------------------------------------------------------
namespace {

struct PjSample {
  float y;
  float2 chi;
};

__device__ inline PjSample pj_sample(const PxField& P, const float* yv, const float2* chi, const float* dy,
                                     const float2* dc, float a, int y, int x) {
#pragma clang fp contract(off)
  const size_t i = (size_t)y * P.f.w + x;
  PjSample q;
  q.y = yv[i];
  q.chi = chi[i];
  if (dy) {
    q.y += a * dy[i];
    const float2 d = dc[i];
    q.chi.x += a * d.x;
    q.chi.y += a * d.y;
  }
  return q;
}

// (dE/dY, dE/dchi) at pixel (y, x) (:615-683).
__device__ inline void pj_grad(const PxField& P, const FieldCG& s, const float* yv, const float2* chi,
                               const float* dyv, const float2* dcv, float a, int y, int x, float* gy_out,
                               float2* gc_out, PjSample* at) {
#pragma clang fp contract(off)
  const DevField& f = P.f;
  const int xl = max(x - 1, 0), xr = min(x + 1, f.w - 1), yu = max(y - 1, 0), yd = min(y + 1, f.h - 1);
  const PjSample q0 = pj_sample(P, yv, chi, dyv, dcv, a, y, x), ql = pj_sample(P, yv, chi, dyv, dcv, a, y, xl),
                 qr = pj_sample(P, yv, chi, dyv, dcv, a, y, xr), qu = pj_sample(P, yv, chi, dyv, dcv, a, yu, x),
                 qd = pj_sample(P, yv, chi, dyv, dcv, a, yd, x);
  if (at) *at = q0;
  const size_t i = (size_t)y * f.w + x;
  const float2 c0 = px_carrier(f, y, x);
  const float r0 = f.s[i] - q0.y - (q0.chi.x * c0.x - q0.chi.y * c0.y);
  const float fr = -2.f * r0;
  float gy = fr;
  float2 gc = make_float2(fr * c0.x, fr * -c0.y);
  for (int j = 0; j < P.nn; ++j) {
    const float4 nb = P.nb[j][i];  // S_w, c_w, wt
    const float rk = nb.x - q0.y - (q0.chi.x * nb.y - q0.chi.y * nb.z);
    const float fk = -2.f * P.nu * nb.w * rk;
    gy += fk;
    gc.x += fk * nb.y;
    gc.y += fk * -nb.z;
  }
  if (P.a_weight && P.nu_a > 0.f) {
    const float fa = 2.f * P.nu_a * P.a_weight[i];
    const float2 ac = P.a_chroma[i];
    gy += fa * (q0.y - P.a_luma[i]);
    gc.x += fa * (q0.chi.x - ac.x);
    gc.y += fa * (q0.chi.y - ac.y);
  }
  const float4 w0 = P.wts[i];
  const float4 wl = x > 0 ? P.wts[i - 1] : make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 wu = y > 0 ? P.wts[i - f.w] : make_float4(0.f, 0.f, 0.f, 0.f);
  // DxT(wx Dx Y), then DyT(wy Dy Y), each added as hvd-core adds them
  gy += 2.f * (wl.x * (q0.y - ql.y) - w0.x * (qr.y - q0.y));
  gy += 2.f * (wu.y * (q0.y - qu.y) - w0.y * (qd.y - q0.y));
  const float2 cx0 = cscale(w0.z, csub(qr.chi, q0.chi)), cxl = cscale(wl.z, csub(q0.chi, ql.chi));
  const float2 cy0 = cscale(w0.w, csub(qd.chi, q0.chi)), cyu = cscale(wu.w, csub(q0.chi, qu.chi));
  gc.x += 2.f * s.mu_h * (cxl.x - cx0.x);
  gc.y += 2.f * s.mu_h * (cxl.y - cx0.y);
  gc.x += 2.f * s.mu_v * (cyu.x - cy0.x);
  gc.y += 2.f * s.mu_v * (cyu.y - cy0.y);
  if (P.wd) {
    const float2 wd0 = P.wd[i];
    const float2 pdr = pj_sample(P, yv, chi, dyv, dcv, a, yd, xr).chi, pdl = pj_sample(P, yv, chi, dyv, dcv, a, yd, xl).chi;
    const float2 pul = pj_sample(P, yv, chi, dyv, dcv, a, yu, xl).chi, pur = pj_sample(P, yv, chi, dyv, dcv, a, yu, xr).chi;
    const float wd1_ul = (y > 0 && x > 0) ? P.wd[i - f.w - 1].x : 0.f;
    const float wd2_ur = (y > 0 && x + 1 < f.w) ? P.wd[i - f.w + 1].y : 0.f;
    const float2 a1 = cscale(wd0.x, csub(pdr, q0.chi)), a1u = cscale(wd1_ul, csub(q0.chi, pul));
    const float2 a2 = cscale(wd0.y, csub(pdl, q0.chi)), a2u = cscale(wd2_ur, csub(q0.chi, pur));
    gc.x += 2.f * s.mu_d * (-a1.x + a1u.x);
    gc.y += 2.f * s.mu_d * (-a1.y + a1u.y);
    gc.x += 2.f * s.mu_d * (-a2.x + a2u.x);
    gc.y += 2.f * s.mu_d * (-a2.y + a2u.y);
  }
  *gy_out = gy;
  *gc_out = gc;
}

}  // namespace

// The joint solve's stage weights (:773-858): Charbonnier on grad Y, the
// coupled chroma diffusivities, the diagonal ones, the neighbours' robust
// weights on S_w - Y - Re[chi c_w] and the prior scale 1 + nu sum wt.
__global__ void __launch_bounds__(256) k_pj_weights(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                   int cur, int stage) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  if (c.chunk == 0 && threadIdx.x == 0) {
    P.px->done[0] = 0;
    P.px->done[1] = 0;
    if (stage < 16) P.px->iters[stage] = s.n_inner;
  }
  const DevField& f = P.f;
  const float2* chi = f.chi[cur];
  const float* yv = P.yv[cur];
  const size_t n = (size_t)f.h * f.w;
  const float e2 = s.eps * s.eps, ec2 = s.eps_c * s.eps_c;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    const int y = (int)(i / f.w), x = (int)(i % f.w);
    const bool hr = x + 1 < f.w, hd = y + 1 < f.h, hl = x > 0;
    const float y0 = yv[i];
    const float2 q0 = chi[i];
    const float gx = hr ? yv[i + 1] - y0 : 0.f;
    const float gyv = hd ? yv[i + f.w] - y0 : 0.f;
    const float2 dcx = hr ? csub(chi[i + 1], q0) : make_float2(0.f, 0.f);
    const float2 dcy = hd ? csub(chi[i + f.w], q0) : make_float2(0.f, 0.f);
    const float mx = sqrtf(cnorm(dcx)), my = sqrtf(cnorm(dcy));
    float4 w;
    w.x = hr ? s.eps / sqrtf(gx * gx + e2) : 0.f;
    w.y = hd ? s.eps / sqrtf(gyv * gyv + e2) : 0.f;
    w.z = hr ? s.eps_c / sqrtf(mx * mx + s.k * gx * gx + ec2) : 0.f;
    w.w = hd ? s.eps_c / sqrtf(my * my + s.k * gyv * gyv + ec2) : 0.f;
    float2 wd = make_float2(0.f, 0.f);
    if (P.wd) {
      if (hd && hr) {
        const float a1 = sqrtf(cnorm(csub(chi[i + f.w + 1], q0))) / s.eps_c;
        wd.x = 1.f / sqrtf(1.f + a1 * a1);
      }
      if (hd && hl) {
        const float a2 = sqrtf(cnorm(csub(chi[i + f.w - 1], q0))) / s.eps_c;
        wd.y = 1.f / sqrtf(1.f + a2 * a2);
      }
    }
    if (P.nn > 0 && P.nu > 0.f) {
      float ks = 1.f;
      for (int j = 0; j < P.nn; ++j) {
        float4 nb = P.nb[j][i];
        const float rk = nb.x - y0 - (q0.x * nb.y - q0.y * nb.z);
        nb.w = P.conf[j][i] * P.et2 / (rk * rk + P.et2);
        P.nb[j][i] = nb;
        ks += P.nu * nb.w;
      }
      w.x *= ks;
      w.y *= ks;
      w.z *= ks;
      w.w *= ks;
      wd.x *= ks;
      wd.y *= ks;
    }
    P.wts[i] = w;
    if (P.wd) P.wd[i] = wd;
  }
}

__global__ void __launch_bounds__(256) k_pj_grad0(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                 int cur) {
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  const size_t n = (size_t)P.f.h * P.f.w;
  double gg = 0.0;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    float gy;
    float2 gc;
    pj_grad(P, s, P.yv[cur], P.f.chi[cur], nullptr, nullptr, 0.f, (int)(i / P.f.w), (int)(i % P.f.w), &gy, &gc,
            nullptr);
    P.gy[i] = gy;
    P.g[i] = gc;
    gg += (double)gy * gy + ((double)gc.x * gc.x + (double)gc.y * gc.y);
  }
  px_store(P, c.chunk, 1, gg, 0.0, 0, 0);
}

// First launch of joint iteration `it`: beta and the tolerance exits, the
// direction (dY, dC) = -(gY, gC) + beta (dY, dC)_prev, and the partials of
// its curvature (:687-755) and of <g, d>.
__global__ void __launch_bounds__(256) k_pj_dir(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                               int it, int stage) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  const bool done_in = P.px->done[0] != 0;
  bool stop = false;
  float beta = 0.f;
  if (!done_in) {
    double gg = 0.0, prev = 0.0;
    px_totals(P, it == 0 ? 1 : 2, it & 1, (it - 1) & 1, &gg, &prev);
    if (it == 0) {
      if (c.chunk == 0 && threadIdx.x == 0) P.px->gg0 = gg;
    } else {
      beta = (float)(gg / (prev > 1e-30 ? prev : 1e-30));
      const double gg0 = P.px->gg0;
      stop = gg < s.tol2 * gg0 || gg < s.gg_floor * s.npix;
    }
  }
  if (c.chunk == 0 && threadIdx.x == 0) {
    P.px->done[1] = (done_in || stop) ? 1 : 0;
    if (!done_in && stop && stage < 16) P.px->iters[stage] = it;
  }
  if (done_in || stop) return;
  const DevField& f = P.f;
  const size_t n = (size_t)f.h * f.w;
  const float2* dcp = f.d[(it + 1) & 1];
  const float* dyp = P.dyv[(it + 1) & 1];
  float2* dcn = f.d[it & 1];
  float* dyn = P.dyv[it & 1];
  auto dir_c = [&](size_t j) {
    const float2 g = P.g[j];
    if (it == 0) return make_float2(-g.x, -g.y);
    const float2 q = dcp[j];
    return make_float2(-g.x + beta * q.x, -g.y + beta * q.y);
  };
  auto dir_y = [&](size_t j) {
    const float g = P.gy[j];
    if (it == 0) return -g;
    return -g + beta * dyp[j];
  };
  double hh = 0.0, hsp = 0.0, hd2 = 0.0, gd = 0.0;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    const int y = (int)(i / f.w), x = (int)(i % f.w);
    const bool hr = x + 1 < f.w, hdn = y + 1 < f.h;
    const float dY = dir_y(i);
    const float2 dC = dir_c(i);
    dyn[i] = dY;
    dcn[i] = dC;
    const float2 c0 = px_carrier(f, y, x);
    const float v = dY + (dC.x * c0.x - dC.y * c0.y);
    hh += (double)v * v;
    for (int j = 0; j < P.nn; ++j) {
      const float4 nb = P.nb[j][i];
      const float dk = dY + (dC.x * nb.y - dC.y * nb.z);
      hh += (double)P.nu * ((double)nb.w * dk * dk);
    }
    if (P.a_weight && P.nu_a > 0.f)
      hh += (double)P.nu_a * ((double)P.a_weight[i] * ((double)dY * dY + (double)cnorm(dC)));
    float t1 = 0.f, t2 = 0.f;
    float2 cx = make_float2(0.f, 0.f), cy = make_float2(0.f, 0.f);
    if (hr) {
      t1 = dir_y(i + 1) - dY;
      cx = csub(dir_c(i + 1), dC);
    }
    if (hdn) {
      t2 = dir_y(i + f.w) - dY;
      cy = csub(dir_c(i + f.w), dC);
    }
    const float4 w = P.wts[i];
    hsp += (double)w.x * t1 * t1;
    hsp += (double)w.y * t2 * t2;
    hsp += (double)s.mu_h * w.z * cnorm(cx);
    hsp += (double)s.mu_v * w.w * cnorm(cy);
    if (P.wd && hdn) {
      const float2 wd = P.wd[i];
      if (hr) hd2 += (double)wd.x * cnorm(csub(dir_c(i + f.w + 1), dC));
      if (x > 0) hd2 += (double)wd.y * cnorm(csub(dir_c(i + f.w - 1), dC));
    }
    const float gY = P.gy[i];
    const float2 gC = P.g[i];
    gd += (double)gY * dY + ((double)gC.x * dC.x + (double)gC.y * dC.y);
  }
  px_store(P, c.chunk, 2, hh + hsp + (double)s.mu_d * hd2, gd, 2, 3);
}

// Second launch of joint iteration `it`: alpha or the curvature exit, the
// step (Y, chi) += alpha (dY, dC) into the other buffers, the gradient there.
__global__ void __launch_bounds__(256) k_pj_step(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                int it, int cur, int stage) {
#pragma clang fp contract(off)
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  const FieldCG s = *P.st;
  const DevField& f = P.f;
  const size_t n = (size_t)f.h * f.w;
  const int src = cur ^ (it & 1), dst = cur ^ ((it + 1) & 1);
  bool stopped = P.px->done[1] != 0;
  double hh = 0.0, gd = 0.0;
  if (!stopped) {
    px_totals(P, 2, 2, 3, &hh, &gd);
    if (hh <= s.hh_floor) {
      stopped = true;
      if (c.chunk == 0 && threadIdx.x == 0 && stage < 16) P.px->iters[stage] = it;
    }
  }
  if (c.chunk == 0 && threadIdx.x == 0) P.px->done[0] = stopped ? 1 : 0;
  if (stopped) {
    for (int k = 0; k < kPxPerLane; ++k) {
      const size_t i = px_index(c, k);
      if (i >= n) break;
      f.chi[dst][i] = f.chi[src][i];
      P.yv[dst][i] = P.yv[src][i];
    }
    return;
  }
  const float alpha = (float)(-0.5 * gd / hh);
  double gg = 0.0;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    float gy;
    float2 gc;
    PjSample q;
    pj_grad(P, s, P.yv[src], f.chi[src], P.dyv[it & 1], f.d[it & 1], alpha, (int)(i / f.w), (int)(i % f.w), &gy,
            &gc, &q);
    P.yv[dst][i] = q.y;
    f.chi[dst][i] = q.chi;
    P.gy[i] = gy;
    P.g[i] = gc;
    gg += (double)gy * gy + ((double)gc.x * gc.x + (double)gc.y * gc.y);
  }
  px_store(P, c.chunk, 1, gg, 0.0, (it + 1) & 1, 0);
}

// (Y, chi) into the caller's buffers when they ended in the workspace copy.
__global__ void __launch_bounds__(256) k_pj_finalize(const PxField* __restrict__ F, const int* __restrict__ cp, int nf,
                                                    int cur) {
  const PxCtx c = px_ctx(cp, nf);
  const PxField& P = F[c.fi];
  if (!cur) return;
  const size_t n = (size_t)P.f.h * P.f.w;
  for (int k = 0; k < kPxPerLane; ++k) {
    const size_t i = px_index(c, k);
    if (i >= n) break;
    P.f.chi[0][i] = P.f.chi[1][i];
    P.yv[0][i] = P.yv[1][i];
  }
}

}  // namespace hvdk
