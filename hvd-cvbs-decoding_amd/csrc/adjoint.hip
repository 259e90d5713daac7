// adjoint.hip -- the adjoint of the 2-D IRLS + CG solve (no hvd-core
// counterpart: the differentiation of VariationalRefine the training path
// needs).  Each IRLS stage k (weights w_k = W(chi_{k-1}, S)) is taken as
// solved exactly, r_k(chi_k) = dE_k / dchi = 0; from gbar = dL/dchi_k the
// host runs an adjoint CG a = (d2E_k)^-1 gbar with the forward's own
// kernels (S = 0, right-hand side gbar), then:
//   k_adj_local      the lambda partials <a, g_data(chi_k)>, <a, g_prior(chi_k)>,
//                    dL/dS += 2 lam_y L_w Re[a c], and per pixel the weight
//                    adjoint chained through W at chi_{k-1};
//   k_adj_transpose  the D^T of those per-pixel values: dL/dS and
//                    gbar <- dL/dchi_{k-1}.
// One thread per pixel; these run once per stage, not per CG iteration.
#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

namespace {

__device__ inline float2 adj_carrier(const AdjField& f, int y, int x) {
  if (f.carrier) return gld(f.carrier, (size_t)y * f.w + x);
  float2 v = cmul(gld(f.rowph, y), gld(f.colph, x));
  if (f.vsw && gld(f.vsw, y) < 0) v.x = -v.x;
  return v;
}

__device__ inline float dot2(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
__device__ inline float2 sub2(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

}  // namespace

// dL/dchi_K = gchi + dL/dY (-c_re, c_im), dL/dS = dL/dY (Y = S - Re[chi c]).
// Grid: (ceil(max px / 256), nfields).
__global__ void k_adj_seed(const AdjField* A) {
  const AdjField& f = A[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.h * f.w) return;
  const int y = i / f.w, x = i - y * f.w;
  const float yb = f.gy ? gld(f.gy, i) : 0.f;
  const float2 c = adj_carrier(f, y, x);
  const float2 g = gld(f.gchi, i);
  gst(f.gbar, i, make_float2(g.x - yb * c.x, g.y + yb * c.y));
  gst(f.sbar, i, yb);
}

// Per pixel after stage k's adjoint solve.  partials: (field, block) pairs of
// <a, g_data>, <a, g_prior>.  Grid: (nbx, nfields), 256 threads.
__global__ void __launch_bounds__(256) k_adj_local(const AdjField* A, const FieldCG* st, int nbx, double* partials) {
  const int fi = blockIdx.y;
  const AdjField f = A[fi];
  const FieldCG s = st[fi];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double dd = 0.0, dp = 0.0;
  if (i < f.h * f.w) {
    const int y = i / f.w, x = i - y * f.w, w = f.w;
    const bool hl = x > 0, hr = x + 1 < w, hu = y > 0, hd = y + 1 < f.h;
    // values at p and its four neighbours (zero where absent; never used there)
    auto at = [&](int dy, int dx, bool ok, float2& a, float2& ck, float& ycv, float& qv) {
      if (!ok) {
        a = ck = make_float2(0.f, 0.f);
        ycv = qv = 0.f;
        return;
      }
      const size_t j = (size_t)(y + dy) * w + (x + dx);
      const float2 c = adj_carrier(f, y + dy, x + dx);
      a = gld(f.abar, j);
      ck = gld(f.chik, j);
      ycv = gld(f.s, j) - re_mul(ck, c);
      qv = re_mul(a, c);
    };
    float2 a0, al, ar, au, ad, k0, kl, kr, ku, kd;
    float y0, yl, yr, yu, ydn, q0, ql, qr, qu, qd;
    at(0, 0, true, a0, k0, y0, q0);
    at(0, -1, hl, al, kl, yl, ql);
    at(0, 1, hr, ar, kr, yr, qr);
    at(-1, 0, hu, au, ku, yu, qu);
    at(1, 0, hd, ad, kd, ydn, qd);
    const float4 w0 = gld(f.wts, (size_t)i);
    const float4 wl = hl ? gld(f.wts, (size_t)i - 1) : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 wu = hu ? gld(f.wts, (size_t)i - w) : make_float4(0.f, 0.f, 0.f, 0.f);
    const float2 c0 = adj_carrier(f, y, x);
    // gradient parts at chi_k (the forward's grad, variational.cpp:286-335)
    const float gx0 = hr ? yr - y0 : 0.f, gxl = hl ? y0 - yl : 0.f;
    const float gy0 = hd ? ydn - y0 : 0.f, gyu = hu ? y0 - yu : 0.f;
    const float img = 2.f * s.lam_y * ((wl.x * gxl - w0.x * gx0) + (wu.y * gyu - w0.y * gy0));
    const float2 gdata = make_float2(-img * c0.x, img * c0.y);
    const float2 cx0 = hr ? sub2(kr, k0) : make_float2(0.f, 0.f), cxl = hl ? sub2(k0, kl) : make_float2(0.f, 0.f);
    const float2 cy0 = hd ? sub2(kd, k0) : make_float2(0.f, 0.f), cyu = hu ? sub2(k0, ku) : make_float2(0.f, 0.f);
    const float2 gprior = make_float2(
        2.f * (s.mu_h * (wl.z * cxl.x - w0.z * cx0.x) + s.mu_v * (wu.w * cyu.x - w0.w * cy0.x)),
        2.f * (s.mu_h * (wl.z * cxl.y - w0.z * cx0.y) + s.mu_v * (wu.w * cyu.y - w0.w * cy0.y)));
    dd = (double)a0.x * gdata.x + (double)a0.y * gdata.y;
    dp = (double)a0.x * gprior.x + (double)a0.y * gprior.y;
    // dL/dS from the data term: 2 lam_y L_w q, q = Re[a c]
    const float qx0 = hr ? qr - q0 : 0.f, qxl = hl ? q0 - ql : 0.f;
    const float qy0 = hd ? qd - q0 : 0.f, qyu = hu ? q0 - qu : 0.f;
    const float sb = 2.f * s.lam_y * ((wl.x * qxl - w0.x * qx0) + (wu.y * qyu - w0.y * qy0));
    gst(f.sbar, (size_t)i, gld(f.sbar, (size_t)i) + sb);
    // weight adjoint at p: wbar = -d<a, r(chi_k; w)>/dw
    const float2 ax0 = hr ? sub2(ar, a0) : make_float2(0.f, 0.f);
    const float2 ay0 = hd ? sub2(ad, a0) : make_float2(0.f, 0.f);
    const float wxb = 2.f * s.lam_y * gx0 * qx0;
    const float wyb = 2.f * s.lam_y * gy0 * qy0;
    const float wcxb = -2.f * s.mu_h * dot2(cx0, ax0);
    const float wcyb = -2.f * s.mu_v * dot2(cy0, ay0);
    // ... through W at chi_{k-1} (variational.cpp:406-433)
    const float2 cpp = gld(f.chip, (size_t)i);
    const float yp0 = gld(f.s, (size_t)i) - re_mul(cpp, c0);
    float gxp = 0.f, gyp = 0.f;
    float2 dxp = make_float2(0.f, 0.f), dyp = dxp;
    if (hr) {
      const float2 c = adj_carrier(f, y, x + 1);
      const float2 cr = gld(f.chip, (size_t)i + 1);
      gxp = (gld(f.s, (size_t)i + 1) - re_mul(cr, c)) - yp0;
      dxp = sub2(cr, cpp);
    }
    if (hd) {
      const float2 c = adj_carrier(f, y + 1, x);
      const float2 cd = gld(f.chip, (size_t)i + w);
      gyp = (gld(f.s, (size_t)i + w) - re_mul(cd, c)) - yp0;
      dyp = sub2(cd, cpp);
    }
    const float e2 = s.eps * s.eps, ec2 = s.eps_c * s.eps_c;
    const float ux = gxp * gxp + e2, uy = gyp * gyp + e2;
    const float vx = dot2(dxp, dxp) + s.k * gxp * gxp + ec2, vy = dot2(dyp, dyp) + s.k * gyp * gyp + ec2;
    const float dwx = -s.eps * gxp / (ux * sqrtf(ux)), dwy = -s.eps * gyp / (uy * sqrtf(uy));
    const float fx = -s.eps_c / (vx * sqrtf(vx)), fy = -s.eps_c / (vy * sqrtf(vy));
    float* sc = f.scratch + (size_t)6 * i;
    sc[0] = wxb * dwx + wcxb * fx * s.k * gxp;
    sc[1] = wyb * dwy + wcyb * fy * s.k * gyp;
    sc[2] = wcxb * fx * dxp.x;
    sc[3] = wcxb * fx * dxp.y;
    sc[4] = wcyb * fy * dyp.x;
    sc[5] = wcyb * fy * dyp.y;
  }
  // deterministic block sums
  __shared__ double lds[4];
  const double bd = block_sum<256>(dd, lds);
  const double bp = block_sum<256>(dp, lds);
  if (threadIdx.x == 0) {
    partials[((size_t)fi * nbx + blockIdx.x) * 2] = bd;
    partials[((size_t)fi * nbx + blockIdx.x) * 2 + 1] = bp;
  }
}

// Per-field sums of the lambda partials, in block order:
// dL/dlam_y -= <a, g_data> / lam_y, dL/dlam_c -= <a, g_prior> / lam_c.
// One block of 256 per field.
__global__ void __launch_bounds__(256) k_adj_reduce(const FieldCG* st, int nbx, const double* partials, double* glam) {
  const int fi = blockIdx.x;
  __shared__ double lds[4];
  double a = 0.0, b = 0.0;
  for (int j = threadIdx.x; j < nbx; j += blockDim.x) {
    a += partials[((size_t)fi * nbx + j) * 2];
    b += partials[((size_t)fi * nbx + j) * 2 + 1];
  }
  a = block_sum<256>(a, lds);
  b = block_sum<256>(b, lds);
  if (threadIdx.x == 0) {
    glam[2 * fi] -= a / (double)st[fi].lam_y;
    glam[2 * fi + 1] -= b / (double)st[fi].mu_v;  // mu_v = lambda_c
  }
}

// The transposed differences of the per-pixel values of k_adj_local:
// ycbar = Dx^T gx' + Dy^T gy', dL/dS += ycbar, and
// gbar <- dL/dchi_{k-1} = ycbar (-c_re, c_im) + Dx^T (Dx chi)bar + Dy^T (Dy chi)bar.
// Grid: (ceil(max px / 256), nfields).
__global__ void k_adj_transpose(const AdjField* A) {
  const AdjField& f = A[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.h * f.w) return;
  const int y = i / f.w, x = i - y * f.w;
  const float* sc = f.scratch + (size_t)6 * i;
  const float* scl = x > 0 ? sc - 6 : nullptr;
  const float* scu = y > 0 ? sc - (size_t)6 * f.w : nullptr;
  const float ycb = (scl ? scl[0] : 0.f) - sc[0] + (scu ? scu[1] : 0.f) - sc[1];
  const float2 dxt = make_float2((scl ? scl[2] : 0.f) - sc[2], (scl ? scl[3] : 0.f) - sc[3]);
  const float2 dyt = make_float2((scu ? scu[4] : 0.f) - sc[4], (scu ? scu[5] : 0.f) - sc[5]);
  gst(f.sbar, (size_t)i, gld(f.sbar, (size_t)i) + ycb);
  const float2 c = adj_carrier(f, y, x);
  gst(f.gbar, (size_t)i, make_float2(-ycb * c.x + dxt.x + dyt.x, ycb * c.y + dxt.y + dyt.y));
}

// Zero fill (float2 planes, n pixels) for the adjoint's starting point.
__global__ void k_zero2(float2* a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    gst(a, i, make_float2(0.f, 0.f));
}

// glam accumulators to zero.
__global__ void k_zero_d(double* a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

}  // namespace hvdk
