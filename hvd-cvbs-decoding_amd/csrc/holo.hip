// holo.hip -- holographic chroma init (engine/holographic_init.cpp:200-317)
// with project-written exact-size FFTs.
//
// The 2-D transforms keep hvd-core's sizes exactly (numpy semantics of
// engine/fft2d.h:30-33, no zero padding to friendly sizes -- that would
// change the circular convolution): NTSC 274 x 760, PAL 320 x 922.  Each 1-D
// transform runs in LDS as a mixed-radix Stockham autosort pass sequence in
// which every output of a radix-R stage is a direct R-term sum
//     y[o] = sum_r x[ib + r M] W_n^{r E(o)},
// so awkward primes (19, 137, 461) are just long stages; every twiddle is a
// lookup into one table of n-th roots computed in FP64 on the host.
//
// Pipeline per field (5 launches for the whole batch):
//   k_fft_rows_fwd  demod S conj(c) + reflect101 row padding + row FFT
//   k_fft_cols_fwd  column FFT in place (8 columns per block)
//   k_fft_cols_inv  per variant: spectrum x Gaussian crop, inverse column
//                   FFT, only the rows that survive the crop are written
//   k_fft_rows_inv  inverse row FFT, x 2 / (hp w)
//   k_holo_blend    residual luma, edge energy (7x7 box in LDS), weighted
//                   blend, Y0 = S - Re[chi0 c]
#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

namespace {

__device__ inline float2 carrier_of(const InitField& f, int y, int x) {
  if (f.carrier) return f.carrier[(size_t)y * f.w + x];
  float2 v = cmul(f.rowph[y], f.colph[x]);
  if (f.vsw && f.vsw[y] < 0) v.x = -v.x;
  return v;
}

// reflect101 (holographic_init.cpp:126-138).
__device__ inline int mirror_row(int y, int h) {
  if (h <= 1) return 0;
  const int period = 2 * h - 2;
  y = ((y % period) + period) % period;
  return y < h ? y : period - y;
}

// In-LDS transform of `ncols` interleaved sequences of length n:
// element (j, c) lives at buf[j * ncols + c].  Returns the buffer holding the
// result (a or b).  sign < 0 forward (e^{-2 pi i/n}), > 0 inverse, unscaled.
//
// A stage of odd radix R >= 11 runs as a length-R DFT of pre-twiddled inputs
// a_r = x_r W^{r k step1} (k the output group's phase), using both of its
// symmetries: the terms r and R - r pair as u_r = a_r + a_{R-r},
// v_r = a_r - a_{R-r}, and the outputs q and R - q share
// U = sum_r u_r cos(2 pi r q / R), V = sum_r v_r sin(2 pi r q / R):
//     y_q = a_0 + U + i V,   y_{R-q} = a_0 + U - i V.
// A first pass forms [a_0, u_1, v_1, ..., u_h, v_h] per group (h = (R - 1) / 2)
// into the spare buffer, a second forms two outputs per work item from it: a
// quarter of the multiply-adds of the direct R-term sums.
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// x * (si i), si = +-1
__device__ inline float2 mul_si(float2 x, float si) { return make_float2(-si * x.y, si * x.x); }

// y = DFT_4(a) with W_4 = si i (si = -1 forward, +1 inverse).
__device__ inline void dft4(float2 a0, float2 a1, float2 a2, float2 a3, float si, float2& y0, float2& y1, float2& y2,
                            float2& y3) {
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_si(csub(a1, a3), si);
  y0 = cadd(t0, t2);
  y2 = csub(t0, t2);
  y1 = cadd(t1, t3);
  y3 = csub(t1, t3);
}

// One Stockham stage of small radix R (2, 3, 4, 5, 7, 8) with one work item
// per output group: the group's R inputs are pre-twiddled by W^{r k step1}
// and transformed by an R-point DFT written out (butterflies for 2, 4, 8; the
// symmetric pair form for 3, 5, 7, whose constants come from the root table).
template <int R, typename Tw>
__device__ inline void small_stage(const float2* src, float2* dst, int n, int p, int M, int step1, float si, int ncols,
                                   Tw twiddle) {
  for (int e = threadIdx.x; e < M * ncols; e += blockDim.x) {
    const int c = e % ncols, ib = e / ncols;
    const int k = ib % p, j = ib / p;
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = src[(ib + r * M) * ncols + c];
      if (r > 0 && k != 0) {
        const float2 w = twiddle(r * k * step1);  // < R (p / p) (n / R) = n
        x = make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
      }
      a[r] = x;
    }
    float2 y[R];
    if constexpr (R == 2) {
      y[0] = cadd(a[0], a[1]);
      y[1] = csub(a[0], a[1]);
    } else if constexpr (R == 4) {
      dft4(a[0], a[1], a[2], a[3], si, y[0], y[1], y[2], y[3]);
    } else if constexpr (R == 8) {
      float2 e0, e1, e2, e3, o0, o1, o2, o3;
      dft4(a[0], a[2], a[4], a[6], si, e0, e1, e2, e3);
      dft4(a[1], a[3], a[5], a[7], si, o0, o1, o2, o3);
      const float h2 = 0.70710678118654752f;
      // W_8^1 = (h2, si h2), W_8^2 = si i, W_8^3 = (-h2, si h2)
      const float2 w1 = make_float2(h2 * (o1.x - si * o1.y), h2 * (o1.y + si * o1.x));
      const float2 w2 = mul_si(o2, si);
      const float2 w3 = make_float2(h2 * (-o3.x - si * o3.y), h2 * (-o3.y + si * o3.x));
      y[0] = cadd(e0, o0);
      y[4] = csub(e0, o0);
      y[1] = cadd(e1, w1);
      y[5] = csub(e1, w1);
      y[2] = cadd(e2, w2);
      y[6] = csub(e2, w2);
      y[3] = cadd(e3, w3);
      y[7] = csub(e3, w3);
    } else {  // odd R: pairs r / R - r and outputs q / R - q
      constexpr int h = (R - 1) / 2;
      float2 u[h + 1], v[h + 1];
      float2 y0 = a[0];
#pragma unroll
      for (int r = 1; r <= h; ++r) {
        u[r] = cadd(a[r], a[R - r]);
        v[r] = csub(a[r], a[R - r]);
        y0 = cadd(y0, u[r]);
      }
      y[0] = y0;
#pragma unroll
      for (int q = 1; q <= h; ++q) {
        float ux = 0.f, uy = 0.f, vx = 0.f, vy = 0.f;
#pragma unroll
        for (int r = 1; r <= h; ++r) {
          const float2 w = twiddle((r * q) % R * M);
          ux = __fmaf_rn(u[r].x, w.x, ux);
          uy = __fmaf_rn(u[r].y, w.x, uy);
          vx = __fmaf_rn(v[r].x, w.y, vx);
          vy = __fmaf_rn(v[r].y, w.y, vy);
        }
        y[q] = make_float2(a[0].x + ux - vy, a[0].y + uy + vx);
        y[R - q] = make_float2(a[0].x + ux + vy, a[0].y + uy - vx);
      }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) dst[((j * R + q) * p + k) * ncols + c] = y[q];
  }
}

__device__ __attribute__((always_inline)) float2* fft_lds(float2* a, float2* b, const float2* root, const FftPlan& P,
                                                          int sign, int ncols) {
  const int n = P.n;
  float2* src = a;
  float2* dst = b;
  int p = 1;
  auto twiddle = [&](int ex) {
    const float2 w = root[ex];
    return make_float2(w.x, sign < 0 ? w.y : -w.y);
  };
  for (int s = 0; s < P.nst; ++s) {
    const int R = P.radix[s];
    const int M = n / R;
    const int step1 = n / (p * R);
    if (R >= 11 && (R & 1)) {
      const int h = (R - 1) / 2;
      // pass 1: group ib (0 <= ib < M) -> dst[ib R ..] = a_0, u_1, v_1, ...
      for (int e = threadIdx.x; e < M * (h + 1) * ncols; e += blockDim.x) {
        const int c = e % ncols, g = e / ncols;
        const int ib = g / (h + 1), r = g % (h + 1);
        const int k = ib % p;
        auto pre = [&](int rr) {  // a_rr = x_rr W^{rr k step1}
          const float2 x = src[(ib + rr * M) * ncols + c];
          if (k == 0) return x;
          const float2 w = twiddle(rr * k * step1);  // k step1 < n / R
          return make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
        };
        if (r == 0) {
          dst[(ib * R) * ncols + c] = pre(0);
          continue;
        }
        const float2 x = pre(r), y = pre(R - r);
        dst[(ib * R + 2 * r - 1) * ncols + c] = make_float2(x.x + y.x, x.y + y.y);
        dst[(ib * R + 2 * r) * ncols + c] = make_float2(x.x - y.x, x.y - y.y);
      }
      __syncthreads();
      // pass 2: outputs q and R - q of group ib into src, kQ values of q per
      // work item so that each u_r, v_r read from LDS feeds 4 kQ products
      constexpr int kQ = 4;
      const int nq = (h + kQ) / kQ;  // q = 0 .. h in blocks of kQ
      for (int e = threadIdx.x; e < M * nq * ncols; e += blockDim.x) {
        const int c = e % ncols, g = e / ncols;
        const int ib = g / nq, q0 = (g % nq) * kQ;
        const int k = ib % p, j = ib / p;
        const float2* gp = dst + (ib * R) * ncols + c;
        const float2 a0 = gp[0];
        float ux[kQ], uy[kQ], vx[kQ], vy[kQ];
        int E[kQ], ex[kQ];
#pragma unroll
        for (int t = 0; t < kQ; ++t) {
          ux[t] = uy[t] = vx[t] = vy[t] = 0.f;
          E[t] = min(q0 + t, h) * M;  // < n (q beyond h repeats h, not stored)
          ex[t] = E[t];
        }
        for (int r = 1; r <= h; ++r) {
          const float2 u = gp[(2 * r - 1) * ncols], v = gp[(2 * r) * ncols];
#pragma unroll
          for (int t = 0; t < kQ; ++t) {
            const float2 w = twiddle(ex[t]);
            ux[t] = __fmaf_rn(u.x, w.x, ux[t]);
            uy[t] = __fmaf_rn(u.y, w.x, uy[t]);
            vx[t] = __fmaf_rn(v.x, w.y, vx[t]);
            vy[t] = __fmaf_rn(v.y, w.y, vy[t]);
            ex[t] += E[t];
            if (ex[t] >= n) ex[t] -= n;
          }
        }
#pragma unroll
        for (int t = 0; t < kQ; ++t) {
          const int q = q0 + t;
          if (q > h) break;
          src[((j * R + q) * p + k) * ncols + c] = make_float2(a0.x + ux[t] - vy[t], a0.y + uy[t] + vx[t]);
          if (q > 0)
            src[((j * R + R - q) * p + k) * ncols + c] = make_float2(a0.x + ux[t] + vy[t], a0.y + uy[t] - vx[t]);
        }
      }
      __syncthreads();
    } else if (R <= 8 && R != 6) {
      const float si = sign < 0 ? -1.f : 1.f;
      switch (R) {
        case 2: small_stage<2>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
        case 3: small_stage<3>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
        case 4: small_stage<4>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
        case 5: small_stage<5>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
        case 7: small_stage<7>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
        default: small_stage<8>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
      }
      __syncthreads();
      float2* t2 = src;
      src = dst;
      dst = t2;
    } else {
      for (int e = threadIdx.x; e < n * ncols; e += blockDim.x) {
        const int c = e % ncols, o = e / ncols;
        const int k = o % p, t = o / p;
        const int q = t % R, ib = (t / R) * p + k;
        int E = k * step1 + q * M;
        if (E >= n) E -= n;
        float ar = 0.f, ai = 0.f;
        int ex = 0;
        for (int r = 0; r < R; ++r) {
          const float2 x = src[(ib + r * M) * ncols + c];
          const float2 w = twiddle(ex);
          ar = __fmaf_rn(x.x, w.x, __fmaf_rn(-x.y, w.y, ar));
          ai = __fmaf_rn(x.x, w.y, __fmaf_rn(x.y, w.x, ai));
          ex += E;
          if (ex >= n) ex -= n;
        }
        dst[o * ncols + c] = make_float2(ar, ai);
      }
      __syncthreads();
      float2* t2 = src;
      src = dst;
      dst = t2;
    }
    p *= R;
  }
  return src;
}

__device__ inline void load_root(float2* lds, const float2* root, int n) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) lds[j] = root[j];
}

// numpy fftfreq(n, d) bin i (holographic_init.cpp:16-25), cycles per unit.
__device__ inline float fftfreq(int i, int n, float inv_nd) {
  const int half = (n % 2 == 0) ? n / 2 : (n + 1) / 2;
  const int k = i < half ? i : i - n;
  return (float)k * inv_nd;
}

}  // namespace

// (1) demod + reflect101 padding + forward row FFT.  One block per padded row.
__global__ void __launch_bounds__(256) k_fft_rows_fwd(const InitField* F, const FftPlan* plans, const float2* roots,
                                                      const int* row_prefix, int nf) {
  extern __shared__ float2 lds[];
  const int fi = field_of_tile(row_prefix, nf, blockIdx.x);
  const InitField& f = F[fi];
  const int yp = blockIdx.x - row_prefix[fi];
  const FftPlan& P = plans[f.row_plan];
  float2* a = lds;
  float2* b = lds + P.n;
  float2* rt = lds + 2 * P.n;
  load_root(rt, roots + P.root_off, P.n);
  const int sy = mirror_row(yp - f.pad, f.h);
  for (int x = threadIdx.x; x < f.w; x += blockDim.x) {
    const float sv = f.s[(size_t)sy * f.w + x];
    const float2 c = carrier_of(f, sy, x);
    a[x] = make_float2(sv * c.x, -sv * c.y);  // S * conj(c)
  }
  __syncthreads();
  const float2* r = fft_lds(a, b, rt, P, -1, 1);
  for (int x = threadIdx.x; x < f.w; x += blockDim.x) f.spec[(size_t)yp * f.w + x] = r[x];
}

// (2) forward column FFT in place, kColGroup columns per block.
__global__ void __launch_bounds__(256) k_fft_cols_fwd(const InitField* F, const FftPlan* plans, const float2* roots,
                                                      const int* grp_prefix, int nf) {
  extern __shared__ float2 lds[];
  const int fi = field_of_tile(grp_prefix, nf, blockIdx.x);
  const InitField& f = F[fi];
  const int x0 = (blockIdx.x - grp_prefix[fi]) * kColGroup;
  const int nc = min(kColGroup, f.w - x0);
  const FftPlan& P = plans[f.col_plan];
  float2* a = lds;
  float2* b = lds + P.n * kColGroup;
  float2* rt = lds + 2 * P.n * kColGroup;
  load_root(rt, roots + P.root_off, P.n);
  for (int e = threadIdx.x; e < P.n * kColGroup; e += blockDim.x) {
    const int c = e % kColGroup, y = e / kColGroup;
    a[e] = c < nc ? f.spec[(size_t)y * f.w + x0 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* r = fft_lds(a, b, rt, P, -1, kColGroup);
  for (int e = threadIdx.x; e < P.n * kColGroup; e += blockDim.x) {
    const int c = e % kColGroup, y = e / kColGroup;
    if (c < nc) f.spec[(size_t)y * f.w + x0 + c] = r[e];
  }
}

// (3) per variant: spectrum x separable Gaussian (GaussianLpf :33-56, the
// vertical cutoff defined against the unpadded height), optional spectral
// symmetry certificate (:378-397), inverse column FFT; rows [pad, pad + h)
// are stored.
__global__ void __launch_bounds__(256) k_fft_cols_inv(const InitField* F, const FftPlan* plans, const float2* roots,
                                                      const int* grp_prefix, int nf) {
  extern __shared__ float2 lds[];
  const int fi = field_of_tile(grp_prefix, nf, blockIdx.x);
  const InitField& f = F[fi];
  const int x0 = (blockIdx.x - grp_prefix[fi]) * kColGroup;
  const int nc = min(kColGroup, f.w - x0);
  const FftPlan& P = plans[f.col_plan];
  const int hp = P.n;
  float2* a = lds;
  float2* b = lds + hp * kColGroup;
  float2* rt = lds + 2 * hp * kColGroup;
  load_root(rt, roots + P.root_off, hp);
  const float vh[3] = {0.8e6f, 1.8e6f, 1.3e6f};
  const float vv[3] = {120.f, 30.f, 60.f};
  const float inv_w = f.sample_rate / (float)f.w;  // fftfreq(w, 1/fs): k / (w / fs)
  const float inv_hp = 1.f / (float)hp;
  for (int v = 0; v < f.nvar; ++v) {
    const float cx = vh[v];
    const float cy = vv[v] / (2.f * (float)f.h);
    for (int e = threadIdx.x; e < hp * kColGroup; e += blockDim.x) {
      const int c = e % kColGroup, y = e / kColGroup;
      float2 val = make_float2(0.f, 0.f);
      if (c < nc) {
        const int x = x0 + c;
        const float rx = fftfreq(x, f.w, inv_w) / cx;
        const float ry = fftfreq(y, hp, inv_hp) / cy;
        float k = expf(-0.5f * ry * ry) * expf(-0.5f * rx * rx);
        const float2 sp = f.spec[(size_t)y * f.w + x];
        if (v == 2) {
          const int ry_ = (hp - y) % hp, rx_ = (f.w - x) % f.w;
          const float2 sr = f.spec[(size_t)ry_ * f.w + rx_];
          const float m = sqrtf(sp.x * sp.x + sp.y * sp.y), mr = sqrtf(sr.x * sr.x + sr.y * sr.y);
          k *= fminf(m, mr) / (m + 1e-6f);
        }
        val = make_float2(sp.x * k, sp.y * k);
      }
      a[e] = val;
    }
    __syncthreads();
    const float2* r = fft_lds(a, b, rt, P, +1, kColGroup);
    for (int e = threadIdx.x; e < hp * kColGroup; e += blockDim.x) {
      const int c = e % kColGroup, y = e / kColGroup - f.pad;
      if (c < nc && y >= 0 && y < f.h) f.var[v][(size_t)y * f.w + x0 + c] = r[e];
    }
    __syncthreads();
  }
}

// (4) inverse row FFT of every variant, scaled by 2 / (hp w) (the numpy
// inverse normalisation of fft2d.cpp:112-118 and the x2 of :358).
__global__ void __launch_bounds__(256) k_fft_rows_inv(const InitField* F, const FftPlan* plans, const float2* roots,
                                                      const int* row_prefix, int nf) {
  extern __shared__ float2 lds[];
  const int fi = field_of_tile(row_prefix, nf, blockIdx.x);
  const InitField& f = F[fi];
  const int y = blockIdx.x - row_prefix[fi];
  const FftPlan& P = plans[f.row_plan];
  float2* a = lds;
  float2* b = lds + P.n;
  float2* rt = lds + 2 * P.n;
  load_root(rt, roots + P.root_off, P.n);
  const float scale = 2.f / ((float)f.hp * (float)f.w);
  for (int v = 0; v < f.nvar; ++v) {
    float2* row = f.var[v] + (size_t)y * f.w;
    for (int x = threadIdx.x; x < f.w; x += blockDim.x) a[x] = row[x];
    __syncthreads();
    const float2* r = fft_lds(a, b, rt, P, +1, 1);
    for (int x = threadIdx.x; x < f.w; x += blockDim.x) row[x] = make_float2(r[x].x * scale, r[x].y * scale);
    __syncthreads();
  }
}

// (5) blend (holographic_init.cpp:171-196, :360-369, :415-422).  Tiles of
// kBlendW x kBlendH pixels; each variant's residual luma is staged with the
// 3 + 1 halo the edge-energy box blur needs.
__global__ void __launch_bounds__(256) k_holo_blend(const InitField* F, const int* tp, int nf) {
  constexpr int RW = kBlendW + 7, RH = kBlendH + 7;  // Y_v region (3 left/up, 4 right/down)
  constexpr int EW = kBlendW + 6, EH = kBlendH + 6;  // edge-energy region
  __shared__ float ys[RH][RW];
  __shared__ float es[EH][EW];
  __shared__ float hs[EH][kBlendW];
  const int fi = field_of_tile(tp, nf, blockIdx.x);
  const InitField& f = F[fi];
  const int lt = blockIdx.x - tp[fi];
  const int tiles_x = (f.w + kBlendW - 1) / kBlendW;
  const int bx0 = (lt % tiles_x) * kBlendW, by0 = (lt / tiles_x) * kBlendH;
  const int tx = threadIdx.x % kBlendW, ty = threadIdx.x / kBlendW;
  const int px = bx0 + tx, py = by0 + ty;
  const bool in = px < f.w && py < f.h;
  float2 num = make_float2(0.f, 0.f);
  float den = 0.f;
  for (int v = 0; v < f.nvar; ++v) {
    const float2* chv = f.var[v];
    for (int e = threadIdx.x; e < RH * RW; e += blockDim.x) {
      const int ry = e / RW, rx = e % RW;
      const int y = by0 - 3 + ry, x = bx0 - 3 + rx;
      float val = 0.f;
      if (y >= 0 && y < f.h && x >= 0 && x < f.w) {
        const size_t i = (size_t)y * f.w + x;
        val = f.s[i] - re_mul(chv[i], carrier_of(f, y, x));
      }
      ys[ry][rx] = val;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < EH * EW; e += blockDim.x) {
      const int ry = e / EW, rx = e % EW;
      const int y = by0 - 3 + ry, x = bx0 - 3 + rx;
      float val = 0.f;
      if (y >= 0 && y < f.h && x >= 0 && x < f.w) {
        const float gx = x + 1 < f.w ? ys[ry][rx + 1] - ys[ry][rx] : 0.f;
        const float gy = y + 1 < f.h ? ys[ry + 1][rx] - ys[ry][rx] : 0.f;
        val = fabsf(gx) + fabsf(gy);
      }
      es[ry][rx] = val;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < EH * kBlendW; e += blockDim.x) {
      const int ry = e / kBlendW, cx = e % kBlendW;
      float acc = 0.f;
      for (int k = 0; k < 7; ++k) acc += es[ry][cx + k];
      hs[ry][cx] = acc;
    }
    __syncthreads();
    if (in) {
      float acc = 0.f;
      for (int k = 0; k < 7; ++k) acc += hs[ty + k][tx];
      const float energy = acc * (1.f / 49.f);
      const float wgt = 1.f / (energy + 0.5f);
      const float2 c = chv[(size_t)py * f.w + px];
      num.x += c.x * wgt;
      num.y += c.y * wgt;
      den += wgt;
    }
    __syncthreads();
  }
  if (in) {
    const size_t i = (size_t)py * f.w + px;
    const float2 chi = make_float2(num.x / den, num.y / den);
    f.chi0[i] = chi;
    if (f.y0) f.y0[i] = f.s[i] - re_mul(chi, carrier_of(f, py, px));
  }
}

// ---- the numpy-convention 2-D DFT of engine/fft2d.h (fft2d.cpp:88-118) ----
// The 2-D DFT's row pass: one block per row, the row's FFT in LDS (sign -1
// forward, +1 inverse).  LDS: 3 n float2.
__global__ void __launch_bounds__(256) k_fft2_rows(const float2* in, float2* out, int h, int w, FftPlan P,
                                                   const float2* roots, int sign) {
  extern __shared__ float2 lds[];
  const int y = blockIdx.x;
  float2* a = lds;
  float2* b = lds + P.n;
  float2* rt = lds + 2 * P.n;
  for (int j = threadIdx.x; j < P.n; j += blockDim.x) rt[j] = roots[P.root_off + j];
  for (int x = threadIdx.x; x < w; x += blockDim.x) a[x] = in[(size_t)y * w + x];
  __syncthreads();
  const float2* r = fft_lds(a, b, rt, P, sign, 1);
  for (int x = threadIdx.x; x < w; x += blockDim.x) out[(size_t)y * w + x] = r[x];
}

// The column pass, in place, kColGroup columns per block, each output times
// `scale` (1 forward, 1 / (h w) inverse).  LDS: (2 kColGroup + 1) n float2.
__global__ void __launch_bounds__(256) k_fft2_cols(float2* io, int h, int w, FftPlan P, const float2* roots, int sign,
                                                   float scale) {
  extern __shared__ float2 lds[];
  const int x0 = blockIdx.x * kColGroup;
  const int nc = min(kColGroup, w - x0);
  float2* a = lds;
  float2* b = lds + P.n * kColGroup;
  float2* rt = lds + 2 * P.n * kColGroup;
  for (int j = threadIdx.x; j < P.n; j += blockDim.x) rt[j] = roots[P.root_off + j];
  for (int e = threadIdx.x; e < P.n * kColGroup; e += blockDim.x) {
    const int c = e % kColGroup, y = e / kColGroup;
    a[e] = c < nc ? io[(size_t)y * w + x0 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* r = fft_lds(a, b, rt, P, sign, kColGroup);
  for (int e = threadIdx.x; e < P.n * kColGroup; e += blockDim.x) {
    const int c = e % kColGroup, y = e / kColGroup;
    if (c < nc) {
      const float2 v = r[e];
      io[(size_t)y * w + x0 + c] = scale == 1.0f ? v : make_float2(v.x * scale, v.y * scale);
    }
  }
}

}  // namespace hvdk
