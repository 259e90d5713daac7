// holo.hip -- holographic chroma init (engine/holographic_init.cpp:200-317)
// with project-written exact-size FFTs.
//
// The 2-D transforms keep hvd-core's sizes exactly (numpy semantics of
// engine/fft2d.h:30-33, no zero padding to friendly sizes -- that would
// change the circular convolution): NTSC 274 x 760, PAL 320 x 922.  Each 1-D
// transform runs in LDS as a mixed-radix Stockham autosort pass sequence in
// which the outputs of a radix-R stage are the R-point DFTs of pre-twiddled
// inputs
//     y[(j R + q) p + k] = sum_r (x[ib + r M] W_n^{r k n / (p R)}) W_R^{r q};
// small radices are written-out butterflies, odd primes below kRaderMin use
// the symmetric pair form, and longer primes (137 in NTSC columns, 461 in PAL
// rows) run by Rader's algorithm.  Every twiddle is a lookup into a table of
// roots computed in FP64 on the host.
//
// Pipeline per field (5 launches for the whole batch):
//   k_fft_rows_fwd  demod S conj(c) + reflect101 row padding + row FFT
//   k_fft_cols_fwd  column FFT in place (8 columns per block)
//   k_fft_cols_inv  per variant: spectrum x Gaussian crop, inverse column
//                   FFT, only the rows that survive the crop are written
//   k_fft_rows_inv  inverse row FFT, x 2 / (hp w)
//   k_holo_blend    residual luma, edge energy (7x7 box in LDS), weighted
//                   blend, Y0 = S - Re[chi0 c]
#include <initializer_list>
#include <map>
#include <mutex>
#include <tuple>

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

// Quotient by a run-time divisor without the integer-division sequence:
// floor(e / d) = (int)((e + 1/2) * fl(1/d)) exactly for 0 <= e < 2^21 and
// 1 <= d <= 4096 (the rounding error stays below 1/(2d); an index here is at
// most n * ncols <= 2^15).
struct Div {
  int d;
  float inv;
  __device__ explicit Div(int d_) : d(d_), inv(1.f / (float)d_) {}
  __device__ int q(int e) const { return (int)(((float)e + 0.5f) * inv); }
};

// A value every lane of the wave holds (a plan field), made a scalar so that
// the sizes and loop bounds derived from it stay in scalar registers.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// x * (si i), si = +-1
__device__ inline float2 mul_si(float2 x, float si) { return make_float2(-si * x.y, si * x.x); }
__device__ inline float2 cmul_tw(float2 x, float2 w) {
  return make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
}

}  // namespace

// The FFT kernels' dynamic LDS.  The transform helpers address it through Lv
// (an offset in float2 slots) rather than through pointers, so that every
// access compiles to an LDS access whichever buffer a stage ends in.
extern __shared__ float2 fsm[];

namespace {

struct Lv {
  int o;
  __device__ float2& operator[](int i) const { return fsm[o + i]; }
  __device__ Lv operator+(int i) const { return Lv{o + i}; }
  __device__ bool operator==(Lv b) const { return o == b.o; }
};

// The roots of a length-n transform: table[j] = e^{-2 pi i j / n}; the
// inverse direction reads the conjugates.
struct Roots {
  Lv t;
  bool inv;
  __device__ float2 operator()(int ex) const {
    const float2 w = t[ex];
    return make_float2(w.x, inv ? -w.y : w.y);
  }
};

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
template <int R>
__device__ inline void small_stage(Lv src, Lv dst, int n, int p, int M, int step1, float si, int ncols,
                                   Roots twiddle) {
  const Div dc(ncols), dp(p);
  for (int e = threadIdx.x; e < M * ncols; e += blockDim.x) {
    const int ib = dc.q(e), c = e - ib * ncols;
    const int j = dp.q(ib), k = ib - j * p;
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = src[(ib + r * M) * ncols + c];
      if (r > 0 && k != 0) x = cmul_tw(x, twiddle(r * k * step1));  // < R (p / p) (n / R) = n
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

// One stage of radix R at stride p (p = product of the earlier radices) of a
// length-n transform of `ncols` interleaved sequences: element (j, c) lives at
// buf[j * ncols + c].  On return *src holds the stage's output (the pointers
// swap when it was written into the spare buffer).  Ends with a barrier.
//
// An odd prime R >= 11 runs as a length-R DFT of pre-twiddled inputs
// a_r = x_r W^{r k step1} (k the output group's phase), using both of its
// symmetries: the terms r and R - r pair as u_r = a_r + a_{R-r},
// v_r = a_r - a_{R-r}, and the outputs q and R - q share
// U = sum_r u_r cos(2 pi r q / R), V = sum_r v_r sin(2 pi r q / R):
//     y_q = a_0 + U + i V,   y_{R-q} = a_0 + U - i V.
// A first pass forms [a_0, u_1, v_1, ..., u_h, v_h] per group (h = (R - 1) / 2)
// into the spare buffer, a second forms two outputs per work item from it: a
// quarter of the multiply-adds of the direct R-term sums.
__device__ inline void plain_stage(Lv& src, Lv& dst, int n, int p, int R, int ncols, Roots twiddle) {
  const int M = n / R;
  const int step1 = n / (p * R);
  const float si = twiddle.inv ? 1.f : -1.f;
  const Div dc(ncols), dp(p);
  if (R >= 11 && (R & 1)) {
    const int h = (R - 1) / 2;
    const Div dh(h + 1);
    // pass 1: group ib (0 <= ib < M) -> dst[ib R ..] = a_0, u_1, v_1, ...
    for (int e = threadIdx.x; e < M * (h + 1) * ncols; e += blockDim.x) {
      const int g = dc.q(e), c = e - g * ncols;
      const int ib = dh.q(g), r = g - ib * (h + 1);
      const int k = ib - dp.q(ib) * p;
      auto pre = [&](int rr) {  // a_rr = x_rr W^{rr k step1}
        const float2 x = src[(ib + rr * M) * ncols + c];
        if (k == 0) return x;
        return cmul_tw(x, twiddle(rr * k * step1));  // k step1 < n / R
      };
      if (r == 0) {
        dst[(ib * R) * ncols + c] = pre(0);
        continue;
      }
      const float2 x = pre(r), y = pre(R - r);
      dst[(ib * R + 2 * r - 1) * ncols + c] = cadd(x, y);
      dst[(ib * R + 2 * r) * ncols + c] = csub(x, y);
    }
    __syncthreads();
    // pass 2: outputs q and R - q of group ib into src, kQ values of q per
    // work item so that each u_r, v_r read from LDS feeds 4 kQ products
    constexpr int kQ = 4;
    const int nq = (h + kQ) / kQ;  // q = 0 .. h in blocks of kQ
    const Div dq(nq);
    for (int e = threadIdx.x; e < M * nq * ncols; e += blockDim.x) {
      const int g = dc.q(e), c = e - g * ncols;
      const int ib = dq.q(g), q0 = (g - ib * nq) * kQ;
      const int j = dp.q(ib), k = ib - j * p;
      const Lv gp = dst + (ib * R) * ncols + c;
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
    return;
  }
  if (R <= 8 && R != 6) {
    switch (R) {
      case 2: small_stage<2>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
      case 3: small_stage<3>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
      case 4: small_stage<4>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
      case 5: small_stage<5>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
      case 7: small_stage<7>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
      default: small_stage<8>(src, dst, n, p, M, step1, si, ncols, twiddle); break;
    }
  } else {  // any other radix: direct R-term sums
    const Div dr(R);
    for (int e = threadIdx.x; e < n * ncols; e += blockDim.x) {
      const int o = dc.q(e), c = e - o * ncols;
      const int t = dp.q(o), k = o - t * p;
      const int tr = dr.q(t), q = t - tr * R, ib = tr * p + k;
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
  }
  __syncthreads();
  const Lv t2 = src;
  src = dst;
  dst = t2;
}

// A transform made of plain stages only; returns the buffer holding the result.
__device__ inline Lv fft_plain(Lv a, Lv b, Lv root, int n, int nst, const int* radix, bool inv, int ncols) {
  const Roots tw{root, inv};
  Lv src = a;
  Lv dst = b;
  int p = 1;
  for (int s = 0; s < nst; ++s) {
    const int R = uni(radix[s]);
    plain_stage(src, dst, n, p, R, ncols, tw);
    p *= R;
  }
  return src;
}

// Rader stage of prime radix R at stride p.  With g a generator of (Z/R)*,
// the outputs of one group are
//     y_0 = sum_r a_r,     y_{g^-m} = a_0 + sum_l a_{g^l} w^{g^(l - m)}
// (w = W_R^{+-1}, l, m in [0, L), L = R - 1): the second sum is the cyclic
// convolution of u_l = a_{g^l} with b_l = w^{g^-l}, formed as
// IDFT_L(DFT_L(u) . bhat) with bhat = DFT_L(b) / L (host table) and the
// length-L transforms run as plain stages over all M * ncols groups at once.
// LDS region `rl`: L roots of L, bhat, the permutation {g^l, g^-l} mod R,
// then a_0 and y_0 of every group.  On return *src holds the output.
__device__ inline void rader_stage(Lv& src, Lv& dst, int n, int p, int R, int ncols, Roots twiddle,
                                   const FftRader& RD, const float2* gtab, Lv rl) {
  const int M = n / R, step1 = n / (p * R), L = uni(RD.L), S = M * ncols;
  const Lv iroot = rl;
  const Lv bhat = rl + L;
  int2* perm = reinterpret_cast<int2*>(fsm + rl.o + 2 * L);
  const Lv a0 = rl + 3 * L;
  const Lv y0 = a0 + S;
  const float2* tab = gtab + uni(RD.tab);
  const int inner_nst = uni(RD.nst);
  for (int j = threadIdx.x; j < L; j += blockDim.x) {
    iroot[j] = tab[j];
    bhat[j] = tab[(twiddle.inv ? 2 : 1) * L + j];
    perm[j] = reinterpret_cast<const int2*>(tab + 3 * L)[j];
  }
  __syncthreads();
  // (A) gather u_l = a_{g^l} into dst[l S + s] and a_0 into a0[s], where
  // s = ib ncols + c, pre-twiddled by W^{r k step1} (< n)
  const Div dS(S), dc(ncols), dp(p);
  for (int e = threadIdx.x; e < R * S; e += blockDim.x) {
    const int l = dS.q(e), s = e - l * S;  // l == L: the a_0 slots
    const bool nz = l < L;
    const int r = nz ? perm[l].x : 0;
    const int ib = dc.q(s), c = s - ib * ncols, k = ib - dp.q(ib) * p;
    float2 x = src[(ib + r * M) * ncols + c];
    if (r > 0 && k != 0) x = cmul_tw(x, twiddle(r * k * step1));
    if (nz)
      dst[e] = x;
    else
      a0[s] = x;
  }
  __syncthreads();
  // (B) U = DFT_L(u) of the S interleaved sequences
  const Lv X = fft_plain(dst, src, iroot, L, inner_nst, RD.radix, false, S);
  const Lv T = X == dst ? src : dst;
  // (C) y_0 = a_0 + U_0; U . bhat
  for (int e = threadIdx.x; e < L * S; e += blockDim.x) {
    const int l = dS.q(e), s = e - l * S;
    const float2 u = X[e];
    if (l == 0) y0[s] = cadd(a0[s], u);
    X[e] = cmul_tw(u, bhat[l]);
  }
  __syncthreads();
  // (D) the convolution: unscaled inverse DFT_L (the 1 / L is in bhat)
  const Lv Z = fft_plain(X, T, iroot, L, inner_nst, RD.radix, true, S);
  const Lv O = Z == X ? T : X;
  // (E) y_{g^-m} = a_0 + c_m and y_0 to their Stockham places
  for (int e = threadIdx.x; e < R * S; e += blockDim.x) {
    const int m = dS.q(e), s = e - m * S;  // m == L: the y_0 slots
    const bool nz = m < L;
    const int q = nz ? perm[m].y : 0;
    const float2 v = nz ? cadd(a0[s], Z[e]) : y0[s];
    const int ib = dc.q(s), c = s - ib * ncols, j = dp.q(ib), k = ib - j * p;
    O[((j * R + q) * p + k) * ncols + c] = v;
  }
  __syncthreads();
  src = O;
  dst = Z;
}

// In-LDS transform of `ncols` interleaved sequences of plan P (element (j, c)
// at a[j * ncols + c]); sign < 0 forward (e^{-2 pi i / n}), > 0 inverse,
// unscaled.  `root` holds the n roots in LDS, `gtab` is the root buffer in
// global memory (the Rader tables' home), `rl` the LDS the Rader stages use
// (fft_lds_slots).  Returns the buffer holding the result (a or b).
__device__ __attribute__((always_inline)) Lv fft_lds(Lv a, Lv b, Lv root, const FftPlan& P, int sign, int ncols,
                                                     const float2* gtab, Lv rl) {
  const Roots tw{root, sign > 0};
  Lv src = a;
  Lv dst = b;
  int p = 1;
  const int n = uni(P.n), nst = uni(P.nst);
  for (int s = 0; s < nst; ++s) {
    const int R = uni(P.radix[s]);
    const int ri = uni(P.rader_of[s]);
    if (ri >= 0)
      rader_stage(src, dst, n, p, R, ncols, tw, P.rader[ri], gtab, rl);
    else
      plain_stage(src, dst, n, p, R, ncols, tw);
    p *= R;
  }
  return src;
}

// ---- transforms whose length, interleave and stages are compile-time ----
// The lengths the decode uses (NTSC 274 x 760, PAL 320 x 922) run through
// these: every index is then a constant expression or a multiply-shift, the
// odd-radix DFTs multiply by literal cos/sin constants instead of reading the
// root table, and a Rader stage's inner transform has a fixed size.  Same
// stage order, tables and Stockham layout as the run-time plan (fft_lds).
constexpr int kFftThreads = 256;  // every FFT block

template <int R> struct St {};               // a plain stage of radix R
template <int R, int... In> struct Rd {};    // a Rader stage of prime R; In: the radices of R - 1
template <class... S> struct Stages {};

constexpr double kTwoPi = 6.283185307179586476925286766559;
// cos and sin of 2 pi j / R in double (Taylor series after reducing the angle
// to [-pi, pi]), rounded once to float.
constexpr double turn_angle(int j, int R) {
  const double x = kTwoPi * (double)(j % R) / (double)R;
  return x > kTwoPi / 2 ? x - kTwoPi : x;
}
constexpr double turn_cos(int j, int R) {
  const double x = turn_angle(j, R);
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 40; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
constexpr double turn_sin(int j, int R) {
  const double x = turn_angle(j, R);
  double term = x, sum = x;
  for (int k = 1; k < 40; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
template <int R> struct Turns {
  float c[R], s[R];
  constexpr Turns() : c(), s() {
    for (int j = 0; j < R; ++j) {
      c[j] = (float)turn_cos(j, R);
      s[j] = (float)turn_sin(j, R);
    }
  }
};

// y = DFT_R(a), W_R = e^{-+2 pi i / R} (inv: +).
template <int R>
__device__ inline void ct_dft(const float2 (&a)[R], float2 (&y)[R], bool inv) {
  const float si = inv ? 1.f : -1.f;
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
  } else {  // odd R: the symmetric pair form with constant cos / sin
    static_assert(R & 1, "odd radix expected");
    constexpr int h = (R - 1) / 2;
    constexpr Turns<R> T{};
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
        const float cw = T.c[(r * q) % R], sw = T.s[(r * q) % R];
        ux = __fmaf_rn(u[r].x, cw, ux);
        uy = __fmaf_rn(u[r].y, cw, uy);
        vx = __fmaf_rn(v[r].x, sw, vx);
        vy = __fmaf_rn(v[r].y, sw, vy);
      }
      vx *= si;  // the sines of W_R^{rq} carry the direction's sign
      vy *= si;
      y[q] = make_float2(a[0].x + ux - vy, a[0].y + uy + vx);
      y[R - q] = make_float2(a[0].x + ux + vy, a[0].y + uy - vx);
    }
  }
}

// A plain stage of radix R at stride P of NC interleaved length-N sequences.
template <int N, int NC, int P, int R>
__device__ inline void ct_stage(Lv& src, Lv& dst, Roots tw) {
  constexpr int M = N / R, STEP1 = N / (P * R);
  for (int e = threadIdx.x; e < M * NC; e += kFftThreads) {
    const int c = e % NC, ib = e / NC;
    const int k = ib % P, j = ib / P;
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      float2 x = src[(ib + r * M) * NC + c];
      if (P > 1 && r > 0) x = cmul_tw(x, tw(r * k * STEP1));  // root 0 is exactly 1
      a[r] = x;
    }
    float2 y[R];
    ct_dft<R>(a, y, tw.inv);
#pragma unroll
    for (int q = 0; q < R; ++q) dst[((j * R + q) * P + k) * NC + c] = y[q];
  }
  __syncthreads();
  const Lv t = src;
  src = dst;
  dst = t;
}

template <int N, int NC, int P, class List> struct CtFft;

// Rader stage (rader_stage with the sizes fixed; the plan's first Rader
// tables).  On return src holds the output.
template <int N, int NC, int P, int R, int... In>
__device__ inline void ct_rader(Lv& src, Lv& dst, Roots tw, const FftPlan& Pl, const float2* gtab, Lv rl) {
  constexpr int M = N / R, STEP1 = N / (P * R), L = R - 1, S = M * NC;
  const Lv iroot = rl, bhat = rl + L;
  int2* perm = reinterpret_cast<int2*>(fsm + rl.o + 2 * L);
  const Lv a0 = rl + 3 * L, y0 = a0 + S;
  const float2* tab = gtab + uni(Pl.rader[0].tab);
  for (int jj = threadIdx.x; jj < L; jj += kFftThreads) {
    iroot[jj] = tab[jj];
    bhat[jj] = tab[(tw.inv ? 2 : 1) * L + jj];
    perm[jj] = reinterpret_cast<const int2*>(tab + 3 * L)[jj];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < R * S; e += kFftThreads) {
    const int l = e / S, s = e % S;  // l == L: the a_0 slots
    const bool nz = l < L;
    const int r = nz ? perm[l].x : 0;
    const int ib = s / NC, c = s % NC, k = ib % P;
    float2 x = src[(ib + r * M) * NC + c];
    if (P > 1) x = cmul_tw(x, tw(r * k * STEP1));
    if (nz)
      dst[e] = x;
    else
      a0[s] = x;
  }
  __syncthreads();
  Lv X = dst, T = src;
  CtFft<L, S, 1, Stages<St<In>...>>::run(X, T, Roots{iroot, false}, Pl, gtab, rl);
  for (int e = threadIdx.x; e < L * S; e += kFftThreads) {
    const int l = e / S, s = e % S;
    const float2 u = X[e];
    if (l == 0) y0[s] = cadd(a0[s], u);
    X[e] = cmul_tw(u, bhat[l]);
  }
  __syncthreads();
  Lv Z = X, O = T;
  CtFft<L, S, 1, Stages<St<In>...>>::run(Z, O, Roots{iroot, true}, Pl, gtab, rl);
  for (int e = threadIdx.x; e < R * S; e += kFftThreads) {
    const int m = e / S, s = e % S;  // m == L: the y_0 slots
    const bool nz = m < L;
    const int q = nz ? perm[m].y : 0;
    const float2 v = nz ? cadd(a0[s], Z[e]) : y0[s];
    const int ib = s / NC, c = s % NC, j = ib / P, k = ib % P;
    O[((j * R + q) * P + k) * NC + c] = v;
  }
  __syncthreads();
  src = O;
  dst = Z;
}

template <int N, int NC, int P> struct CtFft<N, NC, P, Stages<>> {
  static __device__ void run(Lv&, Lv&, Roots, const FftPlan&, const float2*, Lv) {}
};
template <int N, int NC, int P, int R, class... Rest> struct CtFft<N, NC, P, Stages<St<R>, Rest...>> {
  static __device__ void run(Lv& s, Lv& d, Roots tw, const FftPlan& Pl, const float2* g, Lv rl) {
    ct_stage<N, NC, P, R>(s, d, tw);
    CtFft<N, NC, P * R, Stages<Rest...>>::run(s, d, tw, Pl, g, rl);
  }
};
template <int N, int NC, int P, int R, int... In, class... Rest>
struct CtFft<N, NC, P, Stages<Rd<R, In...>, Rest...>> {
  static __device__ void run(Lv& s, Lv& d, Roots tw, const FftPlan& Pl, const float2* g, Lv rl) {
    ct_rader<N, NC, P, R, In...>(s, d, tw, Pl, g, rl);
    CtFft<N, NC, P * R, Stages<Rest...>>::run(s, d, tw, Pl, g, rl);
  }
};

// The transform a kernel instantiation runs: a fixed length (FixedFft) or any
// plan (AnyFft), NC interleaved sequences per block.
// kPer: the elements of one block of NC sequences a thread holds in registers
// (fixed lengths prefetch the next group through them).
template <int N, int NC, class List> struct FixedFft {
  static constexpr int kN = N, kNc = NC, kPer = (N * NC + kFftThreads - 1) / kFftThreads;
  static __device__ Lv run(Lv a, Lv b, Lv rt, const FftPlan& P, int sign, const float2* gtab, Lv rl) {
    Lv s = a, d = b;
    CtFft<N, NC, 1, List>::run(s, d, Roots{rt, sign > 0}, P, gtab, rl);
    return s;
  }
};
template <int NC> struct AnyFft {
  static constexpr int kN = 0, kNc = NC, kPer = 1;
  static __device__ Lv run(Lv a, Lv b, Lv rt, const FftPlan& P, int sign, const float2* gtab, Lv rl) {
    return fft_lds(a, b, rt, P, sign, NC, gtab, rl);
  }
};

using RowsAny = AnyFft<kRowGroup>;
using Rows760 = FixedFft<760, kRowGroup, Stages<St<8>, St<5>, St<19>>>;
using Rows922 = FixedFft<922, kRowGroup, Stages<St<2>, Rd<461, 4, 5, 23>>>;
using ColsAny = AnyFft<kColGroup>;
using Cols274 = FixedFft<274, kColGroup, Stages<St<2>, Rd<137, 8, 17>>>;
using Cols320 = FixedFft<320, kColGroup, Stages<St<8>, St<8>, St<5>>>;

__device__ inline void load_root(Lv lds, const float2* root, int n) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) lds[j] = root[j];
}

// numpy fftfreq(n, d) bin i (holographic_init.cpp:16-25), cycles per unit.
__device__ inline float fftfreq(int i, int n, float inv_nd) {
  const int half = (n % 2 == 0) ? n / 2 : (n + 1) / 2;
  const int k = i < half ? i : i - n;
  return (float)k * inv_nd;
}

}  // namespace

// The four FFT passes of the init run over a list of fields (`fields`, with
// `prefix` the running count of blocks per listed field) that share one
// transform class Fft: a fixed length or any plan.

// A block of a pass takes the groups [g0, g1): blocks are persistent over a
// contiguous run of groups (the blocks of one XCD over neighbouring runs), so
// that a fixed-length class can load its next group into registers while the
// current one is transformed.
__device__ inline void block_groups(int ng, int& g0, int& g1) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  g0 = (int)((long long)b * ng / gridDim.x);
  g1 = (int)((long long)(b + 1) * ng / gridDim.x);
}

// Group g of a pass: its field, its first row or column and how many of its G
// exist, counted along the padded height (kAlongHp), the width (kAlongW) or
// the height (kAlongH).
enum { kAlongHp, kAlongW, kAlongH };
struct Group {
  const InitField* f;
  int lo, cnt;
};
__device__ inline Group group_at(const InitField* F, const int* prefix, const int* fields, int nf, int g, int G,
                                 int along) {
  const int t = field_of_tile(prefix, nf, g);
  const InitField* f = &F[fields[t]];
  const int lo = (g - prefix[t]) * G;
  const int ext = along == kAlongHp ? f->hp : along == kAlongW ? f->w : f->h;
  return Group{f, lo, min(G, ext - lo)};
}

// The plan of the group's field becomes the block's: the roots are reloaded
// when it changes (never within a fixed-length class).
__device__ inline void use_plan(const FftPlan* plans, const float2* roots, int pi, int G, int& plan, int& n) {
  if (pi == plan) return;
  plan = pi;
  n = uni(plans[pi].n);
  load_root(Lv{2 * n * G}, roots + uni(plans[pi].root_off), n);
}

// (1) demod + reflect101 padding + forward row FFT.  Fft::kNc padded rows
// per group, interleaved in LDS.
template <class Fft>
__global__ void __launch_bounds__(kFftThreads) k_fft_rows_fwd(const InitField* F, const FftPlan* plans,
                                                              const float2* roots, const int* prefix,
                                                              const int* fields, int nf) {
  constexpr int G = Fft::kNc, KP = Fft::kPer;
  int g0, g1;
  block_groups(prefix[nf], g0, g1);
  if (g0 >= g1) return;
  static_assert(kFftThreads % G == 0, "a thread keeps one row of a group");
  const int g = threadIdx.x % G;  // the row (of the group) of every element of this thread
  Group cur = group_at(F, prefix, fields, nf, g0, G, kAlongHp);
  // fixed lengths: S and the carrier's column factor (or the explicit
  // carrier) of the next group, and the thread's row factor and V switch;
  // the carrier is formed when the group is staged
  float spre[KP];
  float2 cpre[KP];
  float2 rpre = make_float2(1.f, 0.f);
  bool neg = false;
  auto fetch = [&](const Group& q) {
    const InitField& f = *q.f;
    const bool row = g < q.cnt;
    const int sy = row ? mirror_row(q.lo + g - f.pad, f.h) : 0;
    if (row && !f.carrier) {
      rpre = f.rowph[sy];
      neg = f.vsw && f.vsw[sy] < 0;
    }
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int e = threadIdx.x + k * kFftThreads, x = e / G;
      spre[k] = 0.f;
      cpre[k] = make_float2(0.f, 0.f);
      if (e < Fft::kN * G && row) {
        spre[k] = f.s[(size_t)sy * f.w + x];
        cpre[k] = f.carrier ? f.carrier[(size_t)sy * f.w + x] : f.colph[x];
      }
    }
  };
  if constexpr (Fft::kN > 0) fetch(cur);
  int plan = -1, n = 0;
  for (int gi = g0; gi < g1; ++gi) {
    const Group here = cur;
    const InitField& f = *here.f;
    use_plan(plans, roots, uni(f.row_plan), G, plan, n);  // n == f.w
    const FftPlan& P = plans[plan];
    const Lv a{0}, b{n * G}, rt{2 * n * G};
    if constexpr (Fft::kN > 0) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int e = threadIdx.x + k * kFftThreads;
        if (e >= Fft::kN * G) break;
        float2 v = make_float2(0.f, 0.f);
        if (g < here.cnt) {
          float2 c = cpre[k];
          if (!f.carrier) {  // carrier_of's row factor and V switch
            c = cmul(rpre, c);
            if (neg) c.x = -c.x;
          }
          v = make_float2(spre[k] * c.x, -spre[k] * c.y);  // S * conj(c)
        }
        a[e] = v;
      }
      if (gi + 1 < g1) {
        cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongHp);
        fetch(cur);
      }
    } else {
      for (int e = threadIdx.x; e < n * G; e += kFftThreads) {
        const int x = e / G;
        float2 v = make_float2(0.f, 0.f);
        if (g < here.cnt) {
          const int sy = mirror_row(here.lo + g - f.pad, f.h);
          const float sv = f.s[(size_t)sy * f.w + x];
          const float2 c = carrier_of(f, sy, x);
          v = make_float2(sv * c.x, -sv * c.y);  // S * conj(c)
        }
        a[e] = v;
      }
      if (gi + 1 < g1) cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongHp);
    }
    __syncthreads();
    const Lv r = Fft::run(a, b, rt, P, -1, roots, rt + n);
    for (int e = threadIdx.x; e < n * G; e += kFftThreads) {
      const int x = e / G;
      if (g < here.cnt) f.spec[(size_t)(here.lo + g) * f.w + x] = r[e];
    }
    __syncthreads();
  }
}

// A fixed-length class's next column group of the spectrum into registers:
// element e = threadIdx.x + k kFftThreads is (row e / G, column c).
template <class Fft>
__device__ inline void fetch_spec_cols(float2 (&pre)[Fft::kPer], const Group& q, int c) {
  constexpr int G = Fft::kNc;
#pragma unroll
  for (int k = 0; k < Fft::kPer; ++k) {
    const int e = threadIdx.x + k * kFftThreads, y = e / G;
    pre[k] = e < Fft::kN * G && c < q.cnt ? q.f->spec[(size_t)y * q.f->w + q.lo + c] : make_float2(0.f, 0.f);
  }
}

// (2) forward column FFT in place, Fft::kNc columns per group.
template <class Fft>
__global__ void __launch_bounds__(kFftThreads) k_fft_cols_fwd(const InitField* F, const FftPlan* plans,
                                                              const float2* roots, const int* prefix,
                                                              const int* fields, int nf) {
  constexpr int G = Fft::kNc, KP = Fft::kPer;
  static_assert(kFftThreads % G == 0, "a thread keeps one column");
  int g0, g1;
  block_groups(prefix[nf], g0, g1);
  if (g0 >= g1) return;
  const int c = threadIdx.x % G;
  Group cur = group_at(F, prefix, fields, nf, g0, G, kAlongW);
  float2 pre[KP];
  if constexpr (Fft::kN > 0) fetch_spec_cols<Fft>(pre, cur, c);
  int plan = -1, n = 0;
  for (int gi = g0; gi < g1; ++gi) {
    const Group here = cur;
    const InitField& f = *here.f;
    use_plan(plans, roots, uni(f.col_plan), G, plan, n);  // n == f.hp
    const FftPlan& P = plans[plan];
    const Lv a{0}, b{n * G}, rt{2 * n * G};
    if constexpr (Fft::kN > 0) {
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int e = threadIdx.x + k * kFftThreads;
        if (e < Fft::kN * G) a[e] = pre[k];
      }
      if (gi + 1 < g1) {
        cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongW);
        fetch_spec_cols<Fft>(pre, cur, c);
      }
    } else {
      for (int e = threadIdx.x; e < n * G; e += kFftThreads) {
        const int y = e / G;
        a[e] = c < here.cnt ? f.spec[(size_t)y * f.w + here.lo + c] : make_float2(0.f, 0.f);
      }
      if (gi + 1 < g1) cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongW);
    }
    __syncthreads();
    const Lv r = Fft::run(a, b, rt, P, -1, roots, rt + n);
    for (int e = threadIdx.x; e < n * G; e += kFftThreads) {
      const int y = e / G;
      if (c < here.cnt) f.spec[(size_t)y * f.w + here.lo + c] = r[e];
    }
    __syncthreads();
  }
}

// (3) per variant: spectrum x separable Gaussian (GaussianLpf :33-56, the
// vertical cutoff defined against the unpadded height), optional spectral
// symmetry certificate (:378-397), inverse column FFT; rows [pad, pad + h)
// are stored.
template <class Fft>
__global__ void __launch_bounds__(kFftThreads) k_fft_cols_inv(const InitField* F, const FftPlan* plans,
                                                              const float2* roots, const int* prefix,
                                                              const int* fields, int nf) {
  constexpr int G = Fft::kNc, KP = Fft::kPer;
  static_assert(kFftThreads % G == 0, "a thread keeps one column");
  int g0, g1;
  block_groups(prefix[nf], g0, g1);
  if (g0 >= g1) return;
  const int c = threadIdx.x % G;
  Group cur = group_at(F, prefix, fields, nf, g0, G, kAlongW);
  // fixed lengths: the group's spectrum columns, read once for every variant
  float2 pre[KP];
  if constexpr (Fft::kN > 0) fetch_spec_cols<Fft>(pre, cur, c);
  const float vh[3] = {0.8e6f, 1.8e6f, 1.3e6f};
  const float vv[3] = {120.f, 30.f, 60.f};
  int plan = -1, hp = 0;
  for (int gi = g0; gi < g1; ++gi) {
    const Group here = cur;
    const InitField& f = *here.f;
    use_plan(plans, roots, uni(f.col_plan), G, plan, hp);  // hp == f.hp
    const FftPlan& P = plans[plan];
    const Lv a{0}, b{hp * G}, rt{2 * hp * G};
    const float inv_w = f.sample_rate / (float)f.w;  // fftfreq(w, 1/fs): k / (w / fs)
    const float inv_hp = 1.f / (float)hp;
    // the separable Gaussian's vertical factors, one per spectrum row, after
    // the transform's LDS (the host sizes the block for them)
    float* gy = reinterpret_cast<float*>(&fsm[rt.o + hp + uni(P.rader_fixed) + uni(P.rader_per_col) * G]);
    const int x = here.lo + c;
    const int nvar = f.nvar;
    for (int v = 0; v < nvar; ++v) {
      const float cx = vh[v];
      const float cy = vv[v] / (2.f * (float)f.h);
      for (int y = threadIdx.x; y < hp; y += kFftThreads) {
        const float ry = fftfreq(y, hp, inv_hp) / cy;
        gy[y] = expf(-0.5f * ry * ry);
      }
      const float rx = fftfreq(x, f.w, inv_w) / cx;
      const float gx = expf(-0.5f * rx * rx);
      __syncthreads();
      // spectrum x Gaussian (x the symmetry certificate for variant 2)
      auto weighted = [&](int y, float2 sp) {
        float k = gy[y] * gx;
        if (v == 2) {
          const int ry_ = (hp - y) % hp, rx_ = (f.w - x) % f.w;
          const float2 sr = f.spec[(size_t)ry_ * f.w + rx_];
          const float m = sqrtf(sp.x * sp.x + sp.y * sp.y), mr = sqrtf(sr.x * sr.x + sr.y * sr.y);
          k *= fminf(m, mr) / (m + 1e-6f);
        }
        return make_float2(sp.x * k, sp.y * k);
      };
      if constexpr (Fft::kN > 0) {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
          const int e = threadIdx.x + k * kFftThreads;
          if (e >= Fft::kN * G) break;
          a[e] = c < here.cnt ? weighted(e / G, pre[k]) : make_float2(0.f, 0.f);
        }
        if (v == nvar - 1 && gi + 1 < g1) {
          cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongW);
          fetch_spec_cols<Fft>(pre, cur, c);
        }
      } else {
        for (int e = threadIdx.x; e < hp * G; e += kFftThreads) {
          const int y = e / G;
          a[e] = c < here.cnt ? weighted(y, f.spec[(size_t)y * f.w + x]) : make_float2(0.f, 0.f);
        }
        if (v == nvar - 1 && gi + 1 < g1) cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongW);
      }
      __syncthreads();
      const Lv r = Fft::run(a, b, rt, P, +1, roots, rt + hp);
      for (int e = threadIdx.x; e < hp * G; e += kFftThreads) {
        const int y = e / G - f.pad;
        if (c < here.cnt && y >= 0 && y < f.h) f.var[v][(size_t)y * f.w + x] = r[e];
      }
      __syncthreads();
    }
  }
}

// (4) inverse row FFT of every variant, scaled by 2 / (hp w) (the numpy
// inverse normalisation of fft2d.cpp:112-118 and the x2 of :358); Fft::kNc
// rows per block.
template <class Fft>
__global__ void __launch_bounds__(kFftThreads) k_fft_rows_inv(const InitField* F, const FftPlan* plans,
                                                              const float2* roots, const int* prefix,
                                                              const int* fields, int nf) {
  constexpr int G = Fft::kNc, KP = Fft::kPer;
  int g0, g1;
  block_groups(prefix[nf], g0, g1);
  if (g0 >= g1) return;
  Group cur = group_at(F, prefix, fields, nf, g0, G, kAlongH);
  // fixed lengths: the next (group, variant)'s rows in registers
  float2 pre[KP];
  auto fetch = [&](const Group& q, int v) {
    const float2* rows = q.f->var[v] + (size_t)q.lo * q.f->w;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int e = threadIdx.x + k * kFftThreads, g = e % G, x = e / G;
      pre[k] = e < Fft::kN * G && g < q.cnt ? rows[(size_t)g * q.f->w + x] : make_float2(0.f, 0.f);
    }
  };
  if constexpr (Fft::kN > 0) fetch(cur, 0);
  int plan = -1, n = 0;
  for (int gi = g0; gi < g1; ++gi) {
    const Group here = cur;
    const InitField& f = *here.f;
    use_plan(plans, roots, uni(f.row_plan), G, plan, n);  // n == f.w
    const FftPlan& P = plans[plan];
    const Lv a{0}, b{n * G}, rt{2 * n * G};
    const float scale = 2.f / ((float)f.hp * (float)f.w);
    const int nvar = f.nvar;
    for (int v = 0; v < nvar; ++v) {
      float2* rows = f.var[v] + (size_t)here.lo * f.w;
      if constexpr (Fft::kN > 0) {
#pragma unroll
        for (int k = 0; k < KP; ++k) {
          const int e = threadIdx.x + k * kFftThreads;
          if (e < Fft::kN * G) a[e] = pre[k];
        }
        if (v + 1 < nvar) {
          fetch(here, v + 1);
        } else if (gi + 1 < g1) {
          cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongH);
          fetch(cur, 0);
        }
      } else {
        for (int e = threadIdx.x; e < n * G; e += kFftThreads) {
          const int g = e % G, x = e / G;
          a[e] = g < here.cnt ? rows[(size_t)g * f.w + x] : make_float2(0.f, 0.f);
        }
        if (v + 1 == nvar && gi + 1 < g1) cur = group_at(F, prefix, fields, nf, gi + 1, G, kAlongH);
      }
      __syncthreads();
      const Lv r = Fft::run(a, b, rt, P, +1, roots, rt + n);
      for (int e = threadIdx.x; e < n * G; e += kFftThreads) {
        const int g = e % G, x = e / G;
        const float2 val = r[e];
        if (g < here.cnt) rows[(size_t)g * f.w + x] = make_float2(val.x * scale, val.y * scale);
      }
      __syncthreads();
    }
  }
}

// (5) blend (holographic_init.cpp:171-196, :360-369, :415-422).  Tiles of
// kBlendW x kBlendH pixels; each variant's residual luma is staged with the
// 3 + 1 halo the edge-energy box blur needs.
__global__ void __launch_bounds__(256) k_holo_blend(const InitField* F, const int* tp, int nf) {
  constexpr int RW = kBlendW + 7, RH = kBlendH + 7;  // Y_v region (3 left/up, 4 right/down)
  constexpr int EW = kBlendW + 6, EH = kBlendH + 6;  // edge-energy region
  constexpr int kVar = 3;                            // variants at most
  // every variant is staged at once: S and the carrier are read once per
  // sample, and the tile needs three barriers whatever the variant count
  __shared__ float ys[kVar][RH][RW];
  __shared__ float es[kVar][EH][EW];
  __shared__ float hs[kVar][EH][kBlendW];
  const int bt = xcd_remap(blockIdx.x, gridDim.x);  // neighbouring tiles share an XCD's L2
  const int fi = field_of_tile(tp, nf, bt);
  const InitField& f = F[fi];
  const int lt = bt - tp[fi];
  const int tiles_x = (f.w + kBlendW - 1) / kBlendW;
  const int bx0 = (lt % tiles_x) * kBlendW, by0 = (lt / tiles_x) * kBlendH;
  const int tx = threadIdx.x % kBlendW, ty = threadIdx.x / kBlendW;
  const int px = bx0 + tx, py = by0 + ty;
  const bool in = px < f.w && py < f.h;
  const int nv = min(f.nvar, kVar);
  for (int e = threadIdx.x; e < RH * RW; e += blockDim.x) {
    const int ry = e / RW, rx = e % RW;
    const int y = by0 - 3 + ry, x = bx0 - 3 + rx;
    const bool inside = y >= 0 && y < f.h && x >= 0 && x < f.w;
    const size_t i = inside ? (size_t)y * f.w + x : 0;
    const float sv = inside ? f.s[i] : 0.f;
    const float2 car = inside ? carrier_of(f, y, x) : make_float2(0.f, 0.f);
    for (int v = 0; v < nv; ++v) ys[v][ry][rx] = inside ? sv - re_mul(f.var[v][i], car) : 0.f;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nv * EH * EW; e += blockDim.x) {
    const int v = e / (EH * EW), r = e - v * (EH * EW);
    const int ry = r / EW, rx = r % EW;
    const int y = by0 - 3 + ry, x = bx0 - 3 + rx;
    float val = 0.f;
    if (y >= 0 && y < f.h && x >= 0 && x < f.w) {
      const float gx = x + 1 < f.w ? ys[v][ry][rx + 1] - ys[v][ry][rx] : 0.f;
      const float gy = y + 1 < f.h ? ys[v][ry + 1][rx] - ys[v][ry][rx] : 0.f;
      val = fabsf(gx) + fabsf(gy);
    }
    es[v][ry][rx] = val;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nv * EH * kBlendW; e += blockDim.x) {
    const int v = e / (EH * kBlendW), r = e - v * (EH * kBlendW);
    const int ry = r / kBlendW, cx = r % kBlendW;
    float acc = 0.f;
    for (int k = 0; k < 7; ++k) acc += es[v][ry][cx + k];
    hs[v][ry][cx] = acc;
  }
  __syncthreads();
  float2 num = make_float2(0.f, 0.f);
  float den = 0.f;
  if (in) {
    for (int v = 0; v < nv; ++v) {
      float acc = 0.f;
      for (int k = 0; k < 7; ++k) acc += hs[v][ty + k][tx];
      const float energy = acc * (1.f / 49.f);
      const float wgt = 1.f / (energy + 0.5f);
      const float2 c = f.var[v][(size_t)py * f.w + px];
      num.x += c.x * wgt;
      num.y += c.y * wgt;
      den += wgt;
    }
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
// forward, +1 inverse).  LDS: fft_lds_slots(P, 1) float2.
__global__ void __launch_bounds__(256) k_fft2_rows(const float2* in, float2* out, int h, int w, FftPlan P,
                                                   const float2* roots, int sign) {
  const int y = blockIdx.x;
  const Lv a{0}, b{P.n}, rt{2 * P.n};
  load_root(rt, roots + P.root_off, P.n);
  for (int x = threadIdx.x; x < w; x += blockDim.x) a[x] = in[(size_t)y * w + x];
  __syncthreads();
  const Lv r = fft_lds(a, b, rt, P, sign, 1, roots, rt + P.n);
  for (int x = threadIdx.x; x < w; x += blockDim.x) out[(size_t)y * w + x] = r[x];
}

// The column pass, in place, kColGroup columns per block, each output times
// `scale` (1 forward, 1 / (h w) inverse).  LDS: fft_lds_slots(P, kColGroup)
// float2.
__global__ void __launch_bounds__(256) k_fft2_cols(float2* io, int h, int w, FftPlan P, const float2* roots, int sign,
                                                   float scale) {
  const int x0 = blockIdx.x * kColGroup;
  const int nc = min(kColGroup, w - x0);
  const Lv a{0}, b{P.n * kColGroup}, rt{2 * P.n * kColGroup};
  load_root(rt, roots + P.root_off, P.n);
  for (int e = threadIdx.x; e < P.n * kColGroup; e += blockDim.x) {
    const int c = e % kColGroup, y = e / kColGroup;
    a[e] = c < nc ? io[(size_t)y * w + x0 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const Lv r = fft_lds(a, b, rt, P, sign, kColGroup, roots, rt + P.n);
  for (int e = threadIdx.x; e < P.n * kColGroup; e += blockDim.x) {
    const int c = e % kColGroup, y = e / kColGroup;
    if (c < nc) {
      const float2 v = r[e];
      io[(size_t)y * w + x0 + c] = scale == 1.0f ? v : make_float2(v.x * scale, v.y * scale);
    }
  }
}

// ---- host: the init's FFT passes by transform class ----
namespace {

using FftPassKernel = void (*)(const InitField*, const FftPlan*, const float2*, const int*, const int*, int);

// [pass][class]: rows forward, columns forward, columns inverse, rows
// inverse; class 0 any plan, 1 and 2 the fixed lengths of fft_init_class.
const FftPassKernel kFftPasses[4][3] = {
    {k_fft_rows_fwd<RowsAny>, k_fft_rows_fwd<Rows760>, k_fft_rows_fwd<Rows922>},
    {k_fft_cols_fwd<ColsAny>, k_fft_cols_fwd<Cols274>, k_fft_cols_fwd<Cols320>},
    {k_fft_cols_inv<ColsAny>, k_fft_cols_inv<Cols274>, k_fft_cols_inv<Cols320>},
    {k_fft_rows_inv<RowsAny>, k_fft_rows_inv<Rows760>, k_fft_rows_inv<Rows922>},
};

// The plan runs exactly the stages `radices`, stage `rader_at` (or -1) by
// Rader's algorithm with the inner radices `inner`.
bool plan_matches(const FftPlan& p, int n, std::initializer_list<int> radices, int rader_at,
                  std::initializer_list<int> inner) {
  if (p.n != n || p.nst != (int)radices.size()) return false;
  int s = 0;
  for (int r : radices) {
    if (p.radix[s] != r || (p.rader_of[s] >= 0) != (s == rader_at)) return false;
    ++s;
  }
  if (rader_at < 0) return p.nrader == 0;
  const FftRader& rd = p.rader[p.rader_of[rader_at]];
  if (p.nrader != 1 || p.rader_of[rader_at] != 0 || rd.nst != (int)inner.size()) return false;
  s = 0;
  for (int r : inner)
    if (rd.radix[s++] != r) return false;
  return true;
}

}  // namespace

int fft_init_class(const FftPlan& p, bool rows) {
  if (rows) {
    if (plan_matches(p, 760, {8, 5, 19}, -1, {})) return 1;
    if (plan_matches(p, 922, {2, 461}, 1, {4, 5, 23})) return 2;
  } else {
    if (plan_matches(p, 274, {2, 137}, 1, {8, 17})) return 1;
    if (plan_matches(p, 320, {8, 8, 5}, -1, {})) return 2;
  }
  return 0;
}

void launch_fft_pass(int pass, int cls, int ngroups, size_t lds, hipStream_t st, const InitField* F,
                     const FftPlan* plans, const float2* roots, const int* prefix, const int* fields, int nf) {
  const FftPassKernel k = kFftPasses[pass][cls];
  // as many persistent blocks as the device holds at once (each takes a run
  // of groups); one group per block for the run-time plans
  int blocks = ngroups;
  if (cls != 0) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, size_t>, int> resident;  // (device, pass, class, lds)
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(dev, pass, cls, lds);
    auto it = resident.find(key);
    if (it == resident.end()) {
      int per_cu = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, kFftThreads, lds) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        per_cu = cus = 0;
      it = resident.emplace(key, std::max(per_cu, 1) * std::max(cus, 1)).first;
    }
    blocks = std::min(ngroups, it->second);
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kFftThreads), lds, st, F, plans, roots, prefix, fields, nf);
}

hipError_t fft_allow_lds(int bytes) {
  for (const auto& row : kFftPasses)
    for (FftPassKernel k : row) {
      const hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}

}  // namespace hvdk
