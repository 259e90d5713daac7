// window.hip -- the 3-D window driver on the GPU: hvd-core's
// DecodeFieldWindowWithInits (engine/sequence.cpp:371-1114) over a window of
// equally sized fields, with its noise self-calibration, adaptive temporal
// strength, per-pass pair motion (full search, fast mode's trajectory
// prediction audited by VerifyMotion, the consensus snap), the
// motion-compensated neighbour equations with the coherence and odd-offset
// envelope gates, the PSI init, SynthReference and the anchored joint pass,
// and the output-fidelity luma.
//
// Every per-pixel, per-tile and per-band computation runs on the device: the
// kernels below plus the library's component entry points (EstimateMotion,
// VerifyMotion, MotionCompensatePrev, the warps, ComplexCoherence, BoxBlur and
// the batched solves), all on the caller's stream.  The host keeps the
// scalar decisions hvd-core makes on a handful of values (the median of the
// per-field noise, the p95 of the band energies, the strength map) and the
// order of work: fields of one colour of the Gauss-Seidel sweep (sequence.cpp:
// 1046-1098; fields spaced reach + 1 apart read none of each other's state)
// are solved as one batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "hvd_gpu.h"
#include "hvd_kernels.h"
#include "warp_ops.h"
#include "window.h"

namespace hvdk {

namespace {

constexpr int kSelBlock = 1024;

// Order-preserving map of a float onto an unsigned key (negatives reversed).
__device__ inline unsigned fkey(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float funkey(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// The k-th smallest of v[0 .. n) (std::nth_element's element k), any sign.
// One block of kSelBlock threads; hist: 16 x 256 unsigned in LDS.
__device__ float select_kth(const float* v, int n, int k, unsigned* hist) {
  __shared__ unsigned tot[256];
  __shared__ unsigned s_prefix, s_k;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    s_prefix = 0u;
    s_k = (unsigned)k;
  }
  __syncthreads();
  for (int pass = 3; pass >= 0; --pass) {
    for (int i = threadIdx.x; i < 16 * 256; i += kSelBlock) hist[i] = 0u;
    __syncthreads();
    const unsigned prefix = s_prefix, kk = s_k;
    const int hi_shift = (pass + 1) * 8;
    unsigned* h = hist + wv * 256;
    for (int i = threadIdx.x; i < n; i += kSelBlock) {
      const unsigned u = fkey(v[i]);
      if (hi_shift >= 32 || (u >> hi_shift) == prefix) atomicAdd(&h[(u >> (pass * 8)) & 255u], 1u);
    }
    __syncthreads();
    unsigned cnt = 0u, incl = 0u;
    if (threadIdx.x < 256) {
      for (int w = 0; w < 16; ++w) cnt += hist[w * 256 + threadIdx.x];
      incl = cnt;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      tot[threadIdx.x] = incl;
    }
    __syncthreads();
    if (threadIdx.x < 256) {
      for (int w = 0; w < wv; ++w) incl += tot[w * 64 + 63];
      const unsigned excl = incl - cnt;
      if (excl <= kk && kk < incl) {
        s_k = kk - excl;
        s_prefix = (prefix << 8) | (unsigned)threadIdx.x;
      }
    }
    __syncthreads();
  }
  return funkey(s_prefix);
}

// numpy median (MedianOf, sequence.cpp:28-38) of v[0 .. n) in one block.
__device__ float median_np(const float* v, int n, unsigned* hist) {
#pragma clang fp contract(off)
  if (n <= 0) return 0.f;
  const float m = select_kth(v, n, n / 2, hist);
  if (n % 2) return m;
  __syncthreads();
  const float lo = select_kth(v, n, n / 2 - 1, hist);
  return 0.5f * (m + lo);
}

__device__ inline float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float c_abs(float2 a) { return hypotf(a.x, a.y); }
__device__ inline float re_of(float2 a, float2 b) {
#pragma clang fp contract(off)
  return a.x * b.x - a.y * b.y;
}

}  // namespace

// VerifyMotion's pick (motion.cpp:508-522): the zero vector on ties, else the
// full float prediction.
__global__ void k_verify_pick(int n, const float* pdy, const float* pdx, const float* bp, const float* bz, float* dy,
                              float* dx, float* best) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (bz[i] <= bp[i]) {
    dy[i] = 0.f;
    dx[i] = 0.f;
    best[i] = bz[i];
  } else {
    dy[i] = pdy[i];
    dx[i] = pdx[i];
    best[i] = bp[i];
  }
}

// EstimateNoiseIre (ntsc_geometry.cpp:49-71), one block per field: the
// second differences at the 4-sample period into d, their median (element
// n/2), |d - median| and its element n/4, scaled to sigma.
__global__ void __launch_bounds__(kSelBlock) k_noise_sigma(const float* const* s, int h, int w, float* d_all,
                                                           float* sigma) {
#pragma clang fp contract(off)
  __shared__ unsigned hist[16 * 256];
  const int fi = blockIdx.x;
  const int m = w - 8, n = h * m;
  if (m <= 0 || h <= 0) {
    if (threadIdx.x == 0) sigma[fi] = 0.f;
    return;
  }
  const float* S = s[fi];
  float* d = d_all + (size_t)fi * n;
  for (int i = threadIdx.x; i < n; i += kSelBlock) {
    const int y = i / m, x = i % m;
    const float* r = S + (size_t)y * w + x;
    d[i] = r[0] - 2.0f * r[4] + r[8];
  }
  __syncthreads();
  const float med = select_kth(d, n, n / 2, hist);
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kSelBlock) d[i] = fabsf(d[i] - med);
  __syncthreads();
  const float q = select_kth(d, n, n / 4, hist);
  if (threadIdx.x == 0) sigma[fi] = q / 0.3186f / sqrtf(6.0f);
}

// The adaptive-strength scan's band energies (sequence.cpp:560-630): block
// (pair, band) over the decimated samples of that band's 8 sample rows; a
// sample counts where the box-4 means of the pair agree within nr_eps.
__global__ void __launch_bounds__(256) k_amb_bands(const float* const* s, const float* const* car, int stride, int h,
                                                   int w, float nr_eps, int nbands, double* acc, int* cnt) {
#pragma clang fp contract(off)
  __shared__ double sd[256];
  __shared__ int sn[256];
  const int pr = blockIdx.x / nbands, band = blockIdx.x % nbands;
  const float *Sj = s[pr], *Sk = s[pr + stride];
  const float2 *cj = (const float2*)car[pr], *ck = (const float2*)car[pr + stride];
  const int nx = (w - 7 - 4 + 3) / 4;  // x = 4, 8, ... while x + 7 < w
  const int rows = 8;
  double e_sum = 0.0;
  int n_sum = 0;
  const float kw[7] = {1.f, 2.f, 3.f, 4.f, 3.f, 2.f, 1.f};
  for (int t = threadIdx.x; t < rows * nx; t += 256) {
    const int yi = band * 8 + t / nx;
    const int y = 2 + 3 * yi;
    if (y >= h - 2) continue;
    const int x = 4 + 4 * (t % nx);
    float mj = 0.f, mk = 0.f;
    for (int k4 = 0; k4 < 4; ++k4) {
      mj += Sj[(size_t)y * w + x + k4];
      mk += Sk[(size_t)y * w + x + k4];
    }
    if (fabsf(mj - mk) * 0.25f > nr_eps) continue;
    float2 dj = make_float2(0.f, 0.f), dk = make_float2(0.f, 0.f);
    for (int k = 0; k < 7; ++k) {
      const size_t i = (size_t)y * w + x + k;
      const float aj = kw[k] * Sj[i], ak = kw[k] * Sk[i];
      const float2 c1 = cj[i], c2 = ck[i];
      dj.x += aj * c1.x;
      dj.y += aj * -c1.y;
      dk.x += ak * c2.x;
      dk.y += ak * -c2.y;
    }
    dj = make_float2(dj.x * (1.0f / 16.0f), dj.y * (1.0f / 16.0f));
    dk = make_float2(dk.x * (1.0f / 16.0f), dk.y * (1.0f / 16.0f));
    const float2 e = c_sub(dj, dk);
    e_sum += 0.25 * (double)(e.x * e.x + e.y * e.y);
    n_sum += 1;
  }
  sd[threadIdx.x] = e_sum;
  sn[threadIdx.x] = n_sum;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sd[threadIdx.x] += sd[threadIdx.x + o];
      sn[threadIdx.x] += sn[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    acc[blockIdx.x] = sd[0];
    cnt[blockIdx.x] = sn[0];
  }
}

// Y = S - Re[chi c] (the init's luma when none is given; the output-fidelity
// luma, sequence.cpp:1102-1110).
__global__ void k_fidelity(size_t n, const float* s, const float2* chi, const float2* car, float* y) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = s[i] - re_of(chi[i], car[i]);
}

// Field j's per-tile trajectory prediction (sequence.cpp:702-724): the numpy
// median over the measured offsets of (dy - h_o) / o and dx / o.  dy, dx:
// m planes of nt tiles; ho, off: m entries.
__global__ void k_traj_predict(int m, int nt, TrajTab T, float* vy, float* vx) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  float sy[kMaxNeighbors], sx[kMaxNeighbors];
  for (int j = 0; j < m; ++j) {
    const float o = (float)T.off[j];
    sy[j] = (T.dy[j][t] - T.ho[j]) / o;
    sx[j] = T.dx[j][t] / o;
  }
  // insertion sorts (m <= 16)
  for (int a = 1; a < m; ++a)
    for (int b = a; b > 0 && sy[b] < sy[b - 1]; --b) {
      const float z = sy[b];
      sy[b] = sy[b - 1];
      sy[b - 1] = z;
    }
  for (int a = 1; a < m; ++a)
    for (int b = a; b > 0 && sx[b] < sx[b - 1]; --b) {
      const float z = sx[b];
      sx[b] = sx[b - 1];
      sx[b - 1] = z;
    }
  float my = sy[m / 2], mx = sx[m / 2];
  if (m % 2 == 0) {
    my = 0.5f * (my + sy[m / 2 - 1]);
    mx = 0.5f * (mx + sx[m / 2 - 1]);
  }
  vy[t] = my;
  vx[t] = mx;
}

// The predicted vector of offset o: o v + h_o.
__global__ void k_traj_extrapolate(int nt, const float* vy, const float* vx, float o, float ho, float* pdy,
                                   float* pdx) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  pdy[t] = o * vy[t] + ho;
  pdx[t] = o * vx[t];
}

namespace {

// FitTrajectory (motion.cpp:532-562) at tile t: the numpy median of the
// confident (conf > 0.15, offset != 0) (dy - h) / o and dx / o; 0 where no
// measurement qualifies.
__device__ inline void traj_fit(int m, int t, const TrajTab& T, float& vy, float& vx) {
#pragma clang fp contract(off)
  float sy[kMaxNeighbors], sx[kMaxNeighbors];
  int k = 0;
  for (int j = 0; j < m; ++j) {
    if (T.off[j] == 0 || T.conf[j][t] <= 0.15f) continue;
    const float o = (float)T.off[j];
    sy[k] = (T.dy[j][t] - T.ho[j]) / o;
    sx[k] = T.dx[j][t] / o;
    ++k;
  }
  vy = 0.f;
  vx = 0.f;
  if (k == 0) return;
  for (int a = 1; a < k; ++a)
    for (int b = a; b > 0 && sy[b] < sy[b - 1]; --b) {
      const float z = sy[b];
      sy[b] = sy[b - 1];
      sy[b - 1] = z;
    }
  for (int a = 1; a < k; ++a)
    for (int b = a; b > 0 && sx[b] < sx[b - 1]; --b) {
      const float z = sx[b];
      sx[b] = sx[b - 1];
      sx[b - 1] = z;
    }
  vy = sy[k / 2];
  vx = sx[k / 2];
  if (k % 2 == 0) {
    vy = 0.5f * (vy + sy[k / 2 - 1]);
    vx = 0.5f * (vx + sx[k / 2 - 1]);
  }
}

}  // namespace

// FitTrajectory over m measurements of nt tiles into vy, vx.
__global__ void k_traj_fit(int m, int nt, TrajTab T, float* vy, float* vx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  float a = 0.f, b = 0.f;
  if (m > 0) traj_fit(m, t, T, a, b);
  vy[t] = a;
  vx[t] = b;
}

// TrajectorySnap (motion.cpp:564-601) over one field's m measurements: the
// fit, the consensus count of those within 1.5 px, and the snap where >= 3
// agree.  In place on dy, dx.
__global__ void k_traj_snap(int m, int nt, TrajTab T) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt || m < 2) return;
  float vy, vx;
  traj_fit(m, t, T, vy, vx);
  float agree = 0.f;
  for (int j = 0; j < m; ++j) {
    const float o = (float)T.off[j];
    const float py = o * vy + T.ho[j], px = o * vx;
    const float dev = fabsf(T.dy[j][t] - py) + fabsf(T.dx[j][t] - px);
    if (dev <= 1.5f) agree += 1.f;
  }
  if (agree < 3.f) return;
  for (int j = 0; j < m; ++j) {
    const float o = (float)T.off[j];
    const float py = o * vy + T.ho[j], px = o * vx;
    const float dev = fabsf(T.dy[j][t] - py) + fabsf(T.dx[j][t] - px);
    if (dev <= 1.5f) {
      T.dy[j][t] = py;
      T.dx[j][t] = px;
    }
  }
}

// Copies of n floats from P.src[b] to P.dst[b], block row b.
__global__ void __launch_bounds__(256) k_copy_tiles(int n, CopyTab P) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) P.dst[blockIdx.y][i] = P.src[blockIdx.y][i];
}

// The envelope's per-pixel terms (sequence.cpp:817-852): the one-sided max
// vertical difference energy and the second-difference energy of the
// decoded baseband (Y, chi).
__global__ void k_env_terms(int h, int w, const float* Y, const float2* chi, float* gmag, float* lapmag) {
#pragma clang fp contract(off)
  const size_t n = (size_t)h * w;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int y = (int)(i / w), x = (int)(i % w);
    const size_t im = (size_t)max(0, y - 1) * w + x, ip = (size_t)min(h - 1, y + 1) * w + x;
    const float y0 = Y[i], ym = Y[im], yp = Y[ip];
    const float2 c0 = chi[i], cm = chi[im], cp = chi[ip];
    const float dyu = y0 - ym, dyd = yp - y0;
    const float dcu = c_abs(c_sub(c0, cm)), dcd = c_abs(c_sub(cp, c0));
    const float mu = dyu * dyu + dcu * dcu;
    const float md = dyd * dyd + dcd * dcd;
    gmag[i] = fmaxf(mu, md);
    const float ly = y0 - 0.5f * (ym + yp);
    const float2 hs = make_float2(cm.x + cp.x, cm.y + cp.y);
    const float lc = c_abs(c_sub(c0, make_float2(0.5f * hs.x, 0.5f * hs.y)));
    lapmag[i] = ly * ly + lc * lc;
  }
}

// The 5-tap sliding RMS along each row (sequence.cpp:855-888), in hvd-core's
// order of adds and subtracts (a float running sum: the first three samples,
// then per step the sample entering at x + 2 and the one leaving at x - 3;
// outside the row 0 is added, which leaves the sum, never -0, as it is), for
// a batch of planes: plane b is src[b] -> out[b].  Block: kSmR rows of one
// plane, threads 0 .. kSmR - 1 running one row's sum each; the row segments
// come through LDS in chunks of 64 columns, every thread loading two columns
// of every row with all of a chunk's loads in flight at once, and the next
// chunk's loads issued before the current chunk is summed; the chunk's RMS is
// formed and stored column by column (coalesced).  Grid: (ceil(h / kSmR),
// planes).
constexpr int kSmR = 16;
__global__ void __launch_bounds__(64) k_smooth_rows(int h, int w, EnvRows P) {
#pragma clang fp contract(off)
  constexpr int C = 64, L = 3, SW = C + 5;
  __shared__ float sm[kSmR][SW + 2];
  __shared__ float sa[kSmR][C + 1];
  const int t = threadIdx.x;
  const float* in = P.src[blockIdx.y];
  float* o = P.out[blockIdx.y];
  const int row0 = blockIdx.x * kSmR, y = row0 + t;
  float acc = 0.f;
  // row j's segment of the chunk at x0: thread t holds columns t and t + 64
  float v[kSmR][2];
  auto fetch = [&](int x0) {
#pragma unroll
    for (int u = 0; u < kSmR; ++u) {
      const int rj = row0 + u;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int k = t + 64 * q, col = x0 - L + k;
        v[u][q] = (k < SW && rj < h && col >= 0 && col < w) ? in[(size_t)rj * w + col] : 0.f;
      }
    }
  };
  fetch(0);
  for (int x0 = 0; x0 < w; x0 += C) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kSmR; ++u)
#pragma unroll
      for (int q = 0; q < 2; ++q)
        if (t + 64 * q < SW) sm[u][t + 64 * q] = v[u][q];
    if (x0 + C < w) fetch(x0 + C);
    __syncthreads();
    const int xend = min(x0 + C, w);
    if (t < kSmR && y < h) {
      int x = x0;
      if (x == 0) {
        for (int i = 0; i < min(3, w); ++i) acc += sm[t][i + L];
        sa[t][0] = acc;
        ++x;
      }
#pragma unroll 8
      for (; x < xend; ++x) {
        acc += sm[t][x + 2 - x0 + L];
        acc -= sm[t][x - 3 - x0 + L];
        sa[t][x - x0] = acc;
      }
    }
    __syncthreads();
    const int x = x0 + t;
    if (x < w) {
      const float n_taps = (float)(min(x + 2, w - 1) - max(x - 2, 0) + 1);
      for (int j = 0; j < kSmR; ++j) {
        const int rj = row0 + j;
        if (rj < h) o[(size_t)rj * w + x] = sqrtf(sa[j][t] / n_taps);
      }
    }
  }
}

// The envelope's noise floor (sequence.cpp:889-916), one block per field:
// the median of every seventh vgrad sample, (1.5 median)^2 into floor2; then
// (k_env_apply) per pixel vgrad <- sqrt(max(0, vgrad^2 - floor2)) and lapsm
// = 1 - t^2 with t = clamp(smoothed second difference / vgrad, 0, 1).
__global__ void __launch_bounds__(kSelBlock) k_env_finish(int n, EnvFinish P) {
#pragma clang fp contract(off)
  __shared__ unsigned hist[16 * 256];
  const int fi = blockIdx.x;
  const float* v = P.vg_in[fi];
  float* smp = P.sample[fi];
  const int ns = (n + 6) / 7;
  for (int i = threadIdx.x; i < ns; i += kSelBlock) smp[i] = v[(size_t)i * 7];
  __syncthreads();
  const float med = median_np(smp, ns, hist);
  if (threadIdx.x == 0) P.floor2[fi] = 2.25f * med * med;
}

// The per-pixel part of the envelope (sequence.cpp:917-939) for every field
// of the colour: grid (pixel blocks, fields).
__global__ void __launch_bounds__(256) k_env_apply(int n, EnvFinish P) {
#pragma clang fp contract(off)
  const int fi = blockIdx.y;
  const float fl2 = P.floor2[fi];
  const float* v = P.vg_in[fi];
  float* vg = P.vgrad[fi];
  float* lp = P.lap[fi];
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float g0 = v[i];
    const float g = sqrtf(fmaxf(0.0f, g0 * g0 - fl2));
    vg[i] = g;
    const float t = g > 1e-6f ? fminf(fmaxf(lp[i] / g, 0.0f), 1.0f) : 0.0f;
    lp[i] = 1.0f - t * t;
  }
}

// Complex plane <-> two real planes (the warps take real planes).
__global__ void k_split_cplx(size_t n, const float2* z, float* re, float* im) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    re[i] = z[i].x;
    im[i] = z[i].y;
  }
}
__global__ void k_join_cplx(size_t n, const float* re, const float* im, float2* z) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    z[i] = make_float2(re[i], im[i]);
}

// The neighbour gates on its confidence (sequence.cpp:965-1015): the InSAR
// coherence (1 - a) + a gamma when gamma is given, and for an odd offset
// max(floor lapsm, e2 / (e2 + vgrad^2)) when vgrad is given.
__global__ void k_conf_gates(size_t n, float* conf, const float* gamma, float a, const float* vgrad,
                             const float* lapsm, float e2, float floor_max) {
#pragma clang fp contract(off)
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float c = conf[i];
    if (gamma) c *= (1.0f - a) + a * gamma[i];
    if (vgrad) {
      const float b = vgrad[i];
      const float fl = floor_max * lapsm[i];
      c *= fmaxf(fl, e2 / (e2 + b * b));
    }
    conf[i] = c;
  }
}

// SynthReference (sequence.cpp:52-150): the accumulators start at the
// field's own decode; each neighbour adds w (Yw, cw) with the Geman-McClure
// weight w = conf eps^2 / (|dY|^2 + |dchi|^2 + eps^2).
__global__ void k_synth_init(size_t n, const float* Y, const float2* chi, float* aY, float* aCr, float* aCi,
                             float* aW) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    aY[i] = Y[i];
    aCr[i] = chi[i].x;
    aCi[i] = chi[i].y;
    aW[i] = 1.0f;
  }
}
__global__ void k_synth_accum(size_t n, const float* Y, const float2* chi, const float* conf, const float* Yw,
                              const float* wre, const float* wim, float eps_b, float* aY, float* aCr, float* aCi,
                              float* aW) {
#pragma clang fp contract(off)
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float dy = Yw[i] - Y[i];
    const float2 dc = c_sub(make_float2(wre[i], wim[i]), chi[i]);
    const float d2 = dy * dy + (dc.x * dc.x + dc.y * dc.y);
    const float w = conf[i] * eps_b * eps_b / (d2 + eps_b * eps_b);
    aY[i] += w * Yw[i];
    aCr[i] += w * wre[i];
    aCi[i] += w * wim[i];
    aW[i] += w;
  }
}
__global__ void k_synth_finish(size_t n, const float* s, const float2* car, const float* aY, const float* aCr,
                               const float* aCi, const float* aW, float* luma, float2* chroma, float* resid) {
#pragma clang fp contract(off)
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float l = aY[i] / aW[i];
    const float2 c = make_float2(aCr[i] / aW[i], aCi[i] / aW[i]);
    luma[i] = l;
    chroma[i] = c;
    const float s_hat = l + re_of(c, car[i]);
    resid[i] = fabsf(s[i] - s_hat);
  }
}
__global__ void k_synth_weight(size_t n, const float* r, float eps_b, float* weight) {
#pragma clang fp contract(off)
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    weight[i] = eps_b * eps_b / (r[i] * r[i] + eps_b * eps_b);
}

// PsiClosedForm (sequence.cpp:160-226): per pixel the 3 x 3 weighted normal
// equations over the field's own observation and its neighbours', solved by
// Cramer in FP64, blended with the fallback chi by the neighbour weight.
__global__ void k_psi(size_t n, int nn, const float* s, const float2* car, const float* const* sw,
                      const float2* const* cw, const float* const* conf, float2* chi) {
#pragma clang fp contract(off)
  const float kEpsT = 6.0f;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double a11 = 0, a12 = 0, a13 = 0, a22 = 0, a23 = 0, a33 = 0, b1 = 0, b2 = 0, b3 = 0, wsum = 0;
    const float2 fb = chi[i];
    auto add = [&](float Sk, float2 ck, float wk) {
      const double rc = ck.x, ic = -ck.y;
      a11 += wk;
      a12 += wk * rc;
      a13 += wk * ic;
      a22 += wk * rc * rc;
      a23 += wk * rc * ic;
      a33 += wk * ic * ic;
      const float wsk = wk * Sk;
      b1 += wsk;
      b2 += wsk * rc;
      b3 += wsk * ic;
    };
    const float2 c0 = car[i];
    add(s[i], c0, 1.0f);
    for (int j = 0; j < nn; ++j) {
      const float2 cj = cw[j][i];
      const float2 dc = c_sub(c0, cj);
      const float rt = (sw[j][i] - s[i]) + re_of(fb, dc);
      const float wk = conf[j][i] * kEpsT * kEpsT / (rt * rt + kEpsT * kEpsT);
      add(sw[j][i], cj, wk);
      wsum += wk;
    }
    const double det = a11 * (a22 * a33 - a23 * a23) - a12 * (a12 * a33 - a23 * a13) + a13 * (a12 * a23 - a22 * a13);
    const double a11c = fmax(a11, 1e-9);
    const bool ok = det > 1e-3 * a11c * a11c * a11c * 0.01;
    const double d = ok ? det : 1.0;
    const double q1 = (a11 * (b2 * a33 - b3 * a23) - b1 * (a12 * a33 - a23 * a13) + a13 * (a12 * b3 - b2 * a13)) / d;
    const double q2 = (a11 * (a22 * b3 - a23 * b2) - a12 * (a12 * b3 - b2 * a13) + b1 * (a12 * a23 - a22 * a13)) / d;
    const float t = ok ? fminf(fmaxf((float)(wsum / 2.0), 0.0f), 1.0f) : 0.0f;
    const float pr = (float)q1, pi = (float)q2;
    chi[i] = make_float2(t * pr + (1.0f - t) * fb.x, t * pi + (1.0f - t) * fb.y);
  }
}

// ---- the neighbour equations of a colour, batched -------------------------
// One job per (field j, offset o) with its neighbour k = j + o: hvd-core's
// MotionCompensatePrev of S_k and c_k by the snapped tile motion (temporal.cpp:
// 198-249), the coherence gate on chi_j against the bilinear warp of chi_k
// (:140-196, sequence.cpp:965-995) and the odd-offset envelope gate
// (:996-1015), every job of the colour in the same launches.  The arithmetic
// is the single-plane kernels' (warp_ops.h).

// (1) per tile: the squared confidence (fast mode's map).
__global__ void k_nb_square(int nt, const NbJob* J) {
  const NbJob& j = J[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nt) j.sq[i] = j.mconf[i] * j.mconf[i];
}

// (2) per tile: the outlier snap of the vectors (and, in fast mode, of the
// squared confidence map taken as both components, as hvd-core does).
__global__ void k_nb_snap(int th, int tw, int fast, const NbJob* J) {
  const NbJob& j = J[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= th * tw) return;
  float oy, ox;
  wops::vpp_snap(th, tw, j.mdy, j.mdx, i, oy, ox);
  j.sy[i] = oy;
  j.sx[i] = ox;
  if (fast) {
    float cy, cx;
    wops::vpp_snap(th, tw, j.sq, j.sq, i, cy, cx);
    j.sc[i] = cy;
  }
}

// (3) per pixel: the vectors, the warped composite and carrier, the
// confidence (fast: the interpolated map; else the nearest tile's square for
// the radius-8 blur), and with the gate the coherence terms of chi_j against
// the warped chi_k.
__global__ void k_nb_warp(int h, int w, int th, int tw, int tile, int fast, int gate, const NbJob* J) {
#pragma clang fp contract(off)
  const NbJob& j = J[blockIdx.y];
  const size_t n = (size_t)h * w;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int y = (int)(i / w), x = (int)(i % w);
  const float dy = wops::vpp_interp(th, tw, tile, j.sy, y, x), dx = wops::vpp_interp(th, tw, tile, j.sx, y, x);
  const size_t src = wops::warp_tiles_src(h, w, y, x, dy, dx);
  j.comp[i] = j.s_k[src];
  j.car[i] = j.car_k[src];
  if (fast) {
    j.conf[i] = wops::vpp_interp(th, tw, tile, j.sc, y, x);
  } else {
    const float cf = wops::tile_at(j.mconf, tw, min(y / tile, th - 1), min(x / tile, tw - 1));
    j.cn[i] = cf * cf;
  }
  if (gate) {
    const float2 cw = make_float2(wops::warp_bilinear_c(h, w, j.chi_k, 0, dy, dx, 0.0f, y, x),
                                  wops::warp_bilinear_c(h, w, j.chi_k, 1, dy, dx, 0.0f, y, x));
    float a, b, c, d;
    wops::coh_terms(j.chi_j[i], cw, a, b, c, d);
    j.coh[i] = a;
    j.coh[n + i] = b;
    j.coh[2 * n + i] = c;
    j.coh[3 * n + i] = d;
  }
}

// (4) per pixel: the confidence after the gates (k_conf_gates' arithmetic),
// from the blurred confidence (slow mode) and the blurred coherence terms.
__global__ void k_nb_gates(int h, int w, int fast, int gate, float a, float e2, float floor_max, const NbJob* J) {
#pragma clang fp contract(off)
  const NbJob& j = J[blockIdx.y];
  const size_t n = (size_t)h * w;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float c = fast ? j.conf[i] : j.cb[i];
  if (gate) {
    const float g = wops::coh_out(j.cb4[i], j.cb4[n + i], j.cb4[2 * n + i], j.cb4[3 * n + i]);
    c *= (1.0f - a) + a * g;
  }
  if (j.vgrad) {
    const float b = j.vgrad[i];
    const float fl = floor_max * j.lapsm[i];
    c *= fmaxf(fl, e2 / (e2 + b * b));
  }
  j.conf[i] = c;
}

}  // namespace hvdk

// ---------------------------------------------------------------------------
// The host side of the driver.
// ---------------------------------------------------------------------------
namespace hvd_window {

namespace {

using namespace hvdk;

inline unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 8192); }

constexpr int kJobChunk = 32;  // neighbour equations batched per launch

struct Carver {
  size_t off = 0;
  size_t take(size_t b) {
    const size_t o = off;
    off += (b + 255) & ~size_t(255);
    return o;
  }
};

// numpy median of host floats (MedianOf, sequence.cpp:28-38).
float median_host(std::vector<float> v) {
  if (v.empty()) return 0.f;
  std::sort(v.begin(), v.end());
  const size_t k = v.size() / 2;
  float m = v[k];
  if (v.size() % 2 == 0) m = 0.5F * (m + v[k - 1]);
  return m;
}

struct Ctx {
  hvd_gpu_ctx* ctx;
  hipStream_t st;
  int rc = 0;
  bool ok(int r) {
    if (r != HVD_GPU_OK && rc == 0) rc = r;
    return rc == 0;
  }
  bool launched() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess && rc == 0) rc = fail(ctx, HVD_GPU_ERR_HIP, hipGetErrorString(e));
    return rc == 0;
  }
};

// One motion field on the device (tiles).
struct MF {
  float *dy, *dx, *conf;
};

}  // namespace

int drive(hvd_gpu_ctx* ctx, const hvd_gpu_plane* F, int nf, const int* par, int is_pal, const int* y_given,
          const hvd_solver_params* p_in, const hvd_window_params* wp, hvd_window_diag* diag, hipStream_t st) {
  if (!F || !par || !p_in || !wp || nf < 1) return fail(ctx, HVD_GPU_ERR_INVALID, "null argument or empty window");
  const int h = F[0].lines, w = F[0].width;
  for (int j = 0; j < nf; ++j) {
    if (F[j].lines != h || F[j].width != w) return fail(ctx, HVD_GPU_ERR_INVALID, "all fields of a window share one shape");
    if (!F[j].s || !F[j].carrier || !F[j].chi || !F[j].y)
      return fail(ctx, HVD_GPU_ERR_INVALID, "each field needs s, an explicit carrier, chi and y planes");
  }
  if (h < 4 || w < 4) return fail(ctx, HVD_GPU_ERR_INVALID, "fields must be at least 4 x 4");
  hvd_window_params c = *wp;
  if (c.mc_tile < 1 || c.mc_tile > 256) return fail(ctx, HVD_GPU_ERR_INVALID, "mc_tile outside 1..256");
  if (c.mc_search < 0 || c.mc_search > 64) return fail(ctx, HVD_GPU_ERR_INVALID, "mc_search outside 0..64");
  if (c.nr_radius < 0 || c.nr_radius > 16) return fail(ctx, HVD_GPU_ERR_INVALID, "nr_radius outside 0..16");
  hvd_solver_params p = *p_in;
  p.is_pal = is_pal ? 1 : 0;
  Ctx X{ctx, st};
  const size_t np = (size_t)h * w;
  const int tile = c.mc_tile;
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw;

  // offsets (sequence.cpp:452-476) and the PAL defaults
  std::vector<int> offs;
  if (c.extended_temporal)
    offs = is_pal ? (c.bidirectional ? std::vector<int>{-4, -3, -2, -1, 1, 2, 3, 4} : std::vector<int>{-4, -3, -2, -1})
                  : (c.bidirectional ? std::vector<int>{-3, -2, -1, 1, 2, 3} : std::vector<int>{-3, -2, -1});
  else
    offs = c.bidirectional ? std::vector<int>{-2, -1, 1, 2} : std::vector<int>{-2, -1};
  const int no = (int)offs.size();
  if (is_pal) {
    if ((float)c.coherence_gate == 0.6F) c.coherence_gate = 1.0;
    if (c.nr_radius == 2) c.nr_radius = 4;
  }
  // the motion cache's offsets: the neighbour offsets first (their slot is
  // their index in offs), then the anchor's radius beyond them
  std::vector<int> coffs = offs;
  for (int o = -c.nr_radius; o <= c.nr_radius; ++o)
    if (o != 0 && std::find(coffs.begin(), coffs.end(), o) == coffs.end()) coffs.push_back(o);
  const int nco = (int)coffs.size();
  const int amb_stride = is_pal ? 4 : 2;
  const int ny_amb = h > 4 ? (h - 4 + 2) / 3 : 0;
  const int nbands = (ny_amb + 7) / 8;
  const int npairs = std::max(0, nf - amb_stride);

  // ---- device memory -------------------------------------------------------
  Carver cv;
  const size_t o_sig = cv.take(sizeof(float) * nf);
  const size_t o_d = cv.take(sizeof(float) * nf * (size_t)h * std::max(1, w - 8));
  const size_t o_acc = cv.take(sizeof(double) * (size_t)std::max(1, npairs * nbands));
  const size_t o_cnt = cv.take(sizeof(int) * (size_t)std::max(1, npairs * nbands));
  const size_t o_mc = cv.take(sizeof(float) * 3 * nt * (size_t)nf * nco);     // motion cache
  const size_t o_pm = cv.take(sizeof(float) * 2 * nt * (size_t)nf * no);      // snapped copies
  const size_t o_pred = cv.take(sizeof(float) * 4 * nt * (size_t)nf);
  const size_t o_nb = cv.take(sizeof(float) * 4 * np * (size_t)nf * no);      // comp, car (2), conf
  const size_t o_env = cv.take(sizeof(float) * 2 * np * (size_t)nf);          // vgrad, lapsm
  const size_t o_envt = cv.take(sizeof(float) * 2 * np * (size_t)nf);         // smoothed terms
  const size_t o_envr = cv.take(sizeof(float) * 2 * np * (size_t)nf);         // raw terms
  const size_t o_envs = cv.take(sizeof(float) * (np / 7 + 1) * (size_t)nf);   // median samples
  const size_t o_anc = cv.take(sizeof(float) * 4 * np * (size_t)nf);          // anchor luma, chroma (2), weight
  const size_t o_tmp = cv.take(sizeof(float) * 12 * np);
  const size_t o_smf = cv.take(sizeof(float) * 3 * nt);                       // SynthReference's own estimate
  // the batched neighbour equations: up to kJobChunk jobs at a time
  const size_t job_floats = 4 * (size_t)nt + 14 * np;
  const size_t o_jobs = cv.take(sizeof(NbJob) * kJobChunk);
  const size_t o_jmd = cv.take(sizeof(MotionDesc) * (3 * kJobChunk + 2));
  const size_t o_jscr = cv.take(sizeof(float) * job_floats * kJobChunk);
  const size_t o_ptr = cv.take(sizeof(void*) * 64 * (size_t)(nf + 4));       // pointer tables
  const size_t o_envfl = cv.take(sizeof(float) * kEnvMax);                    // envelope noise floors
  char* base = pool(ctx, cv.off, st);
  if (!base) return HVD_GPU_ERR_NOMEM;
  float* sig = (float*)(base + o_sig);
  float* dnoise = (float*)(base + o_d);
  double* bacc = (double*)(base + o_acc);
  int* bcnt = (int*)(base + o_cnt);
  auto mcache = [&](int j, int oi) {
    float* q = (float*)(base + o_mc) + ((size_t)j * nco + oi) * 3 * nt;
    return MF{q, q + nt, q + 2 * nt};
  };
  auto pmf = [&](int j, int oi) {
    float* q = (float*)(base + o_pm) + ((size_t)j * no + oi) * 2 * nt;
    return MF{q, q + nt, nullptr};
  };
  float* pred = (float*)(base + o_pred);
  auto nbp = [&](int j, int oi) { return (float*)(base + o_nb) + ((size_t)j * no + oi) * 4 * np; };
  auto envp = [&](int j) { return (float*)(base + o_env) + (size_t)j * 2 * np; };
  auto envt = [&](int j) { return (float*)(base + o_envt) + (size_t)j * 2 * np; };
  auto envr = [&](int j) { return (float*)(base + o_envr) + (size_t)j * 2 * np; };
  auto envs = [&](int j) { return (float*)(base + o_envs) + (size_t)j * (np / 7 + 1); };
  auto ancp = [&](int j) { return (float*)(base + o_anc) + (size_t)j * 4 * np; };
  float* tmp = (float*)(base + o_tmp);
  auto T = [&](int k) { return tmp + (size_t)k * np; };
  float* smf = (float*)(base + o_smf);
  void** ptrs = (void**)(base + o_ptr);
  float* envfl = (float*)(base + o_envfl);
  auto cfl = [](const float* q) { return (const float2*)q; };

  // ---- noise self-calibration (sequence.cpp:439-449) ------------------------
  {
    std::vector<const float*> sp(nf);
    for (int j = 0; j < nf; ++j) sp[j] = F[j].s;
    if (hipMemcpyAsync(ptrs, sp.data(), sizeof(void*) * nf, hipMemcpyHostToDevice, st) != hipSuccess)
      return fail(ctx, HVD_GPU_ERR_HIP, "pointer upload failed");
    hipLaunchKernelGGL(k_noise_sigma, dim3(nf), dim3(1024), 0, st, (const float* const*)ptrs, h, w, dnoise, sig);
    if (!X.launched()) return X.rc;
  }
  std::vector<float> hsig(nf);
  if (hipMemcpyAsync(hsig.data(), sig, sizeof(float) * nf, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(ctx, HVD_GPU_ERR_HIP, "noise readback failed");
  const float sigma = median_host(hsig);
  float temporal_eps = (float)c.temporal_eps, nr_eps = (float)c.nr_eps;
  if (temporal_eps <= 0.0F) temporal_eps = std::clamp(7.0F * sigma, 4.0F, 20.0F);
  if (nr_eps <= 0.0F) nr_eps = std::clamp(3.0F * sigma, 3.0F, 12.0F);

  // ---- the working state: Y and chi in the callers' planes -----------------
  for (int j = 0; j < nf; ++j)
    if (!y_given || !y_given[j]) {
      hipLaunchKernelGGL(k_fidelity, dim3(grid_for(np)), dim3(256), 0, st, np, F[j].s, cfl(F[j].chi),
                         cfl(F[j].carrier), F[j].y);
      if (!X.launched()) return X.rc;
    }

  // ---- adaptive strength (sequence.cpp:526-652) ------------------------------
  float strength = (float)c.temporal_strength;
  float amb = 0.F;
  if (strength == 0.0F && p.cg_iterations > 0 && nf > amb_stride) {
    if (npairs > 0 && nbands > 0) {
      std::vector<const float*> sp(nf), cp(nf);
      for (int j = 0; j < nf; ++j) {
        sp[j] = F[j].s;
        cp[j] = F[j].carrier;
      }
      if (hipMemcpyAsync(ptrs, sp.data(), sizeof(void*) * nf, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipMemcpyAsync(ptrs + nf, cp.data(), sizeof(void*) * nf, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(ctx, HVD_GPU_ERR_HIP, "pointer upload failed");
      hipLaunchKernelGGL(k_amb_bands, dim3(npairs * nbands), dim3(256), 0, st, (const float* const*)ptrs,
                         (const float* const*)(ptrs + nf), amb_stride, h, w, nr_eps, nbands, bacc, bcnt);
      if (!X.launched()) return X.rc;
    }
    std::vector<double> ha((size_t)npairs * nbands);
    std::vector<int> hn((size_t)npairs * nbands);
    if (!ha.empty() &&
        (hipMemcpyAsync(ha.data(), bacc, sizeof(double) * ha.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipMemcpyAsync(hn.data(), bcnt, sizeof(int) * hn.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
         hipStreamSynchronize(st) != hipSuccess))
      return fail(ctx, HVD_GPU_ERR_HIP, "band readback failed");
    std::vector<float> ambs;
    for (int pr = 0; pr < npairs; ++pr) {
      double acc = 0.0;
      long n = 0;
      std::vector<float> brms;
      for (int b = 0; b < nbands; ++b) {
        acc += ha[(size_t)pr * nbands + b];
        n += hn[(size_t)pr * nbands + b];
        if (hn[(size_t)pr * nbands + b] >= 8)
          brms.push_back((float)(ha[(size_t)pr * nbands + b] / hn[(size_t)pr * nbands + b]));
      }
      if (n > 0) {
        if (!brms.empty()) {
          const size_t k95 = std::min(brms.size() - 1, (size_t)(0.95 * brms.size()));
          std::nth_element(brms.begin(), brms.begin() + k95, brms.end());
          ambs.push_back(std::sqrt(brms[k95]));
        } else {
          ambs.push_back(std::sqrt((float)(acc / n)));
        }
      }
    }
    amb = median_host(ambs);
    const float sigma_d = nr_eps > 0.0F ? nr_eps / 3.0F : 1.0F;
    const float noise_floor = sigma_d * 0.29F;
    amb = std::sqrt(std::max(0.0F, amb * amb - noise_floor * noise_floor));
    strength = std::clamp(0.5F * amb - 0.05F, 0.0F, 1.5F);
  }
  if (strength < 0.0F) strength = 0.0F;
  if (diag) {
    diag->resolved_strength = std::max(0.0F, strength);
    diag->ambiguity_ire = amb;
    diag->sigma_ire = sigma;
    diag->temporal_eps = temporal_eps;
    diag->nr_eps = nr_eps;
  }
  const bool has_temporal = strength > 0.0F && p.cg_iterations > 0;
  const int n_passes = has_temporal ? std::max(1, c.passes) : 1;
  const bool fast = p.fast != 0;

  // ---- the sweep -------------------------------------------------------------
  std::vector<char> cached((size_t)nf * nco, 0);  // motion cache validity
  std::vector<char> cached_fast((size_t)nf * nco, 0);
  auto oidx = [&](int o) { return (int)(std::find(coffs.begin(), coffs.end(), o) - coffs.begin()); };

  for (int pss = 0; pss < n_passes; ++pss) {
    std::vector<char>& cache = fast ? cached_fast : cached;
    if (!fast) std::fill(cache.begin(), cache.end(), 0);
    // the groups solved as one batch (sequence.cpp:1046-1098)
    std::vector<std::vector<int>> groups;
    if (!has_temporal) {
      groups.push_back({});
      for (int j = 0; j < nf; ++j) groups.back().push_back(j);
    } else if (fast && c.parallel_across_fields) {
      int max_off = 0;
      for (int o : offs) max_off = std::max(max_off, std::abs(o));
      const int reach = std::max(max_off, c.nr_anchor > 0.0 ? c.nr_radius : 0);
      const int stride = reach + 1;
      for (int color = 0; color < stride; ++color) {
        groups.push_back({});
        for (int j = color; j < nf; j += stride) groups.back().push_back(j);
        if (groups.back().empty()) groups.pop_back();
      }
    } else {
      for (int j = 0; j < nf; ++j) groups.push_back({j});
    }
    const bool anchored = pss >= 1 && c.nr_anchor > 0.0 && nf > 1;
    for (const std::vector<int>& G : groups) {
      const int ng = (int)G.size();
      std::vector<std::vector<int>> present(ng);  // offset indices with motion, per field
      if (has_temporal) {
        // (a) full searches: uncached offsets in the searched set
        std::vector<const float*> refs, curs;
        std::vector<float*> ody, odx, ocf;
        std::vector<std::pair<int, int>> filled;
        for (int gi = 0; gi < ng; ++gi) {
          const int j = G[gi];
          for (int oi = 0; oi < no; ++oi) {
            const int o = offs[oi], k = j + o;
            if (k < 0 || k >= nf) continue;
            const bool in_full = !fast || o == -1 || o == 1 || o == 2;
            if (cache[(size_t)j * nco + oi]) {
              present[gi].push_back(oi);
            } else if (!c.trajectory_fit || in_full) {
              const MF m = mcache(j, oi);
              refs.push_back(F[k].y);
              curs.push_back(F[j].y);
              ody.push_back(m.dy);
              odx.push_back(m.dx);
              ocf.push_back(m.conf);
              filled.push_back({j, oi});
              present[gi].push_back(oi);
            }
          }
        }
        if (!refs.empty() &&
            !X.ok(hvd_gpu_estimate_motion(ctx, refs.data(), curs.data(), (int)refs.size(), h, w, tile, c.mc_search,
                                          ody.data(), odx.data(), ocf.data(), st)))
          return X.rc;
        for (auto& f : filled) cache[(size_t)f.first * nco + f.second] = 1;
        // (b) fast mode: predict the remaining offsets from the measured
        // ones and audit them with VerifyMotion (sequence.cpp:696-748)
        if (fast && c.trajectory_fit) {
          std::vector<const float*> vr, vc, vpy, vpx;
          std::vector<float*> vdy, vdx, vcf;
          std::vector<std::pair<int, int>> vfilled;
          for (int gi = 0; gi < ng; ++gi) {
            const int j = G[gi];
            if (present[gi].empty()) continue;
            std::sort(present[gi].begin(), present[gi].end(), [&](int a, int b) { return offs[a] < offs[b]; });
            const int m = (int)present[gi].size();
            if (m > kMaxNeighbors) return fail(ctx, HVD_GPU_ERR_INVALID, "too many measured offsets");
            TrajTab tt{};
            for (int q = 0; q < m; ++q) {
              const int oi = present[gi][q];
              const MF mf = mcache(j, oi);
              tt.dy[q] = mf.dy;
              tt.dx[q] = mf.dx;
              tt.ho[q] = ((float)par[j + offs[oi]] - (float)par[j]) / 2.0F;
              tt.off[q] = offs[oi];
            }
            float* vy = pred + (size_t)j * 4 * nt;
            float* vx = vy + nt;
            hipLaunchKernelGGL(k_traj_predict, dim3((nt + 63) / 64), dim3(64), 0, st, m, nt, tt, vy, vx);
            if (!X.launched()) return X.rc;
            for (int oi = 0; oi < no; ++oi) {
              const int o = offs[oi], k = j + o;
              if (k < 0 || k >= nf) continue;
              if (std::find(present[gi].begin(), present[gi].end(), oi) != present[gi].end()) continue;
              if (cache[(size_t)j * nco + oi]) {
                present[gi].push_back(oi);
                continue;
              }
              const float hoo = ((float)par[k] - (float)par[j]) / 2.0F;
              // one prediction buffer per pending offset, in T(10 .. 11)
              if ((vfilled.size() + 1) * 2 * nt > 2 * np)
                return fail(ctx, HVD_GPU_ERR_INVALID, "too many predicted offsets for the window's scratch");
              float* pbuf = T(10) + (size_t)vfilled.size() * 2 * nt;
              hipLaunchKernelGGL(k_traj_extrapolate, dim3((nt + 63) / 64), dim3(64), 0, st, nt, (const float*)vy,
                                 (const float*)vx, (float)o, hoo, pbuf, pbuf + nt);
              if (!X.launched()) return X.rc;
              const MF mf = mcache(j, oi);
              vr.push_back(F[k].y);
              vc.push_back(F[j].y);
              vpy.push_back(pbuf);
              vpx.push_back(pbuf + nt);
              vdy.push_back(mf.dy);
              vdx.push_back(mf.dx);
              vcf.push_back(mf.conf);
              vfilled.push_back({j, oi});
              present[gi].push_back(oi);
            }
          }
          if (!vr.empty()) {
            if (!X.ok(hvd_gpu_verify_motion(ctx, vr.data(), vc.data(), (int)vr.size(), h, w, vpy.data(), vpx.data(),
                                            tile, vdy.data(), vdx.data(), vcf.data(), st)))
              return X.rc;
            for (auto& f : vfilled) cache[(size_t)f.first * nco + f.second] = 1;
          }
        }
        // (c) the snapped working copies (sequence.cpp:749-777): every copy
        // of the colour in one launch per kCopyMax, then each field's snap
        CopyTab cp{};
        int ncp = 0;
        auto flush_copies = [&]() {
          if (ncp == 0) return true;
          hipLaunchKernelGGL(k_copy_tiles, dim3((2 * nt + 255) / 256, ncp), dim3(256), 0, st, 2 * nt, cp);
          ncp = 0;
          return X.launched();
        };
        for (int gi = 0; gi < ng; ++gi) {
          const int j = G[gi];
          std::sort(present[gi].begin(), present[gi].end(), [&](int a, int b) { return offs[a] < offs[b]; });
          for (int oi : present[gi]) {
            cp.src[ncp] = mcache(j, oi).dy;
            cp.dst[ncp] = pmf(j, oi).dy;
            if (++ncp == kCopyMax && !flush_copies()) return X.rc;
          }
        }
        if (!flush_copies()) return X.rc;
        for (int gi = 0; gi < ng; ++gi) {
          const int j = G[gi];
          const int m = (int)present[gi].size();
          if (c.trajectory_fit && m >= 2) {
            if (m > kMaxNeighbors) return fail(ctx, HVD_GPU_ERR_INVALID, "too many measured offsets");
            TrajTab tt{};
            for (int q = 0; q < m; ++q) {
              const int oi = present[gi][q];
              tt.dy[q] = pmf(j, oi).dy;
              tt.dx[q] = pmf(j, oi).dx;
              tt.conf[q] = mcache(j, oi).conf;
              tt.ho[q] = ((float)par[j + offs[oi]] - (float)par[j]) / 2.0F;
              tt.off[q] = offs[oi];
            }
            hipLaunchKernelGGL(k_traj_snap, dim3((nt + 63) / 64), dim3(64), 0, st, m, nt, tt);
            if (!X.launched()) return X.rc;
          }
        }
      }
      // (d) the odd-offset envelopes (sequence.cpp:790-940)
      std::vector<char> has_env(ng, 0);
      if (has_temporal) {
        std::vector<const float*> vin, rin;
        std::vector<float*> vout, lio, smp, rout;
        for (int gi = 0; gi < ng; ++gi) {
          const int j = G[gi];
          bool any_odd = false;
          for (int o : offs)
            if (j + o >= 0 && j + o < nf && (o & 1)) any_odd = true;
          if (!any_odd) continue;
          has_env[gi] = 1;
          float* sm = envt(j);  // [smoothed gmag | smoothed lapmag]
          float* raw = envr(j);  // [gmag | lapmag]
          hipLaunchKernelGGL(k_env_terms, dim3(grid_for(np)), dim3(256), 0, st, h, w, (const float*)F[j].y,
                             cfl(F[j].chi), raw, raw + np);
          if (!X.launched()) return X.rc;
          rin.push_back(raw);
          rin.push_back(raw + np);
          rout.push_back(sm);
          rout.push_back(sm + np);
          vin.push_back(sm);
          vout.push_back(envp(j));
          lio.push_back(sm + np);
          smp.push_back(envs(j));
        }
        if (!vin.empty()) {
          const int m = (int)vin.size();
          if (m > kEnvMax) return fail(ctx, HVD_GPU_ERR_INVALID, "more than 64 fields in one colour");
          // the smoothing of every plane of the colour in one launch
          EnvRows R{};
          for (int q = 0; q < 2 * m; ++q) {
            R.src[q] = rin[q];
            R.out[q] = rout[q];
          }
          hipLaunchKernelGGL(k_smooth_rows, dim3((h + kSmR - 1) / kSmR, 2 * m), dim3(64), 0, st, h, w, R);
          if (!X.launched()) return X.rc;
          EnvFinish E{};
          for (int q = 0; q < m; ++q) {
            E.vg_in[q] = vin[q];
            E.vgrad[q] = vout[q];
            E.lap[q] = lio[q];
            E.sample[q] = smp[q];
          }
          E.floor2 = envfl;
          hipLaunchKernelGGL(k_env_finish, dim3(m), dim3(1024), 0, st, (int)np, E);
          hipLaunchKernelGGL(k_env_apply, dim3(std::min<size_t>((np + 255) / 256, 1024), m), dim3(256), 0, st,
                             (int)np, E);
          if (!X.launched()) return X.rc;
        }
      }
      // (e) neighbour equations (sequence.cpp:941-1021), the colour's jobs
      // batched kJobChunk at a time
      std::vector<std::vector<hvd_gpu_neighbor>> nbrs(ng);
      if (has_temporal) {
        const bool gate_on = c.coherence_gate > 0.0 && !(fast && pss == 0);
        std::vector<NbJob> jobs;
        for (int gi = 0; gi < ng; ++gi) {
          const int j = G[gi];
          for (int oi = 0; oi < no; ++oi) {
            const int o = offs[oi], k = j + o;
            if (k < 0 || k >= nf) continue;
            const MF mo = pmf(j, oi);
            float* nb = nbp(j, oi);
            NbJob q{};
            q.mdy = mo.dy;
            q.mdx = mo.dx;
            q.mconf = mcache(j, oi).conf;
            q.s_k = F[k].s;
            q.car_k = cfl(F[k].carrier);
            q.chi_k = cfl(F[k].chi);
            q.chi_j = cfl(F[j].chi);
            q.comp = nb;
            q.car = (float2*)(nb + np);
            q.conf = nb + 3 * np;
            if ((o & 1) && has_env[gi]) {
              q.vgrad = envp(j);
              q.lapsm = envt(j) + np;
            }
            jobs.push_back(q);
            nbrs[gi].push_back(hvd_gpu_neighbor{nb, nb + np, nb + 3 * np});
          }
        }
        const float e2 = temporal_eps * temporal_eps;
        const float fmax = std::clamp((float)c.odd_gate_floor, 0.0F, 1.0F);
        NbJob* djobs = (NbJob*)(base + o_jobs);
        MotionDesc* dmd = (MotionDesc*)(base + o_jmd);
        float* scr = (float*)(base + o_jscr);
        for (size_t q0 = 0; q0 < jobs.size(); q0 += kJobChunk) {
          const int m = (int)std::min<size_t>(kJobChunk, jobs.size() - q0);
          std::vector<NbJob> chunk(jobs.begin() + q0, jobs.begin() + q0 + m);
          // scratch: per job 4 tile maps; the blurs read and write planes in
          // job order (coherence: 4 per job; slow confidence: 1 per job)
          float* cn_all = scr;                                   // m + 1 planes
          float* cb_all = cn_all + (size_t)(m + 1) * np;         // m + 1 planes
          float* coh_all = cb_all + (size_t)(m + 1) * np;        // 4 m planes
          float* cb4_all = coh_all + (size_t)4 * m * np;         // 4 m planes
          float* tmp_all = cb4_all + (size_t)4 * m * np;         // 4 m planes (blur scratch)
          float* tiles = tmp_all + (size_t)4 * m * np;           // 4 m tile maps
          for (int q = 0; q < m; ++q) {
            NbJob& jb = chunk[q];
            jb.sq = tiles + (size_t)4 * q * nt;
            jb.sy = jb.sq + nt;
            jb.sx = jb.sy + nt;
            jb.sc = jb.sx + nt;
            jb.cn = cn_all + (size_t)q * np;
            jb.cb = cb_all + (size_t)q * np;
            jb.coh = coh_all + (size_t)4 * q * np;
            jb.cb4 = cb4_all + (size_t)4 * q * np;
          }
          std::vector<MotionDesc> md;
          if (gate_on)
            for (int q = 0; q < m; ++q) {
              md.push_back(MotionDesc{chunk[q].coh, chunk[q].coh + np});
              md.push_back(MotionDesc{chunk[q].coh + 2 * np, chunk[q].coh + 3 * np});
            }
          const size_t md_coh = md.size();
          if (!fast)
            for (int q = 0; q < m; q += 2)
              md.push_back(MotionDesc{chunk[q].cn, q + 1 < m ? chunk[q + 1].cn : chunk[q].cn});
          if (hipMemcpyAsync(djobs, chunk.data(), sizeof(NbJob) * m, hipMemcpyHostToDevice, st) != hipSuccess ||
              (!md.empty() &&
               hipMemcpyAsync(dmd, md.data(), sizeof(MotionDesc) * md.size(), hipMemcpyHostToDevice, st) != hipSuccess))
            return fail(ctx, HVD_GPU_ERR_HIP, "job upload failed");
          const dim3 gt((nt + 63) / 64, m), gp((unsigned)((np + 255) / 256), m);
          if (fast) hipLaunchKernelGGL(k_nb_square, gt, dim3(64), 0, st, nt, (const NbJob*)djobs);
          hipLaunchKernelGGL(k_nb_snap, gt, dim3(64), 0, st, th, tw, fast ? 1 : 0, (const NbJob*)djobs);
          hipLaunchKernelGGL(k_nb_warp, gp, dim3(256), 0, st, h, w, th, tw, tile, fast ? 1 : 0, gate_on ? 1 : 0,
                             (const NbJob*)djobs);
          if (!X.launched()) return X.rc;
          // the confidence's radius-8 blur (slow mode: temporal.cpp:217-233)
          // and the coherence terms' radius-6 blur (:169-196)
          if (!fast) launch_me_blur(st, dmd + md_coh, (m + 1) / 2, h, w, 8, tmp_all, cb_all);
          if (gate_on) launch_me_blur(st, dmd, 2 * m, h, w, 6, tmp_all, cb4_all);
          hipLaunchKernelGGL(k_nb_gates, gp, dim3(256), 0, st, h, w, fast ? 1 : 0, gate_on ? 1 : 0,
                             (float)c.coherence_gate, e2, fmax, (const NbJob*)djobs);
          if (!X.launched()) return X.rc;
          // the job tables and scratch are reused by the next chunk
          if (hipStreamSynchronize(st) != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "sync failed");
        }
      }
      // (f) the PSI init (sequence.cpp:1023-1025)
      if (pss == 0 && c.psi_init)
        for (int gi = 0; gi < ng; ++gi) {
          const int j = G[gi];
          const int m = (int)nbrs[gi].size();
          if (m == 0) continue;
          std::vector<const void*> sw(m), cw(m), cf(m);
          for (int q = 0; q < m; ++q) {
            sw[q] = nbrs[gi][q].composite;
            cw[q] = nbrs[gi][q].carrier;
            cf[q] = nbrs[gi][q].confidence;
          }
          void** pt = ptrs + (size_t)(j + 2) * 64;
          if (hipMemcpyAsync(pt, sw.data(), sizeof(void*) * m, hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemcpyAsync(pt + 16, cw.data(), sizeof(void*) * m, hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemcpyAsync(pt + 32, cf.data(), sizeof(void*) * m, hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(ctx, HVD_GPU_ERR_HIP, "table upload failed");
          hipLaunchKernelGGL(k_psi, dim3(grid_for(np)), dim3(256), 0, st, np, m, F[j].s, cfl(F[j].carrier),
                             (const float* const*)pt, (const float2* const*)(pt + 16), (const float* const*)(pt + 32),
                             (float2*)F[j].chi);
          if (!X.launched()) return X.rc;
          if (hipStreamSynchronize(st) != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "sync failed");
        }
      // (g) SynthReference anchors (sequence.cpp:52-150, :1026-1034)
      std::vector<hvd_gpu_anchor> anchors;
      if (anchored) {
        anchors.resize(ng);
        for (int gi = 0; gi < ng; ++gi) {
          const int j = G[gi];
          float *aY = T(0), *aCr = T(1), *aCi = T(2), *aW = T(3), *cup = T(4), *vdy = T(5), *vdx = T(6), *Yw = T(7),
                *re = T(8), *im = T(9), *wre = T(10), *wim = T(11);
          hipLaunchKernelGGL(k_synth_init, dim3(grid_for(np)), dim3(256), 0, st, np, (const float*)F[j].y,
                             cfl(F[j].chi), aY, aCr, aCi, aW);
          if (!X.launched()) return X.rc;
          for (int k = j - c.nr_radius; k <= j + c.nr_radius; ++k) {
            if (k == j || k < 0 || k >= nf) continue;
            MF mo{smf, smf + nt, smf + 2 * nt};
            const int oi = oidx(k - j);
            if (fast && cached_fast[(size_t)j * nco + oi]) {
              mo = mcache(j, oi);
            } else {
              MF dst = fast ? mcache(j, oi) : mo;
              const float* r1[1] = {F[k].y};
              const float* c1[1] = {F[j].y};
              float* d1[1] = {dst.dy};
              float* x1[1] = {dst.dx};
              float* f1[1] = {dst.conf};
              if (!X.ok(hvd_gpu_estimate_motion(ctx, r1, c1, 1, h, w, tile, c.mc_search, d1, x1, f1, st)))
                return X.rc;
              if (fast) cached_fast[(size_t)j * nco + oi] = 1;
              mo = dst;
            }
            const float row_off = ((float)par[j] - (float)par[k]) / 2.0F;
            if (!X.ok(hvd_gpu_upsample_confidence(ctx, mo.conf, th, tw, tile, h, w, 0, cup, st)) ||
                !X.ok(hvd_gpu_vectors_per_pixel(ctx, mo.dy, mo.dx, th, tw, tile, h, w, vdy, vdx, st)) ||
                !X.ok(hvd_gpu_warp_bilinear(ctx, F[k].y, h, w, vdy, vdx, row_off, h, w, Yw, st)))
              return X.rc;
            hipLaunchKernelGGL(k_split_cplx, dim3(grid_for(np)), dim3(256), 0, st, np, cfl(F[k].chi), re, im);
            if (!X.launched()) return X.rc;
            if (!X.ok(hvd_gpu_warp_bilinear(ctx, re, h, w, vdy, vdx, row_off, h, w, wre, st)) ||
                !X.ok(hvd_gpu_warp_bilinear(ctx, im, h, w, vdy, vdx, row_off, h, w, wim, st)))
              return X.rc;
            hipLaunchKernelGGL(k_synth_accum, dim3(grid_for(np)), dim3(256), 0, st, np, (const float*)F[j].y,
                               cfl(F[j].chi), (const float*)cup, (const float*)Yw, (const float*)wre,
                               (const float*)wim, nr_eps, aY, aCr, aCi, aW);
            if (!X.launched()) return X.rc;
          }
          float* an = ancp(j);
          float *al = an, *ac = an + np, *aw = an + 3 * np;
          float* resid = T(4);
          float* rb = T(5);
          hipLaunchKernelGGL(k_synth_finish, dim3(grid_for(np)), dim3(256), 0, st, np, F[j].s, cfl(F[j].carrier),
                             (const float*)aY, (const float*)aCr, (const float*)aCi, (const float*)aW, al,
                             (float2*)ac, resid);
          if (!X.launched()) return X.rc;
          if (!X.ok(hvd_gpu_box_blur(ctx, resid, h, w, 1, rb, st))) return X.rc;
          hipLaunchKernelGGL(k_synth_weight, dim3(grid_for(np)), dim3(256), 0, st, np, (const float*)rb, nr_eps, aw);
          if (!X.launched()) return X.rc;
          anchors[gi] = hvd_gpu_anchor{al, ac, aw};
        }
      }
      // (h) the solves of the group (sequence.cpp:1026-1040)
      std::vector<hvd_gpu_plane> planes(ng);
      std::vector<hvd_gpu_neighbor> flat;
      std::vector<int> nn(ng);
      for (int gi = 0; gi < ng; ++gi) {
        planes[gi] = F[G[gi]];
        planes[gi].theta = nullptr;
        planes[gi].vswitch = nullptr;
        nn[gi] = (int)nbrs[gi].size();
        flat.insert(flat.end(), nbrs[gi].begin(), nbrs[gi].end());
      }
      if (anchored) {
        if (!X.ok(hvd_gpu_variational_refine_joint(ctx, planes.data(), ng, flat.data(), nn.data(), anchors.data(), &p,
                                                   strength, temporal_eps, (float)c.nr_anchor, nullptr, st)))
          return X.rc;
      } else if (has_temporal) {
        if (!X.ok(hvd_gpu_variational_refine_temporal_batch(ctx, planes.data(), ng, flat.data(), nn.data(), &p,
                                                            strength, temporal_eps, nullptr, st)))
          return X.rc;
      } else {
        if (!X.ok(hvd_gpu_variational_refine(ctx, planes.data(), ng, &p, nullptr, nullptr, st))) return X.rc;
      }
    }
  }
  // ---- output luma (sequence.cpp:1100-1114) ----------------------------------
  if (c.output_fidelity)
    for (int j = 0; j < nf; ++j) {
      hipLaunchKernelGGL(k_fidelity, dim3(grid_for(np)), dim3(256), 0, st, np, F[j].s, cfl(F[j].chi),
                         cfl(F[j].carrier), F[j].y);
      if (!X.launched()) return X.rc;
    }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "window sync failed");
  return X.rc;
}

}  // namespace hvd_window
