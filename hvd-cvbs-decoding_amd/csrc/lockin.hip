// lockin.hip -- per-line colour-burst lock-in, phase-trajectory smoothing,
// burst amplitude (ACC), IRE unpack and active-area preparation.
//
// Restates hvd-core engine/lockin.cpp (BurstLockin :35-62, TridiagSmooth
// :65-99, BurstLockinPhase :101-186, BurstAmplitudeIre :188-255,
// BurstLockinPhasePal :257-349) with FP64 accumulation throughout.  One
// 256-thread block per field: the per-line demodulation runs one wavefront
// per line with FP64 wave reductions; the per-field trajectory work is
// parallel over lines except the two Thomas solves, which are inherently
// serial (one lane, FP64).
#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// e^{-i w0 x}: exact quarter turns on the 4fsc grid, sincos otherwise.
__device__ inline double2 lo_at(const LockinGeom& g, int x) {
  if (g.std_phase) {
    switch (x & 3) {
      case 0: return make_double2(1.0, 0.0);
      case 1: return make_double2(0.0, -1.0);
      case 2: return make_double2(-1.0, 0.0);
      default: return make_double2(0.0, 1.0);
    }
  }
  double sn, cs;
  sincos(-g.w0 * (double)x, &sn, &cs);
  return make_double2(cs, sn);
}

__device__ inline double wrap_pi(double a) { return atan2(sin(a), cos(a)); }

// k-th smallest of v[0..n) (n <= kMaxLines) by parallel rank counting; the
// element whose (value, index) rank is k is unique.  All threads return it.
__device__ double select_kth(const double* v, int n, int k, double* out) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double vi = v[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double vj = v[j];
      rank += (vj < vi) || (vj == vi && j < i);
    }
    if (rank == k) *out = vi;
  }
  __syncthreads();
  const double r = *out;
  __syncthreads();
  return r;
}

// Thomas solve of (diag(a) + lam L) x = a d with Neumann ends (lockin.cpp:65-99),
// FP64, one lane.  The recurrences carry their previous values in registers
// (not re-read from LDS), so each row waits only on its own arithmetic.
__device__ void tridiag_smooth(const double* __restrict__ d, const double* __restrict__ a, int n, double lam,
                               double* __restrict__ x, double* __restrict__ c, double* __restrict__ dd) {
  if (n == 0) return;
  if (n == 1) { x[0] = a[0] > 0.0 ? d[0] : 0.0; return; }
  const double off = -lam;
  auto diag = [&](int i) { return a[i] + 2.0 * lam - ((i == 0 || i == n - 1) ? lam : 0.0); };
  double cp = off / diag(0);
  double dp = a[0] * d[0] / diag(0);
  c[0] = cp;
  dd[0] = dp;
  for (int i = 1; i < n; ++i) {
    const double m = diag(i) - off * cp;
    dp = (a[i] * d[i] - off * dp) / m;
    dd[i] = dp;
    if (i < n - 1) {
      cp = off / m;
      c[i] = cp;
    }
  }
  double xn = dp;
  x[n - 1] = xn;
  for (int i = n - 2; i >= 0; --i) {
    xn = dd[i] - c[i] * xn;
    x[i] = xn;
  }
}

}  // namespace

// (1) of the lock-in for every line of the batch at once, one wavefront per
// line: z[y] = mean_x[(s - mean s) e^{-i w0 x}] (lockin.cpp:35-62) and |z|,
// into z (zr, zi, amp per line; lines numbered by line_prefix).
__global__ void __launch_bounds__(256) k_lockin_demod(const LockinGeom* G, const float* const* fields,
                                                      const int* line_prefix, int nf, double* z) {
  const int lane = threadIdx.x & 63;
  const int line = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (line >= line_prefix[nf]) return;  // uniform per wave
  const int fi = field_of_tile(line_prefix, nf, line);
  const LockinGeom& g = G[fi];
  const int y = line - line_prefix[fi];
  const int x0 = g.x0, n = g.x1 - g.x0 > 0 ? g.x1 - g.x0 : 0;
  const float* row = fields[fi] + (size_t)y * g.fw;
  double m = 0.0;
  for (int k = lane; k < n; k += 64) m += (double)row[x0 + k];
  m = n ? wave_sum(m) / (double)n : 0.0;
  double ar = 0.0, ai = 0.0;
  for (int k = lane; k < n; k += 64) {
    const double v = (double)row[x0 + k] - m;
    const double2 lo = lo_at(g, x0 + k);
    ar += v * lo.x;
    ai += v * lo.y;
  }
  ar = wave_sum(ar);
  ai = wave_sum(ai);
  if (lane == 0) {
    const double zr = n ? ar / (double)n : 0.0, zi = n ? ai / (double)n : 0.0;
    z[3 * (size_t)line] = zr;
    z[3 * (size_t)line + 1] = zi;
    z[3 * (size_t)line + 2] = hypot(zr, zi);
  }
}

// Burst lock-in and phase trajectory for a batch of fields, from the
// per-line demodulation of k_lockin_demod.  PAL-family fields (g.vswitch)
// also resolve the per-line V-switch parity.
__global__ void __launch_bounds__(256) k_lockin(const LockinGeom* G, const double* z, const int* line_prefix,
                                                float* const* theta, int8_t* const* vswitch) {
  __shared__ double zr[kMaxLines], zi[kMaxLines], amp[kMaxLines];
  __shared__ double th[kMaxLines], model[kMaxLines], dv[kMaxLines], av[kMaxLines], xs[kMaxLines];
  __shared__ double t1[kMaxLines], t2[kMaxLines];
  __shared__ int8_t sw[kMaxLines];
  __shared__ char good[kMaxLines];
  __shared__ double s_sel, s_score[2];
  __shared__ int s_first, s_ng, s_hyp;
  const int fi = blockIdx.x;
  const LockinGeom& g = G[fi];
  const int h = g.fh;

  // (1) z[y] and |z[y]| from k_lockin_demod
  const double* zf = z + 3 * (size_t)line_prefix[fi];
  for (int y = threadIdx.x; y < h; y += blockDim.x) {
    zr[y] = zf[3 * y];
    zi[y] = zf[3 * y + 1];
    amp[y] = zf[3 * y + 2];
  }
  __syncthreads();
  // (2) good lines: amp > 0.2 max (lockin.cpp:121-130)
  if (threadIdx.x == 0) {
    double amax = 0.0;
    for (int y = 0; y < h; ++y) amax = fmax(amax, amp[y]);
    const double thr = amax > 0.0 ? amax * 0.2 : 1.0;
    int first = -1, ng = 0;
    for (int y = 0; y < h; ++y) {
      good[y] = amp[y] > thr;
      if (good[y]) {
        if (first < 0) first = y;
        t1[ng++] = amp[y];
      }
    }
    s_first = first;
    s_ng = ng;
  }
  __syncthreads();
  const int first = s_first;
  const double adv = g.line_advance;
  if (first < 0) {
    // no burst anywhere: raw angles (NTSC, :129) or model + alternating parity (PAL, :272-277)
    for (int y = threadIdx.x; y < h; y += blockDim.x) {
      if (g.vswitch) {
        theta[fi][y] = (float)(adv * (double)y);
        vswitch[fi][y] = (y % 2 == 0) ? 1 : -1;
      } else {
        theta[fi][y] = (float)(atan2(zi[y], zr[y]) - 0.5 * kPi);
      }
    }
    return;
  }
  // (3) measured phase, model, deviation
  if (!g.vswitch) {
    for (int y = threadIdx.x; y < h; y += blockDim.x) th[y] = atan2(zi[y], zr[y]) - 0.5 * kPi;
    __syncthreads();
    for (int y = threadIdx.x; y < h; y += blockDim.x) {
      model[y] = th[first] + adv * (double)(y - first);
      dv[y] = wrap_pi(th[y] - model[y]);
    }
  } else {
    // PAL: two parity hypotheses scored by sum cos(dev) amp (lockin.cpp:283-306)
    for (int y = threadIdx.x; y < h; y += blockDim.x) th[y] = atan2(zi[y], zr[y]);
    __syncthreads();
    const double qpi = 0.25 * kPi;
    for (int hyp = 0; hyp < 2; ++hyp) {
      const double sref = hyp == 0 ? 1.0 : -1.0;
      const double tref = th[first] - 0.5 * kPi + sref * qpi;
      double sc = 0.0;
      for (int y = threadIdx.x; y < h; y += blockDim.x) {
        if (!good[y]) continue;
        const double m = tref + adv * (double)(y - first);
        const double sy = (((y - first) % 2) == 0) ? sref : -sref;
        sc += cos(wrap_pi(th[y] - (m + 0.5 * kPi - sy * qpi))) * amp[y];
      }
      // deterministic block sum via LDS scratch
      t2[threadIdx.x] = sc;
      __syncthreads();
      if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int i = 0; i < (int)blockDim.x; ++i) tot += t2[i];
        s_score[hyp] = tot;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) s_hyp = s_score[0] >= s_score[1] ? 0 : 1;  // strict '>' keeps hypothesis 0 on a tie
    __syncthreads();
    const double sref = s_hyp == 0 ? 1.0 : -1.0;
    const double tref = th[first] - 0.5 * kPi + sref * qpi;
    for (int y = threadIdx.x; y < h; y += blockDim.x) {
      model[y] = tref + adv * (double)(y - first);
      const double sy = (((y - first) % 2) == 0) ? sref : -sref;
      sw[y] = sy > 0 ? 1 : -1;
      dv[y] = wrap_pi(th[y] - 0.5 * kPi + sy * qpi - model[y]);
    }
  }
  __syncthreads();
  // (4) amplitude weights: clip(amp / median_good, 0, 2) (lockin.cpp:147-157)
  const double med = select_kth(t1, s_ng, s_ng / 2, &s_sel) + 1e-9;
  for (int y = threadIdx.x; y < h; y += blockDim.x) av[y] = good[y] ? fmin(fmax(amp[y] / med, 0.0), 2.0) : 0.0;
  __syncthreads();
  // (5) smooth, one Huber re-weight, smooth again (lockin.cpp:159-172)
  if (threadIdx.x == 0) tridiag_smooth(dv, av, h, 25.0, xs, t1, t2);
  __syncthreads();
  for (int y = threadIdx.x; y < h; y += blockDim.x) {
    const double r = fabs(dv[y] - xs[y]);
    av[y] = av[y] * 0.15 / fmax(r, 0.15);
  }
  __syncthreads();
  if (threadIdx.x == 0) tridiag_smooth(dv, av, h, 25.0, xs, t1, t2);
  __syncthreads();
  for (int y = threadIdx.x; y < h; y += blockDim.x) {
    theta[fi][y] = (float)(model[y] + xs[y]);
    if (g.vswitch) vswitch[fi][y] = sw[y];
  }
}

// Burst amplitude for ACC (lockin.cpp:188-255): one-cycle sliding demod,
// per-line median of |env|, then the median of the lines above 0.25 x the
// median.  One block per field; requires the window to hold <= 64 cycles
// starts (x1 - x0 - 3 <= 64).
__global__ void __launch_bounds__(256) k_burst_amp(const LockinGeom* G, const float* const* fields, float* out) {
  __shared__ double per[kMaxLines], good[kMaxLines];
  __shared__ double env[4][64];
  __shared__ double s_sel;
  __shared__ int s_ng;
  const int fi = blockIdx.x;
  const LockinGeom& g = G[fi];
  const float* field = fields[fi];
  const int h = g.fh, x0 = g.x0, x1 = g.x1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (x1 - x0 < 4) {
    if (threadIdx.x == 0) out[fi] = g.nominal_burst;
    return;
  }
  const int ne = x1 - x0 - 3;
  for (int y0 = 0; y0 < h; y0 += 4) {
    const int y = y0 + wv;
    if (y < h && lane < ne) {
      const float* row = field + (size_t)y * g.fw;
      double ar = 0.0, ai = 0.0;
      for (int k = 0; k < 4; ++k) {
        const int x = x0 + lane + k;
        const double2 lo = lo_at(g, x);
        ar += (double)row[x] * lo.x;
        ai += (double)row[x] * lo.y;
      }
      env[wv][lane] = 2.0 * hypot(ar, ai) * 0.25;
    }
    __syncthreads();
    if (y < h && lane < ne) {
      const double vi = env[wv][lane];
      int rank = 0;
      for (int j = 0; j < ne; ++j) {
        const double vj = env[wv][j];
        rank += (vj < vi) || (vj == vi && j < lane);
      }
      if (rank == ne / 2) per[y] = vi;
    }
    __syncthreads();
  }
  const double med_all = select_kth(per, h, h / 2, &s_sel);
  if (threadIdx.x == 0) {
    int ng = 0;
    for (int y = 0; y < h; ++y)
      if (per[y] > 0.25 * med_all) good[ng++] = per[y];
    s_ng = ng;
  }
  __syncthreads();
  const int ng = s_ng;
  const double r = ng ? select_kth(good, ng, ng / 2, &s_sel) : (double)g.nominal_burst;
  if (threadIdx.x == 0) out[fi] = (float)r;
}

// TridiagSmooth on its own (lockin.cpp:65-99): FP64 Thomas solve in one lane,
// rounded to float.  scratch: 5 n doubles.  One block of 64.
__global__ void k_tridiag(const float* d, const float* a, int n, float lam, double* scratch, float* out) {
  if (threadIdx.x != 0) return;
  double* dd = scratch;
  double* aa = scratch + n;
  double* x = scratch + 2 * n;
  double* c = scratch + 3 * n;
  double* e = scratch + 4 * n;
  for (int i = 0; i < n; ++i) {
    dd[i] = d[i];
    aa[i] = a[i];
  }
  tridiag_smooth(dd, aa, n, (double)lam, x, c, e);
  for (int i = 0; i < n; ++i) out[i] = (float)x[i];
}

// uint16 codes -> IRE (ntsc_geometry.h:301-304 SampleToIre).
__global__ void k_u16_to_ire(const uint16_t* raw, size_t n, float blanking, float cpi, float* out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = ((float)raw[i] - blanking) / cpi;
}

// The same for a batch of fields in one launch: grid (x blocks, fields).
__global__ void k_u16_to_ire_batch(const IreDesc* D, float blanking, float cpi) {
  const IreDesc d = D[blockIdx.y];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < d.n; i += (size_t)gridDim.x * blockDim.x)
    d.out[i] = ((float)d.raw[i] - blanking) / cpi;
}

// Active crop + factored carrier tables for a batch (sequence.cpp:229-268):
// s = field[fal.., a0..]; rowph[y] = e^{i theta[fal + y]} with the chroma
// phase trim signed by the V-switch (ntsc_geometry.h:265-276); colph[x] =
// e^{i w0 (a0 + x)}.  Grid: (max(lines, width) / 256 + 1, n_fields).
__global__ void k_prepare(const PrepDesc* P) {
  const PrepDesc& p = P[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p.lines) {
    const int8_t s = p.vswitch_field ? p.vswitch_field[p.fal + i] : (int8_t)1;
    float tf = p.theta_field[p.fal + i];
    if (p.phase_offset != 0.0f) {
      const float off = p.phase_offset;  // already -deg * pi / 180
      tf = tf + ((p.signed_by_parity && s < 0) ? -off : off);
    }
    if (p.th_active) p.th_active[i] = tf;
    const double t = (double)tf;
    double sn, cs;
    sincos(t, &sn, &cs);
    p.rowph[i] = make_float2((float)cs, (float)sn);
    if (p.vsw_active) p.vsw_active[i] = s;
  }
  if (i < p.width) {
    const int sample = p.a0 + i;
    if (p.std_phase) {
      const float tab[4][2] = {{1.f, 0.f}, {0.f, 1.f}, {-1.f, 0.f}, {0.f, -1.f}};
      p.colph[i] = make_float2(tab[sample & 3][0], tab[sample & 3][1]);
    } else {
      double sn, cs;
      sincos(p.w0 * (double)sample, &sn, &cs);
      p.colph[i] = make_float2((float)cs, (float)sn);
    }
  }
}

// Active crop: s[y][x] = field[fal + y][a0 + x].  Grid (ceil(width/256), lines, n).
__global__ void k_crop(const PrepDesc* P) {
  const PrepDesc& p = P[blockIdx.z];
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x < p.width && y < p.lines) p.s[(size_t)y * p.width + x] = p.field[(size_t)(p.fal + y) * p.fw + p.a0 + x];
}

}  // namespace hvdk
