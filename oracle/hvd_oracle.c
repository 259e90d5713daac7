/*
 * hvd_oracle.c -- CPU restatement of hvd-core's Y/C-separation hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see hvd_oracle.h).  Each function names the
 * hvd-core routine it restates (paths relative to the hvd-core checkout).
 * Build: `make -C oracle` produces liboracle64.so (OR_REAL=double) and
 * liboracle32.so (OR_REAL=float).
 */
#include "hvd_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef OR_REAL
#define OR_REAL double
#endif
typedef OR_REAL R;

#if OR_FP32
#define OR_FN(n) or32_##n
#else
#define OR_FN(n) or64_##n
#endif

#define PI_D 3.14159265358979323846

/* Scalar math in the build's precision: the FP32 build calls the float
 * functions hvd-core's float code resolves to (cosf, atan2f, hypotf ...). */
#if OR_FP32
#define RCOS cosf
#define RSIN sinf
#define RATAN2 atan2f
#define RHYPOT hypotf
#else
#define RCOS cos
#define RSIN sin
#define RATAN2 atan2
#define RHYPOT hypot
#endif

typedef struct { R re, im; } C;
typedef struct { double re, im; } CD;

static inline C cmul(C a, C b) { C r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline C cmulconj(C a, C b) { C r = {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; return r; } /* a*conj(b) */
static inline R re_prod(C a, C b) { return a.re * b.re - a.im * b.im; } /* Re[a*b] */
static inline R cabs2(C a) { return a.re * a.re + a.im * a.im; }
static inline R cabsr(C a) { return RHYPOT(a.re, a.im); }

static void *xmalloc(size_t n) { void *p = calloc(n ? n : 1, 1); if (!p) abort(); return p; }

/* ------------------------------------------------------------------------ */
/* geometry: engine/ntsc_geometry.h:200-251                                 */
/* ------------------------------------------------------------------------ */
static int uses_vswitch(const or_geom *g) { return g->standard == 1 || g->standard == 2; }

double OR_FN(phase_per_sample)(const or_geom *g) {
  if (g->subcarrier_hz <= 0.0) return 0.5 * PI_D;
  return 2.0 * PI_D * g->subcarrier_hz / g->sample_rate;
}

double OR_FN(line_advance)(const or_geom *g) {
  if (g->subcarrier_hz <= 0.0) {
    if (g->standard == 1) return 2.0 * PI_D * 0.7516;
    if (g->standard == 2) return 0.5 * PI_D;
    return PI_D;
  }
  double cycles = g->subcarrier_hz * (double)g->field_width / g->sample_rate;
  return 2.0 * PI_D * (cycles - floor(cycles));
}

static double nominal_burst_ire(const or_geom *g) { return g->standard == 1 ? 21.43 : 20.0; }

/* ------------------------------------------------------------------------ */
/* order statistics (std::nth_element semantics: the k-th smallest value)   */
/* ------------------------------------------------------------------------ */
static int cmp_r(const void *a, const void *b) {
  R x = *(const R *)a, y = *(const R *)b;
  return (x > y) - (x < y);
}
static R kth_smallest(const R *v, int n, int k) {
  R *t = (R *)xmalloc(sizeof(R) * (size_t)n);
  memcpy(t, v, sizeof(R) * (size_t)n);
  qsort(t, (size_t)n, sizeof(R), cmp_r);
  R r = t[k];
  free(t);
  return r;
}
/* lockin.cpp:13-18 MedianOf: element n/2 (upper median), 0 on empty. */
static R median_upper(const R *v, int n) { return n ? kth_smallest(v, n, n / 2) : (R)0; }

/* motion.cpp:80-91 Median: numpy median (average of the two middle values). */
static R median_np(const R *v, int n) {
  if (!n) return 0;
  R *t = (R *)xmalloc(sizeof(R) * (size_t)n);
  memcpy(t, v, sizeof(R) * (size_t)n);
  qsort(t, (size_t)n, sizeof(R), cmp_r);
  R m = t[n / 2];
  if (n % 2 == 0) m = (R)0.5 * (m + t[n / 2 - 1]);
  free(t);
  return m;
}

static R wrap_pi(R a) { return RATAN2(RSIN(a), RCOS(a)); }

/* ------------------------------------------------------------------------ */
/* lock-in: engine/lockin.cpp                                              */
/* ------------------------------------------------------------------------ */

/* lockin.cpp:35-62 BurstLockin: z[y] = mean_x[(s - mean(s)) e^{-i w0 x}],
 * x the absolute sample index.  Accumulates in R. */
void OR_FN(burst_lockin)(const R *field, int h, int w, int x0, int x1, double w0d,
                         R *z /* 2h */, R *amp /* h */) {
  const int n = x1 - x0 > 0 ? x1 - x0 : 0;
  const R w0 = (R)w0d;
  for (int y = 0; y < h; ++y) {
    z[2 * y] = z[2 * y + 1] = 0;
    amp[y] = 0;
    if (n == 0) continue;
    const R *row = field + (size_t)y * w;
    R mean = 0;
    for (int x = x0; x < x1; ++x) mean += row[x];
    mean /= (R)n;
    R ar = 0, ai = 0;
    for (int x = x0; x < x1; ++x) {
      const R ph = -w0 * (R)x;
      const R v = row[x] - mean;
      ar += v * RCOS(ph);
      ai += v * RSIN(ph);
    }
    ar /= (R)n;
    ai /= (R)n;
    z[2 * y] = ar;
    z[2 * y + 1] = ai;
    amp[y] = RHYPOT(ar, ai);
  }
}

/* lockin.cpp:65-99 TridiagSmooth: (diag(a) + lam L) x = a d, Neumann ends,
 * Thomas algorithm. */
void OR_FN(tridiag_smooth)(const R *d, const R *a, int n, R lam, R *x) {
  if (n == 0) return;
  if (n == 1) { x[0] = a[0] > 0 ? d[0] : 0; return; }
  R *diag = (R *)xmalloc(sizeof(R) * (size_t)n);
  R *c = (R *)xmalloc(sizeof(R) * (size_t)n);
  R *dd = (R *)xmalloc(sizeof(R) * (size_t)n);
  for (int i = 0; i < n; ++i) diag[i] = a[i] + 2 * lam;
  diag[0] -= lam;
  diag[n - 1] -= lam;
  const R off = -lam;
  c[0] = off / diag[0];
  dd[0] = a[0] * d[0] / diag[0];
  for (int i = 1; i < n; ++i) {
    const R m = diag[i] - off * c[i - 1];
    if (i < n - 1) c[i] = off / m;
    dd[i] = (a[i] * d[i] - off * dd[i - 1]) / m;
  }
  x[n - 1] = dd[n - 1];
  for (int i = n - 2; i >= 0; --i) x[i] = dd[i] - c[i] * x[i + 1];
  free(diag); free(c); free(dd);
}

/* Shared tail of both lock-ins (lockin.cpp:160-185 / :326-348): amplitude
 * weights, smooth, one Huber re-weight, smooth again, theta = model + x. */
static void smooth_trajectory(const R *amp, const char *good, const R *d, const R *model, int h, R *theta) {
  R *ga = (R *)xmalloc(sizeof(R) * (size_t)h);
  int ng = 0;
  for (int y = 0; y < h; ++y) if (good[y]) ga[ng++] = amp[y];
  const R med = median_upper(ga, ng) + (R)1e-9;
  R *a = (R *)xmalloc(sizeof(R) * (size_t)h);
  R *a2 = (R *)xmalloc(sizeof(R) * (size_t)h);
  R *x = (R *)xmalloc(sizeof(R) * (size_t)h);
  for (int y = 0; y < h; ++y) {
    R v = good[y] ? amp[y] / med : 0;
    a[y] = v < 0 ? 0 : (v > 2 ? 2 : v);
  }
  const R lam = 25;
  OR_FN(tridiag_smooth)(d, a, h, lam, x);
  for (int y = 0; y < h; ++y) {
    R r = (R)fabs((double)(d[y] - x[y]));
    a2[y] = a[y] * (R)0.15 / (r > (R)0.15 ? r : (R)0.15);
  }
  OR_FN(tridiag_smooth)(d, a2, h, lam, x);
  for (int y = 0; y < h; ++y) theta[y] = model[y] + x[y];
  free(ga); free(a); free(a2); free(x);
}

/* lockin.cpp:101-186 BurstLockinPhase (NTSC). */
void OR_FN(burst_lockin_phase)(const R *field, const or_geom *g, R *theta) {
  const int h = g->field_height, w = g->field_width;
  R *z = (R *)xmalloc(sizeof(R) * 2 * (size_t)h);
  R *amp = (R *)xmalloc(sizeof(R) * (size_t)h);
  OR_FN(burst_lockin)(field, h, w, g->colour_burst_start, g->colour_burst_end,
                      OR_FN(phase_per_sample)(g), z, amp);
  const R half_pi = (R)(0.5 * PI_D);
  for (int y = 0; y < h; ++y) theta[y] = RATAN2(z[2 * y + 1], z[2 * y]) - half_pi;
  R amax = 0;
  for (int y = 0; y < h; ++y) if (amp[y] > amax) amax = amp[y];
  const R thr = amax > 0 ? amax * (R)0.2 : 1;
  char *good = (char *)xmalloc((size_t)h);
  int first = -1;
  for (int y = 0; y < h; ++y) {
    good[y] = amp[y] > thr;
    if (good[y] && first < 0) first = y;
  }
  if (first >= 0) {
    const R adv = (R)OR_FN(line_advance)(g);
    R *model = (R *)xmalloc(sizeof(R) * (size_t)h);
    R *d = (R *)xmalloc(sizeof(R) * (size_t)h);
    for (int y = 0; y < h; ++y) model[y] = theta[first] + adv * (R)(y - first);
    for (int y = 0; y < h; ++y) d[y] = wrap_pi(theta[y] - model[y]);
    smooth_trajectory(amp, good, d, model, h, theta);
    free(model); free(d);
  }
  free(z); free(amp); free(good);
}

/* lockin.cpp:257-349 BurstLockinPhasePal: joint phase + V-switch parity. */
void OR_FN(burst_lockin_phase_pal)(const R *field, const or_geom *g, R *theta, int8_t *vswitch) {
  const int h = g->field_height, w = g->field_width;
  R *z = (R *)xmalloc(sizeof(R) * 2 * (size_t)h);
  R *amp = (R *)xmalloc(sizeof(R) * (size_t)h);
  OR_FN(burst_lockin)(field, h, w, g->colour_burst_start, g->colour_burst_end,
                      OR_FN(phase_per_sample)(g), z, amp);
  const R adv = (R)OR_FN(line_advance)(g);
  const R qpi = (R)(0.25 * PI_D), half_pi = (R)(0.5 * PI_D);
  R amax = 0;
  for (int y = 0; y < h; ++y) if (amp[y] > amax) amax = amp[y];
  const R thr = amax > 0 ? amax * (R)0.2 : 1;
  char *good = (char *)xmalloc((size_t)h);
  int first = -1;
  for (int y = 0; y < h; ++y) {
    good[y] = amp[y] > thr;
    if (good[y] && first < 0) first = y;
  }
  if (first < 0) {
    for (int y = 0; y < h; ++y) { theta[y] = adv * (R)y; vswitch[y] = (y % 2 == 0) ? 1 : -1; }
    free(z); free(amp); free(good);
    return;
  }
  R *am = (R *)xmalloc(sizeof(R) * (size_t)h);
  for (int y = 0; y < h; ++y) am[y] = RATAN2(z[2 * y + 1], z[2 * y]);
  R best = (R)-1e30;
  R *model = (R *)xmalloc(sizeof(R) * (size_t)h);
  R *m = (R *)xmalloc(sizeof(R) * (size_t)h);
  int8_t *s = (int8_t *)xmalloc((size_t)h);
  for (int hyp = 0; hyp < 2; ++hyp) {
    const R sref = hyp == 0 ? 1 : -1;
    const R tref = am[first] - half_pi + sref * qpi;
    R score = 0;
    for (int y = 0; y < h; ++y) {
      m[y] = tref + adv * (R)(y - first);
      const int same = ((y - first) % 2) == 0;
      const R sy = same ? sref : -sref;
      s[y] = sy > 0 ? 1 : -1;
      if (good[y]) {
        const R dev = wrap_pi(am[y] - (m[y] + half_pi - sy * qpi));
        score += RCOS(dev) * amp[y];
      }
    }
    if (score > best) {
      best = score;
      memcpy(model, m, sizeof(R) * (size_t)h);
      memcpy(vswitch, s, (size_t)h);
    }
  }
  R *d = (R *)xmalloc(sizeof(R) * (size_t)h);
  for (int y = 0; y < h; ++y) {
    const R sy = vswitch[y] > 0 ? 1 : -1;
    d[y] = wrap_pi(am[y] - half_pi + sy * qpi - model[y]);
  }
  smooth_trajectory(amp, good, d, model, h, theta);
  free(z); free(amp); free(good); free(am); free(model); free(m); free(s); free(d);
}

/* lockin.cpp:188-255 BurstAmplitudeIre: sliding one-cycle demod, median per
 * line, median of the lines above 0.25 x the all-line median. */
R OR_FN(burst_amplitude_ire)(const R *field, const or_geom *g) {
  const int h = g->field_height, w = g->field_width;
  const int x0 = g->colour_burst_start, x1 = g->colour_burst_end;
  const R w0 = (R)OR_FN(phase_per_sample)(g);
  if (x1 - x0 < 4) return (R)nominal_burst_ire(g);
  const int ne = x1 - x0 - 3;
  R *env = (R *)xmalloc(sizeof(R) * (size_t)ne);
  R *per = (R *)xmalloc(sizeof(R) * (size_t)h);
  for (int y = 0; y < h; ++y) {
    const R *row = field + (size_t)y * w;
    for (int s0 = x0; s0 + 4 <= x1; ++s0) {
      R ar = 0, ai = 0;
      for (int k = 0; k < 4; ++k) {
        const int x = s0 + k;
        const R ph = -w0 * (R)x;
        ar += row[x] * RCOS(ph);
        ai += row[x] * RSIN(ph);
      }
      env[s0 - x0] = 2 * RHYPOT(ar, ai) * (R)0.25;
    }
    per[y] = median_upper(env, ne);
  }
  const R med_all = median_upper(per, h);
  int ng = 0;
  for (int y = 0; y < h; ++y) if (per[y] > (R)0.25 * med_all) per[ng++] = per[y];
  R out = ng ? median_upper(per, ng) : (R)nominal_burst_ire(g);
  free(env); free(per);
  return out;
}

/* colour.cpp:24-27 AccGain: clamp(nominal / max(measured, 1), 0.5, 2). */
R OR_FN(acc_gain)(R measured, R nominal) {
  const R den = measured > 1 ? measured : 1;
  const R gn = nominal / den;
  return gn < (R)0.5 ? (R)0.5 : (gn > 2 ? 2 : gn);
}

/* ntsc_geometry.cpp:29-47 MakeCarrier and :75-93 MakeCarrierPal:
 * c = polar(1, theta[y] + w0 (a0 + x)); switched PAL lines get -conj. */
void OR_FN(make_carrier)(const R *theta, const int8_t *vswitch /* may be NULL */, int lines,
                         const or_geom *g, R *carrier) {
  const int a0 = g->active_video_start, width = g->active_video_end - g->active_video_start;
  const R w0 = (R)OR_FN(phase_per_sample)(g);
  for (int y = 0; y < lines; ++y) {
    const int sw = vswitch && vswitch[y] < 0;
    for (int x = 0; x < width; ++x) {
      /* the phase argument is formed in R, as hvd-core forms it in float */
      const R ph = theta[y] + w0 * (R)(a0 + x);
#if OR_FP32
      R cr = cosf(ph), ci = sinf(ph);
#else
      R cr = cos(ph), ci = sin(ph);
#endif
      if (sw) cr = -cr;
      carrier[2 * ((size_t)y * width + x)] = cr;
      carrier[2 * ((size_t)y * width + x) + 1] = ci;
    }
  }
}

/* ntsc_geometry.h:265-276 ApplyChromaPhase. */
void OR_FN(apply_chroma_phase)(R deg, const or_geom *g, const int8_t *vswitch, int n, R *theta) {
  if (deg == 0) return;
  const R off = -deg * (R)PI_D / 180;
  const int signed_by = uses_vswitch(g) && vswitch;
  for (int i = 0; i < n; ++i) theta[i] += (signed_by && vswitch[i] < 0) ? -off : off;
}

/* ------------------------------------------------------------------------ */
/* FFT (engine/fft2d.h:30-33 numpy convention: forward unnormalised with    */
/* sign -1, inverse sign +1 divided by h*w).  Exact-size mixed-radix        */
/* Cooley-Tukey in FP64 with a direct DFT for prime lengths.                */
/* ------------------------------------------------------------------------ */
static void fft_rec(const CD *in, int n, int istride, CD *out, const CD *root, int rn, int sign, CD *work) {
  /* root[j] = exp(-2 pi i j / rn); this level's twiddles step by rn/n. */
  if (n == 1) { out[0] = in[0]; return; }
  int p = 2;
  while (p * p <= n && n % p) ++p;
  if (n % p) p = n;
  const int m = n / p;
  const int step = rn / n;
  if (p == n) {
    for (int k = 0; k < n; ++k) {
      double ar = 0, ai = 0;
      for (int j = 0; j < n; ++j) {
        const int e = (int)(((long)j * k) % n) * step;
        const CD wv = root[e];
        const double wi = sign < 0 ? wv.im : -wv.im;
        const CD x = in[(size_t)j * istride];
        ar += x.re * wv.re - x.im * wi;
        ai += x.re * wi + x.im * wv.re;
      }
      out[k].re = ar;
      out[k].im = ai;
    }
    return;
  }
  for (int r = 0; r < p; ++r) fft_rec(in + (size_t)r * istride, m, istride * p, out + (size_t)r * m, root, rn, sign, work);
  memcpy(work, out, sizeof(CD) * (size_t)n);
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < p; ++q) {
      const int kk = k + q * m;
      double ar = 0, ai = 0;
      for (int r = 0; r < p; ++r) {
        const int e = (int)(((long)r * kk) % n) * step;
        const CD wv = root[e];
        const double wi = sign < 0 ? wv.im : -wv.im;
        const CD y = work[(size_t)r * m + k];
        ar += y.re * wv.re - y.im * wi;
        ai += y.re * wi + y.im * wv.re;
      }
      out[kk].re = ar;
      out[kk].im = ai;
    }
  }
}

static CD *make_root(int n) {
  CD *r = (CD *)xmalloc(sizeof(CD) * (size_t)n);
  for (int j = 0; j < n; ++j) {
    const double a = -2.0 * PI_D * (double)j / (double)n;
    r[j].re = cos(a);
    r[j].im = sin(a);
  }
  return r;
}

/* 2-D transform of an h x w interleaved complex FP64 plane, in place. */
static void fft2d_cd(CD *a, int h, int w, int sign) {
  const int mx = h > w ? h : w;
  CD *in = (CD *)xmalloc(sizeof(CD) * (size_t)mx);
  CD *out = (CD *)xmalloc(sizeof(CD) * (size_t)mx);
  CD *wk = (CD *)xmalloc(sizeof(CD) * (size_t)mx);
  CD *rw = make_root(w), *rh = make_root(h);
  for (int y = 0; y < h; ++y) {
    memcpy(in, a + (size_t)y * w, sizeof(CD) * (size_t)w);
    fft_rec(in, w, 1, out, rw, w, sign, wk);
    memcpy(a + (size_t)y * w, out, sizeof(CD) * (size_t)w);
  }
  for (int x = 0; x < w; ++x) {
    for (int y = 0; y < h; ++y) in[y] = a[(size_t)y * w + x];
    fft_rec(in, h, 1, out, rh, h, sign, wk);
    for (int y = 0; y < h; ++y) a[(size_t)y * w + x] = out[y];
  }
  if (sign > 0) {
    const double s = 1.0 / ((double)h * (double)w);
    for (size_t i = 0; i < (size_t)h * w; ++i) { a[i].re *= s; a[i].im *= s; }
  }
  free(in); free(out); free(wk); free(rw); free(rh);
}

/* fft2d.cpp:88-118 Forward (sign -1) / Inverse (sign +1, / h w); interleaved
 * R planes, computed in FP64. */
void OR_FN(fft2d)(R *a, int h, int w, int sign) {
  CD *t = (CD *)xmalloc(sizeof(CD) * (size_t)h * w);
  for (size_t i = 0; i < (size_t)h * w; ++i) { t[i].re = a[2 * i]; t[i].im = a[2 * i + 1]; }
  fft2d_cd(t, h, w, sign);
  for (size_t i = 0; i < (size_t)h * w; ++i) { a[2 * i] = (R)t[i].re; a[2 * i + 1] = (R)t[i].im; }
  free(t);
}

/* ------------------------------------------------------------------------ */
/* gradients: engine/gradients.h (Neumann forward differences + adjoints)   */
/* ------------------------------------------------------------------------ */
void OR_FN(dx)(const R *a, int h, int w, R *d) {
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x + 1 < w; ++x) d[(size_t)y * w + x] = a[(size_t)y * w + x + 1] - a[(size_t)y * w + x];
    if (w > 0) d[(size_t)y * w + w - 1] = 0;
  }
}
void OR_FN(dy)(const R *a, int h, int w, R *d) {
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x)
      d[(size_t)y * w + x] = y + 1 < h ? a[(size_t)(y + 1) * w + x] - a[(size_t)y * w + x] : 0;
}
void OR_FN(dxt)(const R *a, int h, int w, R *d) {
  for (int y = 0; y < h; ++y) {
    const R *ar = a + (size_t)y * w;
    R *dr = d + (size_t)y * w;
    if (w < 2) { for (int x = 0; x < w; ++x) dr[x] = 0; continue; }
    dr[0] = -ar[0];
    for (int x = 1; x < w - 1; ++x) dr[x] = ar[x - 1] - ar[x];
    dr[w - 1] = ar[w - 2];
  }
}
void OR_FN(dyt)(const R *a, int h, int w, R *d) {
  if (h < 2) { memset(d, 0, sizeof(R) * (size_t)h * w); return; }
  for (int x = 0; x < w; ++x) d[x] = -a[x];
  for (int y = 1; y < h - 1; ++y)
    for (int x = 0; x < w; ++x) d[(size_t)y * w + x] = a[(size_t)(y - 1) * w + x] - a[(size_t)y * w + x];
  for (int x = 0; x < w; ++x) d[(size_t)(h - 1) * w + x] = a[(size_t)(h - 2) * w + x];
}

/* Complex-plane versions (interleaved) of the four operators plus the
 * diagonal pair of gradients.h:195-277. */
static void cdx(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    if (x + 1 < w) { d[i].re = a[i + 1].re - a[i].re; d[i].im = a[i + 1].im - a[i].im; }
    else { d[i].re = 0; d[i].im = 0; }
  }
}
static void cdy(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    if (y + 1 < h) { d[i].re = a[i + w].re - a[i].re; d[i].im = a[i + w].im - a[i].im; }
    else { d[i].re = 0; d[i].im = 0; }
  }
}
static void cdxt(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    C v = {0, 0};
    if (w >= 2) {
      if (x == 0) { v.re = -a[i].re; v.im = -a[i].im; }
      else if (x == w - 1) v = a[i - 1];
      else { v.re = a[i - 1].re - a[i].re; v.im = a[i - 1].im - a[i].im; }
    }
    d[i] = v;
  }
}
static void cdyt(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    C v = {0, 0};
    if (h >= 2) {
      if (y == 0) { v.re = -a[i].re; v.im = -a[i].im; }
      else if (y == h - 1) v = a[i - w];
      else { v.re = a[i - w].re - a[i].re; v.im = a[i - w].im - a[i].im; }
    }
    d[i] = v;
  }
}
static void cd1(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    if (y + 1 < h && x + 1 < w) { d[i].re = a[i + w + 1].re - a[i].re; d[i].im = a[i + w + 1].im - a[i].im; }
    else { d[i].re = 0; d[i].im = 0; }
  }
}
static void cd1t(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    C v = {0, 0};
    if (y + 1 < h && x + 1 < w) { v.re -= a[i].re; v.im -= a[i].im; }
    if (y >= 1 && x >= 1) { v.re += a[i - w - 1].re; v.im += a[i - w - 1].im; }
    d[i] = v;
  }
}
static void cd2(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    if (y + 1 < h && x >= 1) { d[i].re = a[i + w - 1].re - a[i].re; d[i].im = a[i + w - 1].im - a[i].im; }
    else { d[i].re = 0; d[i].im = 0; }
  }
}
static void cd2t(const C *a, int h, int w, C *d) {
  for (int y = 0; y < h; ++y) for (int x = 0; x < w; ++x) {
    size_t i = (size_t)y * w + x;
    C v = {0, 0};
    if (y + 1 < h && x >= 1) { v.re -= a[i].re; v.im -= a[i].im; }
    if (y >= 1 && x + 1 < w) { v.re += a[i - w + 1].re; v.im += a[i - w + 1].im; }
    d[i] = v;
  }
}
void OR_FN(cplx_op)(int op, const R *a, int h, int w, R *d) {
  switch (op) {
    case 0: cdx((const C *)a, h, w, (C *)d); break;
    case 1: cdy((const C *)a, h, w, (C *)d); break;
    case 2: cdxt((const C *)a, h, w, (C *)d); break;
    case 3: cdyt((const C *)a, h, w, (C *)d); break;
    case 4: cd1((const C *)a, h, w, (C *)d); break;
    case 5: cd1t((const C *)a, h, w, (C *)d); break;
    case 6: cd2((const C *)a, h, w, (C *)d); break;
    default: cd2t((const C *)a, h, w, (C *)d); break;
  }
}

/* ------------------------------------------------------------------------ */
/* holographic init: engine/holographic_init.cpp                            */
/* ------------------------------------------------------------------------ */

/* Mutation hooks for tests/test_oracle_init_pins.py: each macro, when given
 * on the compiler line, plants one plausible mistake in the init so the test
 * can show that its closed-form checks catch it.  Never set in lib/. */
#ifndef OR_MUT_BOX_R
#define OR_MUT_BOX_R 3 /* :195 EdgeEnergy's blur radius */
#endif
#ifndef OR_MUT_WEIGHT_BIAS
#define OR_MUT_WEIGHT_BIAS 0.5 /* :259,302 weight = 1/(E + 0.5) */
#endif
#ifndef OR_MUT_PAD_CAP
#define OR_MUT_PAD_CAP 16 /* :220 pad = min(16, h-1) */
#endif

/* :126-131 MirrorRow (reflect101: the edge row is not repeated). */
static int mirror_row(int y, int h) {
  if (h <= 1) return 0;
#ifdef OR_MUT_SYMMETRIC_MIRROR
  const int period = 2 * h; /* edge row repeated */
  y = ((y % period) + period) % period;
  return y < h ? y : period - 1 - y;
#else
  const int period = 2 * h - 2;
  y = ((y % period) + period) % period;
  return y < h ? y : period - y;
#endif
}

/* :141-150 PadRowsMirror: `pad` reflected rows above and below a complex
 * plane (interleaved re, im).  out is (h + 2 pad) x w. */
void OR_FN(pad_rows_mirror)(const R *a, int h, int w, int pad, R *out) {
  for (int y = 0; y < h + 2 * pad; ++y) {
    const int sy = mirror_row(y - pad, h);
    memcpy(out + (size_t)y * w * 2, a + (size_t)sy * w * 2, sizeof(R) * 2 * (size_t)w);
  }
}

/* :16-25 FftFreq (numpy fftfreq). */
static double fftfreq(int i, int n, double d) {
  const int half = (n % 2 == 0) ? n / 2 : (n + 1) / 2;
  const int k = i < half ? i : i - n;
  return (double)k / ((double)n * d);
}

/* :58-113 BoxBlur: zero-padded 'same', constant 1/(2r+1); rows pass (along y)
 * then columns pass (along x). */
void OR_FN(box_blur_zero)(R *a, int h, int w, int r) {
  const R norm = (R)1 / (R)(2 * r + 1);
  R *col = (R *)xmalloc(sizeof(R) * (size_t)(h > w ? h : w));
  for (int x = 0; x < w; ++x) {
    for (int y = 0; y < h; ++y) col[y] = a[(size_t)y * w + x];
    for (int y = 0; y < h; ++y) {
      R acc = 0;
      for (int k = -r; k <= r; ++k) if (y + k >= 0 && y + k < h) acc += col[y + k];
      a[(size_t)y * w + x] = acc * norm;
    }
  }
  for (int y = 0; y < h; ++y) {
    memcpy(col, a + (size_t)y * w, sizeof(R) * (size_t)w);
    for (int x = 0; x < w; ++x) {
      R acc = 0;
      for (int k = -r; k <= r; ++k) if (x + k >= 0 && x + k < w) acc += col[x + k];
      a[(size_t)y * w + x] = acc * norm;
    }
  }
  free(col);
}

/* :170-196 ResidualLuma + EdgeEnergy, then the :258-262 blend weight. */
static void blend_variant(const R *s, const C *carrier, const C *chi_v, int h, int w, C *chi_num, R *w_sum) {
  const size_t n = (size_t)h * w;
  R *yv = (R *)xmalloc(sizeof(R) * n);
  R *gx = (R *)xmalloc(sizeof(R) * n);
  R *gy = (R *)xmalloc(sizeof(R) * n);
  for (size_t i = 0; i < n; ++i) yv[i] = s[i] - re_prod(chi_v[i], carrier[i]);
  OR_FN(dx)(yv, h, w, gx);
  OR_FN(dy)(yv, h, w, gy);
  for (size_t i = 0; i < n; ++i) yv[i] = (R)fabs((double)gx[i]) + (R)fabs((double)gy[i]);
  OR_FN(box_blur_zero)(yv, h, w, OR_MUT_BOX_R);
  for (size_t i = 0; i < n; ++i) {
    const R wt = (R)1 / (yv[i] + (R)OR_MUT_WEIGHT_BIAS);
    chi_num[i].re += chi_v[i].re * wt;
    chi_num[i].im += chi_v[i].im * wt;
    w_sum[i] += wt;
  }
  free(yv); free(gx); free(gy);
}

/* :224-306, :308-315 the blend: chi0 = sum_v w_v chi_v / sum_v w_v with
 * w_v = 1 / (E_v + 0.5), E_v the edge energy of S - Re[chi_v c]; then
 * Y0 = S - Re[chi0 c].  chi_v holds nvar complex planes back to back. */
void OR_FN(init_blend)(const R *s, const R *carrier_r, const R *chi_v_r, int nvar, int h, int w,
                       R *y0, R *chi0_r) {
  const C *carrier = (const C *)carrier_r;
  const C *chi_v = (const C *)chi_v_r;
  const size_t n = (size_t)h * w;
  C *chi_num = (C *)xmalloc(sizeof(C) * n);
  R *w_sum = (R *)xmalloc(sizeof(R) * n);
  for (int v = 0; v < nvar; ++v) blend_variant(s, carrier, chi_v + (size_t)v * n, h, w, chi_num, w_sum);
  C *chi0 = (C *)chi0_r;
  for (size_t i = 0; i < n; ++i) {
    chi0[i].re = chi_num[i].re / w_sum[i];
    chi0[i].im = chi_num[i].im / w_sum[i];
    y0[i] = s[i] - re_prod(chi0[i], carrier[i]);
  }
  free(chi_num); free(w_sum);
}

/* :207-252, :265-295 the crop variants: demod = S conj(c), reflect101-padded
 * by min(16, h-1) rows, one forward 2-D DFT; per variant the spectrum times
 * the separable Gaussian (the vertical cutoff against the unpadded h), the
 * inverse DFT, the rows cropped back, x2.  The optional third variant also
 * scales each bin by min(|D[k]|, |D[-k]|) / (|D[k]| + 1e-6).  chi_v receives
 * 2 (or 3) complex h x w planes back to back. */
int OR_FN(init_variants)(const R *s, const R *carrier_r, int h, int w, double sample_rate,
                         int symmetry_variant, R *chi_v_r) {
  const C *carrier = (const C *)carrier_r;
  C *chi_v = (C *)chi_v_r;
  const size_t n = (size_t)h * w;
  const int pad = h > 1 ? (OR_MUT_PAD_CAP < h - 1 ? OR_MUT_PAD_CAP : h - 1) : 0;
  const int hp = h + 2 * pad;
  const size_t np = (size_t)hp * w;
  C *demod = (C *)xmalloc(sizeof(C) * n);
  C *padded = (C *)xmalloc(sizeof(C) * np);
  for (size_t i = 0; i < n; ++i) {
    demod[i].re = s[i] * carrier[i].re;
    demod[i].im = -s[i] * carrier[i].im;
  }
  OR_FN(pad_rows_mirror)((const R *)demod, h, w, pad, (R *)padded);
  CD *spec = (CD *)xmalloc(sizeof(CD) * np);
  for (size_t i = 0; i < np; ++i) { spec[i].re = (double)padded[i].re; spec[i].im = (double)padded[i].im; }
  fft2d_cd(spec, hp, w, -1);

  CD *work = (CD *)xmalloc(sizeof(CD) * np);
  const double variants[3][2] = {{0.8, 120.0}, {1.8, 30.0}, {1.3, 60.0}};
  const int nvar = symmetry_variant ? 3 : 2;
  double *mag = NULL;
  if (symmetry_variant) {
    mag = (double *)xmalloc(sizeof(double) * np);
    for (size_t i = 0; i < np; ++i) mag[i] = hypot(spec[i].re, spec[i].im);
  }
#ifdef OR_MUT_HREF_PADDED
  const int h_ref = hp;
#else
  const int h_ref = h;
#endif
  for (int v = 0; v < nvar; ++v) {
    const double cx = variants[v][0] * 1.0e6;
    const double cy = variants[v][1] / (2.0 * (double)h_ref);
    for (int y = 0; y < hp; ++y) {
      const double ry = fftfreq(y, hp, 1.0) / cy;
      const double gy = exp(-0.5 * ry * ry);
      for (int x = 0; x < w; ++x) {
        const double rx = fftfreq(x, w, 1.0 / sample_rate) / cx;
        double k = gy * exp(-0.5 * rx * rx);
        const size_t i = (size_t)y * w + x;
        if (v == 2) {
          /* :281-289 spectral-symmetry certification; bin k pairs with -k */
#ifdef OR_MUT_SYM_NOROLL
          const int ry_ = hp - 1 - y, rx_ = w - 1 - x;
#else
          const int ry_ = (hp - y) % hp, rx_ = (w - x) % w;
#endif
          const double m = mag[i], mr = mag[(size_t)ry_ * w + rx_];
#ifdef OR_MUT_SYM_NOMIN
          k *= mr / (m + 1e-6);
#else
          k *= (m < mr ? m : mr) / (m + 1e-6);
#endif
        }
        work[i].re = spec[i].re * k;
        work[i].im = spec[i].im * k;
      }
    }
    fft2d_cd(work, hp, w, +1);
    C *out = chi_v + (size_t)v * n;
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const CD c = work[(size_t)(y + pad) * w + x];
        out[(size_t)y * w + x].re = (R)(2.0 * c.re);
        out[(size_t)y * w + x].im = (R)(2.0 * c.im);
      }
  }
  free(demod); free(padded); free(spec); free(work); free(mag);
  return nvar;
}

/* :200-317 HolographicInit = the crop variants, then their blend.
 * s (h x w), carrier (h x w complex) -> y0, chi0. */
void OR_FN(holographic_init)(const R *s, const R *carrier_r, int h, int w, double sample_rate,
                             int symmetry_variant, R *y0, R *chi0_r) {
  const size_t n = (size_t)h * w;
  R *chi_v = (R *)xmalloc(sizeof(C) * n * 3);
  const int nvar = OR_FN(init_variants)(s, carrier_r, h, w, sample_rate, symmetry_variant, chi_v);
  OR_FN(init_blend)(s, carrier_r, chi_v, nvar, h, w, y0, chi0_r);
  free(chi_v);
}

/* ------------------------------------------------------------------------ */
/* variational solve: engine/variational.cpp                                */
/* ------------------------------------------------------------------------ */

/* :130-196 ResolveChromaAniso. */
R OR_FN(resolve_chroma_aniso)(const R *chi0_r, int h, int w, int is_pal, R forced) {
  if (forced > 0) return forced;
  const C *chi0 = (const C *)chi0_r;
  const int group = is_pal ? 4 : 2;
  const int hh = h / group;
  const size_t cap = (size_t)(hh / 2 + 1) * (size_t)(w / 3 + 1);
  R *gx = (R *)xmalloc(sizeof(R) * cap);
  R *gy = (R *)xmalloc(sizeof(R) * cap);
  int n = 0;
  for (int t = 1; t < hh; t += 2) {
    for (int x = 1; x < w; x += 3) {
      C p = {0, 0}, pl = {0, 0}, pu = {0, 0};
      for (int k = 0; k < group; ++k) {
        const C a = chi0[(size_t)(group * t + k) * w + x];
        const C b = chi0[(size_t)(group * t + k) * w + x - 1];
        const C c = chi0[(size_t)(group * (t - 1) + k) * w + x];
        p.re += a.re; p.im += a.im;
        pl.re += b.re; pl.im += b.im;
        pu.re += c.re; pu.im += c.im;
      }
      const R inv = (R)1 / (R)group;
      p.re *= inv; p.im *= inv; pl.re *= inv; pl.im *= inv; pu.re *= inv; pu.im *= inv;
      C ex = {p.re - pl.re, p.im - pl.im}, ey = {p.re - pu.re, p.im - pu.im};
      gx[n] = cabsr(ex);
      gy[n] = cabsr(ey);
      ++n;
    }
  }
  R px = 0, py = 0;
  if (n) {
    const int k = n - 1 - n / 50;
    px = kth_smallest(gx, n, k);
    py = kth_smallest(gy, n, k);
  }
  free(gx); free(gy);
  const R r = py / (px > (R)1e-6 ? px : (R)1e-6);
  R a;
  if (is_pal) { a = (R)0.7 + (R)0.9 * (r - (R)1.3); if (a < (R)0.55) a = (R)0.55; if (a > 1) a = 1; }
  else { a = (R)0.5 + (R)1.1 * (r - (R)1.3); if (a < (R)0.5) a = (R)0.5; if (a > 1) a = 1; }
  return a;
}

typedef struct {
  int h, w;
  size_t n;
  const R *s;
  const C *car;
  R *wx, *wy, *wcx, *wcy, *wd1, *wd2;
  R mu_h, mu_v, mu_d, lam_y;
  /* temporal neighbours (3-D; nn = 0: the 2-D solve): composite S_w, the
   * fixed delta dc = carrier - c_w, confidence and the stage's robust
   * weight wt of each (variational.cpp:255-281, :452-461) */
  int nn;
  const R *const *sw;
  C **dc;
  const R *const *conf;
  R **wt;
  R nu, et2;
  R *rt;
  /* scratch */
  R *yc, *t1, *t2, *img;
  C *c1, *c2, *c3, *cp;
} solver;

/* :270-281 temporal residual (S_w - S) + Re[c dc] of neighbour j into sv->rt. */
static void sv_temporal_residual(solver *sv, const C *c, int j) {
  for (size_t i = 0; i < sv->n; ++i) sv->rt[i] = (sv->sw[j][i] - sv->s[i]) + re_prod(c[i], sv->dc[j][i]);
}

/* :286-347 grad: the 2-D terms and, per neighbour, 2 nu wt rt conj(dc). */
static void sv_grad(solver *sv, const C *c, C *out) {
  const int h = sv->h, w = sv->w;
  const size_t n = sv->n;
  for (size_t i = 0; i < n; ++i) sv->yc[i] = sv->s[i] - re_prod(c[i], sv->car[i]);
  OR_FN(dx)(sv->yc, h, w, sv->t1);
  for (size_t i = 0; i < n; ++i) sv->t1[i] *= sv->wx[i];
  OR_FN(dxt)(sv->t1, h, w, sv->img);
  OR_FN(dy)(sv->yc, h, w, sv->t2);
  for (size_t i = 0; i < n; ++i) sv->t2[i] *= sv->wy[i];
  OR_FN(dyt)(sv->t2, h, w, sv->t1);
  for (size_t i = 0; i < n; ++i) sv->img[i] = 2 * sv->lam_y * (sv->img[i] + sv->t1[i]);

  cdx(c, h, w, sv->c1);
  for (size_t i = 0; i < n; ++i) { sv->c1[i].re *= sv->wcx[i]; sv->c1[i].im *= sv->wcx[i]; }
  cdxt(sv->c1, h, w, sv->cp);
  cdy(c, h, w, sv->c2);
  for (size_t i = 0; i < n; ++i) { sv->c2[i].re *= sv->wcy[i]; sv->c2[i].im *= sv->wcy[i]; }
  cdyt(sv->c2, h, w, sv->c3);
  for (size_t i = 0; i < n; ++i) {
    sv->cp[i].re = 2 * (sv->mu_h * sv->cp[i].re + sv->mu_v * sv->c3[i].re);
    sv->cp[i].im = 2 * (sv->mu_h * sv->cp[i].im + sv->mu_v * sv->c3[i].im);
  }
  if (sv->mu_d > 0) {
    cd1(c, h, w, sv->c1);
    for (size_t i = 0; i < n; ++i) { sv->c1[i].re *= sv->wd1[i]; sv->c1[i].im *= sv->wd1[i]; }
    cd1t(sv->c1, h, w, sv->c2);
    cd2(c, h, w, sv->c1);
    for (size_t i = 0; i < n; ++i) { sv->c1[i].re *= sv->wd2[i]; sv->c1[i].im *= sv->wd2[i]; }
    cd2t(sv->c1, h, w, sv->c3);
    for (size_t i = 0; i < n; ++i) {
      sv->cp[i].re += 2 * sv->mu_d * (sv->c2[i].re + sv->c3[i].re);
      sv->cp[i].im += 2 * sv->mu_d * (sv->c2[i].im + sv->c3[i].im);
    }
  }
  /* out = -img * conj(carrier) + cprior */
  for (size_t i = 0; i < n; ++i) {
    out[i].re = -sv->img[i] * sv->car[i].re + sv->cp[i].re;
    out[i].im = sv->img[i] * sv->car[i].im + sv->cp[i].im;
  }
  for (int j = 0; j < sv->nn; ++j) {
    sv_temporal_residual(sv, c, j);
    for (size_t i = 0; i < n; ++i) {
      const R k = 2 * sv->nu * sv->wt[j][i] * sv->rt[i];
      out[i].re += k * sv->dc[j][i].re;
      out[i].im -= k * sv->dc[j][i].im;
    }
  }
}

/* :352-399 curv: the quadratic form of the fixed-weight objective along d. */
static double sv_curv(solver *sv, const C *d) {
  const int h = sv->h, w = sv->w;
  const size_t n = sv->n;
  for (size_t i = 0; i < n; ++i) sv->yc[i] = -re_prod(d[i], sv->car[i]);
  OR_FN(dx)(sv->yc, h, w, sv->t1);
  OR_FN(dy)(sv->yc, h, w, sv->t2);
  cdx(d, h, w, sv->c1);
  cdy(d, h, w, sv->c2);
  double acc = 0;
  for (size_t i = 0; i < n; ++i) {
    acc += (double)(sv->lam_y * sv->wx[i]) * sv->t1[i] * sv->t1[i];
    acc += (double)(sv->lam_y * sv->wy[i]) * sv->t2[i] * sv->t2[i];
    acc += (double)sv->mu_h * sv->wcx[i] * cabs2(sv->c1[i]);
    acc += (double)sv->mu_v * sv->wcy[i] * cabs2(sv->c2[i]);
  }
  if (sv->mu_d > 0) {
    cd1(d, h, w, sv->c1);
    cd2(d, h, w, sv->c2);
    double ad = 0;
    for (size_t i = 0; i < n; ++i) {
      ad += (double)sv->wd1[i] * cabs2(sv->c1[i]);
      ad += (double)sv->wd2[i] * cabs2(sv->c2[i]);
    }
    acc += sv->mu_d * ad;
  }
  for (int j = 0; j < sv->nn; ++j) {  /* :385-397 nu sum wt Re[d dc]^2 */
    double at = 0;
    for (size_t i = 0; i < n; ++i) {
      const R re = re_prod(d[i], sv->dc[j][i]);
      at += (double)sv->wt[j][i] * re * re;
    }
    acc += (double)sv->nu * at;
  }
  return acc;
}

/* :72-117 FP64 reductions. */
static double sum_abs2(const C *g, size_t n) {
  double acc = 0;
  for (size_t i = 0; i < n; ++i) acc += (double)g[i].re * g[i].re + (double)g[i].im * g[i].im;
  return acc;
}
static double dot_real(const C *a, const C *b, size_t n) {
  double acc = 0;
  for (size_t i = 0; i < n; ++i) acc += (double)a[i].re * b[i].re + (double)a[i].im * b[i].im;
  return acc;
}

/* :406-451 IRLS weights at the current chi (lagged diffusivity). */
static void sv_weights(solver *sv, const C *chi, R eps, R eps_c, R k) {
  const int h = sv->h, w = sv->w;
  const size_t n = sv->n;
  const R e2 = eps * eps, ec2 = eps_c * eps_c;
  for (size_t i = 0; i < n; ++i) sv->yc[i] = sv->s[i] - re_prod(chi[i], sv->car[i]);
  OR_FN(dx)(sv->yc, h, w, sv->t1);
  OR_FN(dy)(sv->yc, h, w, sv->t2);
  cdx(chi, h, w, sv->c1);
  cdy(chi, h, w, sv->c2);
  for (size_t i = 0; i < n; ++i) {
    const R gx = sv->t1[i], gy = sv->t2[i];
    sv->wx[i] = eps / (R)sqrt((double)(gx * gx + e2));
    sv->wy[i] = eps / (R)sqrt((double)(gy * gy + e2));
    const R mx = cabsr(sv->c1[i]), my = cabsr(sv->c2[i]);
    sv->wcx[i] = eps_c / (R)sqrt((double)(mx * mx + k * gx * gx + ec2));
    sv->wcy[i] = eps_c / (R)sqrt((double)(my * my + k * gy * gy + ec2));
  }
  if (sv->mu_d > 0) {
    cd1(chi, h, w, sv->c1);
    cd2(chi, h, w, sv->c2);
    for (size_t i = 0; i < n; ++i) {
      const R a1 = cabsr(sv->c1[i]) / eps_c, a2 = cabsr(sv->c2[i]) / eps_c;
      sv->wd1[i] = (R)1 / (R)sqrt((double)(1 + a1 * a1));
      sv->wd2[i] = (R)1 / (R)sqrt((double)(1 + a2 * a2));
    }
  }
  if (sv->nn > 0 && sv->nu > 0) {
    /* :452-461 robust temporal weights wt = conf et2 / (rt^2 + et2) */
    for (int j = 0; j < sv->nn; ++j) {
      sv_temporal_residual(sv, chi, j);
      for (size_t i = 0; i < n; ++i)
        sv->wt[j][i] = sv->conf[j][i] * sv->et2 / (sv->rt[i] * sv->rt[i] + sv->et2);
    }
    /* :463-516 prior renormalisation: the 2-D diffusivities times
     * 1 + nu sum_j wt_j |dc_j|^2 */
    for (size_t i = 0; i < n; ++i) sv->yc[i] = 1;
    for (int j = 0; j < sv->nn; ++j)
      for (size_t i = 0; i < n; ++i) sv->yc[i] += sv->nu * sv->wt[j][i] * cabs2(sv->dc[j][i]);
    for (size_t i = 0; i < n; ++i) {
      const R ki = sv->yc[i];
      sv->wx[i] *= ki;
      sv->wy[i] *= ki;
      sv->wcx[i] *= ki;
      sv->wcy[i] *= ki;
      if (sv->mu_d > 0) {
        sv->wd1[i] *= ki;
        sv->wd2[i] *= ki;
      }
    }
  }
}

/* :198-556 VariationalRefine, 2-D path (neighbors empty).  chi_r is chi0 on
 * entry and chi on return; y receives S - Re[chi c].  stage_iters (optional,
 * irls_outer entries) receives the CG iterations each stage ran; the return
 * value is the resolved aniso. */
/* As variational_refine, and optionally (stages non-NULL) saves chi at the
 * start of each IRLS stage (irls_outer planes of h*w complex): the weights
 * of stage k are a function of stage k's starting chi and S, which is all
 * the adjoint needs to recompute them.  hh_floor and gg_floor are the two
 * absolute CG exits (curv <= hh_floor, |g|^2 < gg_floor N); hvd-core's are
 * 1e-12 and 1e-10.  Tests that need each stage solved to round-off pass
 * smaller floors. */
static R refine_impl(const R *s, const R *carrier_r, int h, int w, const or_cfg *cfg, R *chi_r, R *y,
                     int *stage_iters, R *stages, double hh_floor, double gg_floor, int nn, const R *sw,
                     const R *cw, const R *conf, R nu, R eps_t) {
  solver sv;
  memset(&sv, 0, sizeof sv);
  sv.h = h; sv.w = w; sv.n = (size_t)h * w; sv.s = s; sv.car = (const C *)carrier_r;
  const size_t n = sv.n;
  /* neighbours: dc = carrier - c_w, fixed for the solve (:255-268); nu is
   * the temporal strength when neighbours are given (:224) */
  const R *swp[16];
  const R *confp[16];
  C *dcp[16];
  R *wtp[16];
  if (nn > 16) nn = 16;
  sv.nn = nn;
  sv.nu = nn > 0 ? nu : 0;
  sv.et2 = eps_t * eps_t;
  for (int j = 0; j < nn; ++j) {
    swp[j] = sw + (size_t)j * n;
    confp[j] = conf + (size_t)j * n;
    dcp[j] = (C *)xmalloc(sizeof(C) * n);
    wtp[j] = (R *)xmalloc(sizeof(R) * n); /* zero until the first weights pass (nu > 0) */
    const C *cwj = (const C *)cw + (size_t)j * n;
    for (size_t i = 0; i < n; ++i) {
      dcp[j][i].re = sv.car[i].re - cwj[i].re;
      dcp[j][i].im = sv.car[i].im - cwj[i].im;
    }
  }
  sv.sw = swp; sv.conf = confp; sv.dc = dcp; sv.wt = wtp;
  sv.rt = (R *)xmalloc(sizeof(R) * n);
  C *chi = (C *)chi_r;
  const R eps = (R)cfg->charbonnier_eps, eps_c = (R)cfg->chroma_eps, k = (R)cfg->structure_coupling;
  const R lambda_c = (R)cfg->lambda_c;
  const R aniso = OR_FN(resolve_chroma_aniso)(chi_r, h, w, cfg->is_pal, (R)cfg->chroma_aniso);
  R mu_h = lambda_c * aniso, mu_v = lambda_c, mu_d = lambda_c * (R)cfg->diag_prior * (R)0.5;
  {
    const R tot0 = lambda_c * (1 + aniso);
    const R sc = mu_d > 0 ? tot0 / (tot0 + 2 * mu_d) : 1;
    mu_h *= sc; mu_v *= sc; mu_d *= sc;
  }
  sv.mu_h = mu_h; sv.mu_v = mu_v; sv.mu_d = mu_d;
  sv.lam_y = (R)(cfg->lambda_y > 0 ? cfg->lambda_y : 1.0);
  const int irls_outer = cfg->irls_outer > 1 ? cfg->irls_outer : 1;
  const int cg_total = cfg->fast ? cfg->cg_iterations * 2 / 3 : cfg->cg_iterations;
  const int n_inner = cg_total / irls_outer > 1 ? cg_total / irls_outer : 1;
  const R cg_tol = cfg->cg_tol > 0 ? (R)cfg->cg_tol : (cfg->fast ? (R)0.10 : (R)0.02);

  R *planes[10];
  for (int i = 0; i < 10; ++i) planes[i] = (R *)xmalloc(sizeof(R) * n);
  sv.wx = planes[0]; sv.wy = planes[1]; sv.wcx = planes[2]; sv.wcy = planes[3];
  sv.wd1 = planes[4]; sv.wd2 = planes[5]; sv.yc = planes[6]; sv.t1 = planes[7]; sv.t2 = planes[8]; sv.img = planes[9];
  C *cpl[7];
  for (int i = 0; i < 7; ++i) cpl[i] = (C *)xmalloc(sizeof(C) * n);
  sv.c1 = cpl[0]; sv.c2 = cpl[1]; sv.c3 = cpl[2]; sv.cp = cpl[3];
  C *g = cpl[4], *gn = cpl[5], *d = cpl[6];

  for (int outer = 0; outer < irls_outer; ++outer) {
    if (stages) memcpy(stages + (size_t)outer * 2 * n, chi, sizeof(C) * n);
    sv_weights(&sv, chi, eps, eps_c, k);
    sv_grad(&sv, chi, g);
    for (size_t i = 0; i < n; ++i) { d[i].re = -g[i].re; d[i].im = -g[i].im; }
    double gg = sum_abs2(g, n);
    const double gg0 = gg;
    int it = 0;
    for (; it < n_inner; ++it) {
      const double hh = sv_curv(&sv, d);
      if (hh <= hh_floor) break;
      const double gd = dot_real(g, d, n);
      const R alpha = (R)(-0.5 * gd / hh);
      for (size_t i = 0; i < n; ++i) { chi[i].re += alpha * d[i].re; chi[i].im += alpha * d[i].im; }
      sv_grad(&sv, chi, gn);
      const double ggn = sum_abs2(gn, n);
      const R beta = (R)(ggn / (gg > 1e-30 ? gg : 1e-30));
      for (size_t i = 0; i < n; ++i) {
        d[i].re = -gn[i].re + beta * d[i].re;
        d[i].im = -gn[i].im + beta * d[i].im;
      }
      memcpy(g, gn, sizeof(C) * n);
      gg = ggn;
      if (gg < (double)cg_tol * cg_tol * gg0) { ++it; break; }
      if (gg < gg_floor * (double)n) { ++it; break; }
    }
    if (stage_iters) stage_iters[outer] = it;
  }
  if (y)
    for (size_t i = 0; i < n; ++i) y[i] = s[i] - re_prod(chi[i], sv.car[i]);
  for (int i = 0; i < 10; ++i) free(planes[i]);
  for (int i = 0; i < 7; ++i) free(cpl[i]);
  for (int j = 0; j < nn; ++j) { free(dcp[j]); free(wtp[j]); }
  free(sv.rt);
  return aniso;
}

R OR_FN(variational_refine_ex)(const R *s, const R *carrier_r, int h, int w, const or_cfg *cfg,
                               R *chi_r, R *y, int *stage_iters, R *stages, double hh_floor, double gg_floor) {
  return refine_impl(s, carrier_r, h, w, cfg, chi_r, y, stage_iters, stages, hh_floor, gg_floor, 0, NULL, NULL,
                     NULL, 0, 0);
}

R OR_FN(variational_refine)(const R *s, const R *carrier_r, int h, int w, const or_cfg *cfg,
                            R *chi_r, R *y, int *stage_iters) {
  return OR_FN(variational_refine_ex)(s, carrier_r, h, w, cfg, chi_r, y, stage_iters, NULL, 1e-12, 1e-10);
}

/* :198-556 VariationalRefine with nn (<= 16) motion-compensated temporal
 * neighbours (NeighborTerm, variational.h:54-58): sw, conf nn planes of
 * h*w, cw nn complex planes; nu = temporal_strength, eps_t = temporal_eps
 * (> 0).  The 2-D solve's other arguments as variational_refine. */
R OR_FN(variational_refine_nbr)(const R *s, const R *carrier_r, int h, int w, const or_cfg *cfg, int nn,
                                const R *sw, const R *cw, const R *conf, R nu, R eps_t, R *chi_r, R *y,
                                int *stage_iters) {
  return refine_impl(s, carrier_r, h, w, cfg, chi_r, y, stage_iters, NULL, 1e-12, 1e-10, nn, sw, cw, conf, nu,
                     eps_t);
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the 2-D solve (no hvd-core counterpart: the differentiation   */
/* of VariationalRefine the training path needs).                           */
/*                                                                          */
/* Each IRLS stage k (weights w_k = W(chi_{k-1}, S), lagged diffusivity)     */
/* is taken as solved exactly: r_k(chi_k) = dE_k/dchi = 0.  Implicit         */
/* differentiation then gives, from gbar = dL/dchi_k:                       */
/*   a      = (d2E_k)^-1 gbar                  (adjoint solve, same operator) */
/*   dL/dlam_y -= <a, g_data(chi_k)> / lam_y,  dL/dlam_c -= <a, g_prior> / lam_c */
/*   dL/dS  += 2 lam_y L_w Re[a c]             (S in the data term)          */
/*   wbar   = -d<a, r_k>/dw,  chained through W into dL/dS and dL/dchi_{k-1} */
/* and gbar for stage k-1 is dL/dchi_{k-1}; after stage 1 it is dL/dchi0.    */
/* The output luma Y = S - Re[chi c] seeds gbar = dL/dchi + dL/dY (-c_re,    */
/* c_im) and dL/dS = dL/dY.  aniso (an order statistic of chi0) and the      */
/* stage weights' dependence on nothing else are taken as is: aniso is      */
/* locally constant.                                                        */
/* ------------------------------------------------------------------------ */

/* Standard CG on the SPD operator v -> d2E v (= grad with S = 0). */
static int adj_solve(solver *sv, const C *b, C *x, int iters, double tol, C *r, C *p, C *ap) {
  const size_t n = sv->n;
  for (size_t i = 0; i < n; ++i) { x[i].re = 0; x[i].im = 0; r[i] = b[i]; p[i] = b[i]; }
  double rr = sum_abs2(r, n);
  const double rr0 = rr;
  int it = 0;
  for (; it < iters && rr > tol * tol * rr0 && rr > 1e-28 * (double)n * (rr0 > 0 ? 1.0 : 0.0); ++it) {
    sv_grad(sv, p, ap);
    const double pap = dot_real(p, ap, n);
    if (!(pap > 0)) break;
    const double a = rr / pap;
    for (size_t i = 0; i < n; ++i) {
      x[i].re += (R)(a * p[i].re); x[i].im += (R)(a * p[i].im);
      r[i].re -= (R)(a * ap[i].re); r[i].im -= (R)(a * ap[i].im);
    }
    const double rn = sum_abs2(r, n);
    const double be = rn / rr;
    for (size_t i = 0; i < n; ++i) { p[i].re = r[i].re + (R)be * p[i].re; p[i].im = r[i].im + (R)be * p[i].im; }
    rr = rn;
  }
  return it;
}

/* stages: irls_outer planes of chi at each stage's start (as saved by
 * variational_refine_ex); chi_final: chi after the last stage; gchi, gy:
 * dL/dchi and dL/dY of the outputs (gy may be NULL).  Writes dL/dS (gs),
 * dL/dchi0 (gchi0) and dL/dlambda_y, dL/dlambda_c (glam[2]); returns the
 * total adjoint CG iterations. */
int OR_FN(refine_adjoint)(const R *s, const R *carrier_r, int h, int w, const or_cfg *cfg, const R *stages,
                          const R *chi_final, const R *gchi, const R *gy, int adj_iters, double adj_tol, R *gs,
                          R *gchi0, double *glam) {
  solver sv;
  memset(&sv, 0, sizeof sv);
  sv.h = h; sv.w = w; sv.n = (size_t)h * w; sv.car = (const C *)carrier_r;
  const size_t n = sv.n;
  const C *car = sv.car;
  const R eps = (R)cfg->charbonnier_eps, eps_c = (R)cfg->chroma_eps, kc = (R)cfg->structure_coupling;
  const R lambda_c = (R)cfg->lambda_c;
  const R aniso = OR_FN(resolve_chroma_aniso)(stages, h, w, cfg->is_pal, (R)cfg->chroma_aniso);
  sv.mu_h = lambda_c * aniso; sv.mu_v = lambda_c; sv.mu_d = 0;
  sv.lam_y = (R)(cfg->lambda_y > 0 ? cfg->lambda_y : 1.0);
  const int K = cfg->irls_outer > 1 ? cfg->irls_outer : 1;
  R *zero = (R *)xmalloc(sizeof(R) * n);
  memset(zero, 0, sizeof(R) * n);
  R *planes[16];
  for (int i = 0; i < 16; ++i) planes[i] = (R *)xmalloc(sizeof(R) * n);
  sv.wx = planes[0]; sv.wy = planes[1]; sv.wcx = planes[2]; sv.wcy = planes[3];
  sv.wd1 = planes[4]; sv.wd2 = planes[5]; sv.yc = planes[6]; sv.t1 = planes[7]; sv.t2 = planes[8];
  sv.img = planes[9];
  R *q = planes[10], *qx = planes[11], *qy = planes[12], *gxb = planes[13], *gyb = planes[14], *ycb = planes[15];
  C *cpl[14];
  for (int i = 0; i < 14; ++i) cpl[i] = (C *)xmalloc(sizeof(C) * n);
  sv.c1 = cpl[0]; sv.c2 = cpl[1]; sv.c3 = cpl[2]; sv.cp = cpl[3];
  C *gbar = cpl[4], *a = cpl[5], *r = cpl[6], *p = cpl[7], *ap = cpl[8], *dxa = cpl[9], *dya = cpl[10];
  C *dxc = cpl[11], *dyc = cpl[12], *tmpc = cpl[13];
  R *sbar = gs;
  double lam_y_bar = 0, lam_c_bar = 0;
  int total = 0;
  /* seed: Y = S - Re[chi c] */
  const C *gc = (const C *)gchi;
  for (size_t i = 0; i < n; ++i) {
    const R yb = gy ? gy[i] : 0;
    gbar[i].re = gc[i].re - yb * car[i].re;
    gbar[i].im = gc[i].im + yb * car[i].im;
    sbar[i] = yb;
  }
  for (int k = K; k >= 1; --k) {
    const C *chi_prev = (const C *)(stages + (size_t)(k - 1) * 2 * n);
    const C *chi_k = k == K ? (const C *)chi_final : (const C *)(stages + (size_t)k * 2 * n);
    /* weights of stage k from its starting chi */
    sv.s = s;
    sv_weights(&sv, chi_prev, eps, eps_c, kc);
    /* adjoint solve with the stage operator (S = 0 makes grad linear) */
    sv.s = zero;
    total += adj_solve(&sv, gbar, a, adj_iters, adj_tol, r, p, ap);
    /* parameter gradients: data and prior parts of grad at chi_k */
    sv.s = s;
    sv_grad(&sv, chi_k, tmpc);            /* full grad; sv.img holds the data term's L_w yc factor */
    double dd = 0, dp = 0;
    for (size_t i = 0; i < n; ++i) {
      const double gdr = -(double)sv.img[i] * car[i].re, gdi = (double)sv.img[i] * car[i].im;
      dd += (double)a[i].re * gdr + (double)a[i].im * gdi;
      dp += (double)a[i].re * ((double)tmpc[i].re - gdr) + (double)a[i].im * ((double)tmpc[i].im - gdi);
    }
    lam_y_bar -= dd / (double)sv.lam_y;
    lam_c_bar -= dp / (double)lambda_c;
    /* dL/dS through the data term: 2 lam_y L_w q, q = Re[a c] */
    for (size_t i = 0; i < n; ++i) q[i] = re_prod(a[i], car[i]);
    OR_FN(dx)(q, h, w, qx);
    OR_FN(dy)(q, h, w, qy);
    for (size_t i = 0; i < n; ++i) sv.t1[i] = sv.wx[i] * qx[i];
    OR_FN(dxt)(sv.t1, h, w, sv.img);
    for (size_t i = 0; i < n; ++i) sv.t2[i] = sv.wy[i] * qy[i];
    OR_FN(dyt)(sv.t2, h, w, sv.t1);
    for (size_t i = 0; i < n; ++i) sbar[i] += 2 * sv.lam_y * (sv.img[i] + sv.t1[i]);
    /* weight adjoint at chi_k, then through W(chi_prev, S) */
    for (size_t i = 0; i < n; ++i) sv.yc[i] = s[i] - re_prod(chi_k[i], car[i]);
    OR_FN(dx)(sv.yc, h, w, sv.t1);
    OR_FN(dy)(sv.yc, h, w, sv.t2);
    cdx(chi_k, h, w, dxc);
    cdy(chi_k, h, w, dyc);
    cdx(a, h, w, dxa);
    cdy(a, h, w, dya);
    for (size_t i = 0; i < n; ++i) {
      const R wxb = 2 * sv.lam_y * sv.t1[i] * qx[i];
      const R wyb = 2 * sv.lam_y * sv.t2[i] * qy[i];
      const R wcxb = -2 * sv.mu_h * (dxc[i].re * dxa[i].re + dxc[i].im * dxa[i].im);
      const R wcyb = -2 * sv.mu_v * (dyc[i].re * dya[i].re + dyc[i].im * dya[i].im);
      /* reuse: stash the weight adjoints in the weight planes */
      sv.wx[i] = wxb; sv.wy[i] = wyb; sv.wcx[i] = wcxb; sv.wcy[i] = wcyb;
    }
    for (size_t i = 0; i < n; ++i) sv.yc[i] = s[i] - re_prod(chi_prev[i], car[i]);
    OR_FN(dx)(sv.yc, h, w, sv.t1);
    OR_FN(dy)(sv.yc, h, w, sv.t2);
    cdx(chi_prev, h, w, dxc);
    cdy(chi_prev, h, w, dyc);
    const R e2 = eps * eps, ec2 = eps_c * eps_c;
    for (size_t i = 0; i < n; ++i) {
      const R gx = sv.t1[i], gy_ = sv.t2[i];
      const R mx2 = (R)cabs2(dxc[i]), my2 = (R)cabs2(dyc[i]);
      const R ux = gx * gx + e2, uy = gy_ * gy_ + e2;
      const R vx = mx2 + kc * gx * gx + ec2, vy = my2 + kc * gy_ * gy_ + ec2;
      const R dwx = -eps * gx / (ux * (R)sqrt((double)ux));
      const R dwy = -eps * gy_ / (uy * (R)sqrt((double)uy));
      const R fx = -eps_c / (vx * (R)sqrt((double)vx)), fy = -eps_c / (vy * (R)sqrt((double)vy));
      gxb[i] = sv.wx[i] * dwx + sv.wcx[i] * fx * kc * gx;
      gyb[i] = sv.wy[i] * dwy + sv.wcy[i] * fy * kc * gy_;
      dxc[i].re *= sv.wcx[i] * fx; dxc[i].im *= sv.wcx[i] * fx;
      dyc[i].re *= sv.wcy[i] * fy; dyc[i].im *= sv.wcy[i] * fy;
    }
    OR_FN(dxt)(gxb, h, w, ycb);
    OR_FN(dyt)(gyb, h, w, sv.t1);
    for (size_t i = 0; i < n; ++i) ycb[i] += sv.t1[i];
    cdxt(dxc, h, w, dxa);
    cdyt(dyc, h, w, dya);
    for (size_t i = 0; i < n; ++i) {
      sbar[i] += ycb[i];
      gbar[i].re = -ycb[i] * car[i].re + dxa[i].re + dya[i].re;
      gbar[i].im = ycb[i] * car[i].im + dxa[i].im + dya[i].im;
    }
  }
  memcpy(gchi0, gbar, sizeof(C) * n);
  glam[0] = lam_y_bar;
  glam[1] = lam_c_bar;
  free(zero);
  for (int i = 0; i < 16; ++i) free(planes[i]);
  for (int i = 0; i < 14; ++i) free(cpl[i]);
  return total;
}

/* ------------------------------------------------------------------------ */
/* field driver: sequence.cpp:229-268 PrepareFieldObs, :270-281 + :1046    */
/* DecodeFieldWindow 2-D path (one field; temporal strength off), output   */
/* luma = S - Re[chi c] (:1102-1114).                                      */
/* ------------------------------------------------------------------------ */

/* uint16 codes -> IRE (ntsc_geometry.h:301-304 SampleToIre). */
void OR_FN(sample_to_ire_u16)(const uint16_t *raw, size_t n, R blanking, R white, R *out) {
  const R cpi = (white - blanking) / 100;
  for (size_t i = 0; i < n; ++i) out[i] = ((R)raw[i] - blanking) / cpi;
}

/* Lock-in + crop + carrier for one field (sequence.cpp:229-268).  Writes the
 * active composite s (lines x width), the carrier, theta (field_height) and
 * vswitch (field_height, PAL family only; may be NULL on NTSC). */
void OR_FN(prepare_field)(const R *field, const or_geom *g, R chroma_phase_deg, R *s, R *carrier,
                          R *theta, int8_t *vswitch) {
  const int fh = g->field_height, fw = g->field_width;
  const int pal = uses_vswitch(g);
  if (pal) OR_FN(burst_lockin_phase_pal)(field, g, theta, vswitch);
  else OR_FN(burst_lockin_phase)(field, g, theta);
  OR_FN(apply_chroma_phase)(chroma_phase_deg, g, pal ? vswitch : NULL, fh, theta);
  const int fal = g->first_active_field_line;
  const int lal = g->last_active_field_line ? g->last_active_field_line : fh;
  const int a0 = g->active_video_start, a1 = g->active_video_end;
  const int lines = lal - fal > 0 ? lal - fal : 0, width = a1 - a0 > 0 ? a1 - a0 : 0;
  for (int y = 0; y < lines; ++y)
    for (int x = 0; x < width; ++x) s[(size_t)y * width + x] = field[(size_t)(fal + y) * fw + a0 + x];
  OR_FN(make_carrier)(theta + fal, pal ? vswitch + fal : NULL, lines, g, carrier);
}

/* One field through the whole 2-D path: prepare -> init -> refine.  Returns
 * the resolved aniso. */
R OR_FN(decode_field)(const R *field, const or_geom *g, const or_cfg *cfg, R chroma_phase_deg,
                      R *y_out, R *chi_out, R *theta, int8_t *vswitch, int *stage_iters) {
  const int fh = g->field_height;
  const int lal = g->last_active_field_line ? g->last_active_field_line : fh;
  const int lines = lal - g->first_active_field_line, width = g->active_video_end - g->active_video_start;
  const size_t n = (size_t)lines * width;
  R *s = (R *)xmalloc(sizeof(R) * n);
  R *car = (R *)xmalloc(sizeof(R) * 2 * n);
  OR_FN(prepare_field)(field, g, chroma_phase_deg, s, car, theta, vswitch);
  OR_FN(holographic_init)(s, car, lines, width, g->sample_rate, 0, y_out, chi_out);
  /* sequence.cpp:1046 runs VariationalRefine unconditionally on the field
   * path (with cg_iterations == 0 every stage still takes one step). */
  or_cfg c2 = *cfg;
  c2.is_pal = uses_vswitch(g);
  const R aniso = OR_FN(variational_refine)(s, car, lines, width, &c2, chi_out, y_out, stage_iters);
  free(s); free(car);
  return aniso;
}

/* ------------------------------------------------------------------------ */
/* motion: engine/motion.cpp                                                */
/* ------------------------------------------------------------------------ */

/* :21-55 + :391-395 BoxBlur via prefix sums (double accumulators). */
void OR_FN(box_blur_prefix)(R *a, int h, int w, int r) {
  const R norm = (R)1 / (R)(2 * r + 1);
  double *c = (double *)xmalloc(sizeof(double) * (size_t)((h > w ? h : w) + 1));
  for (int x = 0; x < w; ++x) {
    c[0] = 0;
    for (int y = 0; y < h; ++y) c[y + 1] = c[y] + a[(size_t)y * w + x];
    for (int y = 0; y < h; ++y) {
      const int y0 = y - r > 0 ? y - r : 0, y1 = y + r + 1 < h ? y + r + 1 : h;
      a[(size_t)y * w + x] = (R)((c[y1] - c[y0]) * norm);
    }
  }
  for (int y = 0; y < h; ++y) {
    c[0] = 0;
    for (int x = 0; x < w; ++x) c[x + 1] = c[x] + a[(size_t)y * w + x];
    for (int x = 0; x < w; ++x) {
      const int x0 = x - r > 0 ? x - r : 0, x1 = x + r + 1 < w ? x + r + 1 : w;
      a[(size_t)y * w + x] = (R)((c[x1] - c[x0]) * norm);
    }
  }
  free(c);
}

/* :59-78 Decimate: f x f block average, remainder cropped. */
static R *decimate(const R *a, int h, int w, int f, int *hh, int *ww) {
  *hh = h / f; *ww = w / f;
  R *o = (R *)xmalloc(sizeof(R) * (size_t)(*hh) * (size_t)(*ww));
  const R norm = (R)1 / (R)(f * f);
  for (int y = 0; y < *hh; ++y)
    for (int x = 0; x < *ww; ++x) {
      R acc = 0;
      for (int dy = 0; dy < f; ++dy)
        for (int dx = 0; dx < f; ++dx) acc += a[(size_t)(y * f + dy) * w + x * f + dx];
      o[(size_t)y * (*ww) + x] = acc * norm;
    }
  return o;
}

/* :286-298 RangeFromCentre: 0, -1, +1, -2, +2, ... within [lo, hi]. */
static int range_from_centre(int lo, int hi, int *v) {
  int n = 0;
  if (lo > hi) return 0;
  if (lo <= 0 && 0 <= hi) v[n++] = 0;
  const int mx = -lo > hi ? -lo : hi;
  for (int d = 1; d <= mx; ++d) {
    if (-d >= lo && -d <= hi) v[n++] = -d;
    if (d >= lo && d <= hi) v[n++] = d;
  }
  if (!n) for (int i = lo; i <= hi; ++i) v[n++] = i;
  return n;
}

/* :105-260 BmPass, exhaustive (no partial-distortion elimination: PDE only
 * skips candidates that cannot win, so the selected field is the same).
 * A = current, B = reference; candidate (dy, dx) compares A(y, x) with
 * B(y - dy, x - dx), B edge-clamped. */
static void bm_pass(const R *A, const R *B, int h, int w, int tile, const int *dys, int ndy, const int *dxs, int ndx,
                    const R *base_dy, const R *base_dx, R *mdy, R *mdx, R *best, R *cost0) {
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile;
  for (int i = 0; i < th * tw; ++i) { mdy[i] = 0; mdx[i] = 0; best[i] = (R)INFINITY; if (cost0) cost0[i] = 0; }
  for (int a = 0; a < ndy; ++a) {
    for (int b = 0; b < ndx; ++b) {
      const int dy = dys[a], dx = dxs[b];
      const R bias = (R)1 + (R)0.02 * (R)(abs(dy) + abs(dx));
      for (int ty = 0; ty < th; ++ty) {
        const int ye = (ty + 1) * tile < h ? (ty + 1) * tile : h;
        for (int tx = 0; tx < tw; ++tx) {
          const int xe = (tx + 1) * tile < w ? (tx + 1) * tile : w;
          int sy = dy, sx = dx;
          if (base_dy) {
            sy = (int)lround((double)base_dy[ty * tw + tx]) + dy;
            sx = (int)lround((double)base_dx[ty * tw + tx]) + dx;
          }
          R acc = 0;
          for (int y = ty * tile; y < ye; ++y) {
            int yy = y - sy; yy = yy < 0 ? 0 : (yy >= h ? h - 1 : yy);
            for (int x = tx * tile; x < xe; ++x) {
              int xx = x - sx; xx = xx < 0 ? 0 : (xx >= w ? w - 1 : xx);
              const R df = A[(size_t)y * w + x] - B[(size_t)yy * w + xx];
              acc += df * df;
            }
          }
          if (cost0 && dy == 0 && dx == 0) cost0[ty * tw + tx] = acc;
          const R cost = acc * bias;
          if (cost < best[ty * tw + tx]) {
            best[ty * tw + tx] = cost;
            if (base_dy) {
              mdy[ty * tw + tx] = base_dy[ty * tw + tx] + (R)dy;
              mdx[ty * tw + tx] = base_dx[ty * tw + tx] + (R)dx;
            } else {
              mdy[ty * tw + tx] = (R)dy;
              mdx[ty * tw + tx] = (R)dx;
            }
          }
        }
      }
    }
  }
}

/* :321-387 MotionConfidence. */
static void motion_confidence(const R *A, int h, int w, const R *best, int tile, int th, int tw, R *conf) {
  const int nt = th * tw;
  R *sum = (R *)xmalloc(sizeof(R) * (size_t)nt), *sum2 = (R *)xmalloc(sizeof(R) * (size_t)nt);
  R *npix = (R *)xmalloc(sizeof(R) * (size_t)nt), *ratios = (R *)xmalloc(sizeof(R) * (size_t)nt);
  for (int y = 0; y < h; ++y) {
    const int ty = y / tile < th - 1 ? y / tile : th - 1;
    for (int x = 0; x < w; ++x) {
      const int tx = x / tile < tw - 1 ? x / tile : tw - 1;
      const R v = A[(size_t)y * w + x];
      sum[ty * tw + tx] += v;
      sum2[ty * tw + tx] += v * v;
      npix[ty * tw + tx] += 1;
    }
  }
  for (int i = 0; i < nt; ++i) {
    const R np = npix[i] > 1 ? npix[i] : 1;
    const R norm = (R)1 / np;
    const R mean = sum[i] * norm;
    const R var = sum2[i] * norm - mean * mean;
    const R energy = (var + 1) * np;
    R c = 1 - best[i] / energy;
    conf[i] = c < 0 ? 0 : (c > 1 ? 1 : c);
    ratios[i] = best[i] / np;
  }
  const R med = median_np(ratios, nt);
  for (int i = 0; i < nt; ++i) {
    const R ex = ratios[i] - med > 0 ? ratios[i] - med : 0;
    conf[i] *= (med + 1) / (med + 1 + ex);
    conf[i] *= (R)6 / ((R)6 + med);
  }
  free(sum); free(sum2); free(npix); free(ratios);
}

/* :397-488 EstimateMotion: pre-blur r=2, 4x-decimated coarse search,
 * +/-3 refine, parabolic half-pel, prefer-zero margin, confidence.
 * Outputs are th x tw tiles. */
void OR_FN(estimate_motion)(const R *y_ref, const R *y_cur, int h, int w, int tile, int search,
                            R *out_dy, R *out_dx, R *out_conf) {
  const size_t n = (size_t)h * w;
  R *A = (R *)xmalloc(sizeof(R) * n), *B = (R *)xmalloc(sizeof(R) * n);
  memcpy(A, y_cur, sizeof(R) * n);
  memcpy(B, y_ref, sizeof(R) * n);
  OR_FN(box_blur_prefix)(A, h, w, 2);
  OR_FN(box_blur_prefix)(B, h, w, 2);
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw;
  const int f = 4;
  const int cs = (search + f - 1) / f > 1 ? (search + f - 1) / f : 1;
  int hh, ww;
  R *Ad = decimate(A, h, w, f, &hh, &ww), *Bd = decimate(B, h, w, f, &hh, &ww);
  const int ct = tile / f > 4 ? tile / f : 4;
  const int cth = (hh + ct - 1) / ct, ctw = (ww + ct - 1) / ct;
  int rng[64];
  const int nr = range_from_centre(-cs, cs, rng);
  R *cdy = (R *)xmalloc(sizeof(R) * (size_t)(cth * ctw)), *cdx = (R *)xmalloc(sizeof(R) * (size_t)(cth * ctw));
  R *cbest = (R *)xmalloc(sizeof(R) * (size_t)(cth * ctw));
  bm_pass(Ad, Bd, hh, ww, ct, rng, nr, rng, nr, NULL, NULL, cdy, cdx, cbest, NULL);
  R *bdy = (R *)xmalloc(sizeof(R) * (size_t)nt), *bdx = (R *)xmalloc(sizeof(R) * (size_t)nt);
  for (int ty = 0; ty < th; ++ty) {
    const int gy = (ty * cth) / th < cth - 1 ? (ty * cth) / th : cth - 1;
    for (int tx = 0; tx < tw; ++tx) {
      const int gx = (tx * ctw) / tw < ctw - 1 ? (tx * ctw) / tw : ctw - 1;
      bdy[ty * tw + tx] = cdy[gy * ctw + gx] * (R)f;
      bdx[ty * tw + tx] = cdx[gy * ctw + gx] * (R)f;
    }
  }
  int r3[8];
  const int n3 = range_from_centre(-3, 3, r3);
  R *mdy = (R *)xmalloc(sizeof(R) * (size_t)nt), *mdx = (R *)xmalloc(sizeof(R) * (size_t)nt);
  R *best = (R *)xmalloc(sizeof(R) * (size_t)nt);
  bm_pass(A, B, h, w, tile, r3, n3, r3, n3, bdy, bdx, mdy, mdx, best, NULL);
  R *cm[4];
  const int offs[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
  R *tdy = (R *)xmalloc(sizeof(R) * (size_t)nt), *tdx = (R *)xmalloc(sizeof(R) * (size_t)nt);
  R *jy = (R *)xmalloc(sizeof(R) * (size_t)nt), *jx = (R *)xmalloc(sizeof(R) * (size_t)nt);
  int zero = 0;
  for (int k = 0; k < 4; ++k) {
    cm[k] = (R *)xmalloc(sizeof(R) * (size_t)nt);
    for (int i = 0; i < nt; ++i) { tdy[i] = mdy[i] + (R)offs[k][0]; tdx[i] = mdx[i] + (R)offs[k][1]; }
    bm_pass(A, B, h, w, tile, &zero, 1, &zero, 1, tdy, tdx, jy, jx, cm[k], NULL);
  }
  R *sdy = (R *)xmalloc(sizeof(R) * (size_t)nt), *sdx = (R *)xmalloc(sizeof(R) * (size_t)nt);
  for (int i = 0; i < nt; ++i) {
    const R c0 = best[i];
    R den = cm[0][i] - 2 * c0 + cm[1][i], sub = 0;
    if (den > (R)1e-9) sub = (R)0.5 * (cm[0][i] - cm[1][i]) / den;
    sdy[i] = sub < (R)-0.5 ? (R)-0.5 : (sub > (R)0.5 ? (R)0.5 : sub);
    den = cm[2][i] - 2 * c0 + cm[3][i]; sub = 0;
    if (den > (R)1e-9) sub = (R)0.5 * (cm[2][i] - cm[3][i]) / den;
    sdx[i] = sub < (R)-0.5 ? (R)-0.5 : (sub > (R)0.5 ? (R)0.5 : sub);
  }
  R *best0 = (R *)xmalloc(sizeof(R) * (size_t)nt), *cost0 = (R *)xmalloc(sizeof(R) * (size_t)nt);
  bm_pass(A, B, h, w, tile, &zero, 1, &zero, 1, NULL, NULL, jy, jx, best0, cost0);
  for (int i = 0; i < nt; ++i) {
    if (best0[i] < best[i]) { mdy[i] = 0; mdx[i] = 0; best[i] = best0[i]; }
    if (best[i] > (R)0.85 * cost0[i]) { mdy[i] = 0; mdx[i] = 0; best[i] = cost0[i]; sdy[i] = 0; sdx[i] = 0; }
  }
  motion_confidence(A, h, w, best, tile, th, tw, out_conf);
  for (int i = 0; i < nt; ++i) { out_dy[i] = mdy[i] + sdy[i]; out_dx[i] = mdx[i] + sdx[i]; }
  free(A); free(B); free(Ad); free(Bd); free(cdy); free(cdx); free(cbest); free(bdy); free(bdx);
  free(mdy); free(mdx); free(best); free(tdy); free(tdx); free(jy); free(jx);
  for (int k = 0; k < 4; ++k) free(cm[k]);
  free(sdy); free(sdx); free(best0); free(cost0);
}

/* ------------------------------------------------------------------------ */
/* Integer-SAD block matching over uint16 samples.  This operator has no    */
/* hvd-core counterpart (the engine matches float SSD on 32 px tiles); it   */
/* is the project's own specification:                                      */
/*   blocks bw x bh tile the plane (edge blocks clipped to the plane);      */
/*   candidate (dy, dx), |dy| <= sy, |dx| <= sx, compares cur(y, x) with    */
/*   ref(y - dy, x - dx), ref edge-clamped (the BmPass convention);         */
/*   cost = sum |cur - ref| in int32;                                       */
/*   winner = least cost, then least |dy| + |dx|, then first in raster      */
/*   order (dy ascending, then dx ascending).                               */
/* ------------------------------------------------------------------------ */
void OR_FN(block_match_sad_u16)(const uint16_t *ref, const uint16_t *cur, int h, int w, int bh, int bw, int sy,
                                int sx, int16_t *vy, int16_t *vx, int32_t *sad) {
  const int nby = (h + bh - 1) / bh, nbx = (w + bw - 1) / bw;
  for (int by = 0; by < nby; ++by)
    for (int bx = 0; bx < nbx; ++bx) {
      int32_t best = INT32_MAX;
      int bdy = 0, bdx = 0;
      const int y0 = by * bh, y1 = y0 + bh < h ? y0 + bh : h;
      const int x0 = bx * bw, x1 = x0 + bw < w ? x0 + bw : w;
      for (int dy = -sy; dy <= sy; ++dy)
        for (int dx = -sx; dx <= sx; ++dx) {
          int32_t acc = 0;
          for (int y = y0; y < y1; ++y) {
            int yy = y - dy; yy = yy < 0 ? 0 : (yy >= h ? h - 1 : yy);
            for (int x = x0; x < x1; ++x) {
              int xx = x - dx; xx = xx < 0 ? 0 : (xx >= w ? w - 1 : xx);
              acc += abs((int)cur[(size_t)y * w + x] - (int)ref[(size_t)yy * w + xx]);
            }
          }
          const int m = abs(dy) + abs(dx), bm = abs(bdy) + abs(bdx);
          if (acc < best || (acc == best && m < bm)) { best = acc; bdy = dy; bdx = dx; }
        }
      vy[by * nbx + bx] = (int16_t)bdy;
      vx[by * nbx + bx] = (int16_t)bdx;
      sad[by * nbx + bx] = best;
    }
}

/* ------------------------------------------------------------------------ */
/* temporal neighbours' warps: engine/temporal.cpp                          */
/* ------------------------------------------------------------------------ */

/* :19-40 3x3 component-wise median, edge-clamped; grids under 3x3 copied. */
static void median3(const R *v, int th, int tw, R *out) {
  if (th < 3 || tw < 3) { memcpy(out, v, sizeof(R) * (size_t)th * tw); return; }
  for (int y = 0; y < th; ++y)
    for (int x = 0; x < tw; ++x) {
      R a[9];
      int k = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int yy = y + dy < 0 ? 0 : (y + dy > th - 1 ? th - 1 : y + dy);
          const int xx = x + dx < 0 ? 0 : (x + dx > tw - 1 ? tw - 1 : x + dx);
          a[k++] = v[(size_t)yy * tw + xx];
        }
      for (int i = 1; i < 9; ++i) /* insertion sort: the 5th smallest */
        for (int j = i; j > 0 && a[j - 1] > a[j]; --j) { const R t = a[j]; a[j] = a[j - 1]; a[j - 1] = t; }
      out[(size_t)y * tw + x] = a[4];
    }
}

/* :73-87 the tile-centre bracket of a pixel coordinate. */
static void tile_bracket(int coord, int n_tiles, int tile, int *lo, R *frac) {
  if (n_tiles > 1) {
    const float rel = (float)coord / (float)tile - 0.5f;
    int l = (int)floorf(rel);
    l = l < 0 ? 0 : (l > n_tiles - 2 ? n_tiles - 2 : l);
    const float centre = ((float)l + 0.5f) * (float)tile;
    float f = ((float)coord - centre) / (float)tile;
    f = f < 0.f ? 0.f : (f > 1.f ? 1.f : f);
    *lo = l;
    *frac = (R)f;
  } else {
    *lo = 0;
    *frac = 0;
  }
}

/* :44-113 VectorsPerPixel: 3x3-median outlier snap (|dy-my| + |dx-mx| > 3
 * takes the median), then bilinear interpolation between tile centres. */
void OR_FN(vectors_per_pixel)(const R *mdy_in, const R *mdx_in, int th, int tw, int tile, int out_h, int out_w,
                              R *dy, R *dx) {
  const size_t nt = (size_t)th * tw;
  R *my = (R *)xmalloc(sizeof(R) * nt), *mx = (R *)xmalloc(sizeof(R) * nt);
  R *vy = (R *)xmalloc(sizeof(R) * nt), *vx = (R *)xmalloc(sizeof(R) * nt);
  median3(mdy_in, th, tw, my);
  median3(mdx_in, th, tw, mx);
  for (size_t i = 0; i < nt; ++i) {
    const R dis = (R)fabs((double)(mdy_in[i] - my[i])) + (R)fabs((double)(mdx_in[i] - mx[i]));
    vy[i] = dis > 3 ? my[i] : mdy_in[i];
    vx[i] = dis > 3 ? mx[i] : mdx_in[i];
  }
  for (int y = 0; y < out_h; ++y) {
    int iy;
    R fy;
    tile_bracket(y, th, tile, &iy, &fy);
    const int iy1 = iy + 1 < th - 1 ? iy + 1 : th - 1;
    for (int x = 0; x < out_w; ++x) {
      int ix;
      R fx;
      tile_bracket(x, tw, tile, &ix, &fx);
      const int ix1 = ix + 1 < tw - 1 ? ix + 1 : tw - 1;
      for (int c = 0; c < 2; ++c) {
        const R *v = c ? vx : vy;
        const R v00 = v[(size_t)iy * tw + ix], v10 = v[(size_t)iy1 * tw + ix];
        const R v01 = v[(size_t)iy * tw + ix1], v11 = v[(size_t)iy1 * tw + ix1];
        (c ? dx : dy)[(size_t)y * out_w + x] =
            v00 * (1 - fy) * (1 - fx) + v10 * fy * (1 - fx) + v01 * (1 - fy) * fx + v11 * fy * fx;
      }
    }
  }
  free(my); free(mx); free(vy); free(vx);
}

/* :117-132, :198-215 WarpByTiles / WarpByTilesComplex: out(y, x) =
 * a(clamp(y - lround(dy)), clamp(x - lround(dx))), `ch` values per pixel. */
void OR_FN(warp_by_tiles)(const R *a, int h, int w, int ch, const R *dy, const R *dx, R *out) {
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const size_t i = (size_t)y * w + x;
      int sy = y - (int)lround((double)dy[i]), sx = x - (int)lround((double)dx[i]);
      sy = sy < 0 ? 0 : (sy > h - 1 ? h - 1 : sy);
      sx = sx < 0 ? 0 : (sx > w - 1 ? w - 1 : sx);
      for (int c = 0; c < ch; ++c) out[i * ch + c] = a[((size_t)sy * w + sx) * ch + c];
    }
}

/* :140-166 WarpBilinearTiles with per-pixel vectors on the output grid. */
void OR_FN(warp_bilinear)(const R *a, int hs, int ws, const R *dy, const R *dx, R row_offset, int out_h, int out_w,
                          R *out) {
  for (int y = 0; y < out_h; ++y)
    for (int x = 0; x < out_w; ++x) {
      const size_t i = (size_t)y * out_w + x;
      R sy = (R)y - dy[i] + row_offset, sx = (R)x - dx[i];
      const R my = (R)hs - (R)1.001f, mx = (R)ws - (R)1.001f;
      sy = sy < 0 ? 0 : (sy > my ? my : sy);
      sx = sx < 0 ? 0 : (sx > mx ? mx : sx);
      const int y0 = (int)sy, x0 = (int)sx;
      const R fy = sy - (R)y0, fx = sx - (R)x0;
      const int y1 = y0 + 1 < hs - 1 ? y0 + 1 : hs - 1, x1 = x0 + 1 < ws - 1 ? x0 + 1 : ws - 1;
      out[i] = a[(size_t)y0 * ws + x0] * (1 - fy) * (1 - fx) + a[(size_t)y1 * ws + x0] * fy * (1 - fx) +
               a[(size_t)y0 * ws + x1] * (1 - fy) * fx + a[(size_t)y1 * ws + x1] * fy * fx;
    }
}

/* :217-249 UpsampleConfidence (nearest tile, squared, BoxBlur radius 8) and
 * the fast form (squared tile map through VectorsPerPixel). */
void OR_FN(upsample_confidence)(const R *conf, int th, int tw, int tile, int out_h, int out_w, int fast, R *out) {
  if (fast) {
    const size_t nt = (size_t)th * tw;
    R *c2 = (R *)xmalloc(sizeof(R) * nt), *tmp = (R *)xmalloc(sizeof(R) * (size_t)out_h * out_w);
    for (size_t i = 0; i < nt; ++i) c2[i] = conf[i] * conf[i];
    OR_FN(vectors_per_pixel)(c2, c2, th, tw, tile, out_h, out_w, out, tmp);
    free(c2); free(tmp);
    return;
  }
  for (int y = 0; y < out_h; ++y) {
    const int ty = y / tile < th - 1 ? y / tile : th - 1;
    for (int x = 0; x < out_w; ++x) {
      const int tx = x / tile < tw - 1 ? x / tile : tw - 1;
      const R c = conf[(size_t)ty * tw + tx];
      out[(size_t)y * out_w + x] = c * c;
    }
  }
  OR_FN(box_blur_prefix)(out, out_h, out_w, 8);
}

/* :169-196 ComplexCoherence: |<z1 conj z2>| / (sqrt(<|z1|^2><|z2|^2>) + 1e-6),
 * box-blurred averages of radius r, clamped to [0, 1]. */
void OR_FN(complex_coherence)(const R *z1, const R *z2, int h, int w, int r, R *out) {
  const size_t n = (size_t)h * w;
  R *p = (R *)xmalloc(sizeof(R) * n * 4);
  R *cre = p, *cim = p + n, *m1 = p + 2 * n, *m2 = p + 3 * n;
  for (size_t i = 0; i < n; ++i) {
    const R ar = z1[2 * i], ai = z1[2 * i + 1], br = z2[2 * i], bi = z2[2 * i + 1];
    cre[i] = ar * br + ai * bi;
    cim[i] = ai * br - ar * bi;
    m1[i] = ar * ar + ai * ai;
    m2[i] = br * br + bi * bi;
  }
  for (int k = 0; k < 4; ++k) OR_FN(box_blur_prefix)(p + k * n, h, w, r);
  for (size_t i = 0; i < n; ++i) {
    const R num = (R)sqrt((double)(cre[i] * cre[i] + cim[i] * cim[i]));
    const R pr = m1[i] * m2[i];
    const R den = (R)sqrt((double)(pr > 0 ? pr : 0)) + (R)1e-6f;
    const R g = num / den;
    out[i] = g < 0 ? 0 : (g > 1 ? 1 : g);
  }
  free(p);
}

/* :253-287 MotionCompensatePrev with a given motion field: one per-pixel
 * vector interpolation shared by the composite and carrier warps, and the
 * confidence upsample.  car / out_car are interleaved complex. */
void OR_FN(motion_compensate_prev)(const R *comp, const R *car, int h, int w, const R *mdy, const R *mdx,
                                   const R *mconf, int th, int tw, int tile, int fast, R *out_comp, R *out_car,
                                   R *out_conf) {
  const size_t n = (size_t)h * w;
  R *dy = (R *)xmalloc(sizeof(R) * n), *dx = (R *)xmalloc(sizeof(R) * n);
  OR_FN(vectors_per_pixel)(mdy, mdx, th, tw, tile, h, w, dy, dx);
  OR_FN(warp_by_tiles)(comp, h, w, 1, dy, dx, out_comp);
  OR_FN(warp_by_tiles)(car, h, w, 2, dy, dx, out_car);
  OR_FN(upsample_confidence)(mconf, th, tw, tile, h, w, fast, out_conf);
  free(dy); free(dx);
}

/* ------------------------------------------------------------------------ */
/* 3-D sequence pieces: engine/ntsc_geometry.cpp, motion.cpp, variational.cpp */
/* ------------------------------------------------------------------------ */

/* ntsc_geometry.cpp:49-71 EstimateNoiseIre: the second difference at the
 * 4-sample subcarrier period, d = S[x] - 2 S[x+4] + S[x+8]; the robust
 * spread P25(|d - median d|) / 0.3186 / sqrt(6).  Median and P25 are the
 * elements n/2 and n/4 of the sorted set (std::nth_element). */
R OR_FN(estimate_noise_ire)(const R *s, int h, int w) {
  if (w <= 8 || h <= 0) return 0;
  const int m = w - 8, n = h * m;
  R *d = (R *)xmalloc(sizeof(R) * (size_t)n);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < m; ++x) {
      const R *r = s + (size_t)y * w + x;
      d[y * m + x] = r[0] - (R)2 * r[4] + r[8];
    }
  const R med = median_upper(d, n);
  for (int i = 0; i < n; ++i) d[i] = (R)fabs((double)(d[i] - med));
  const R q = kth_smallest(d, n, n / 4);
  free(d);
  return q / (R)0.3186 / (R)sqrt((double)(R)6);
}

/* motion.cpp:490-525 VerifyMotion: the rounded trajectory prediction and the
 * zero vector, one SSD each per tile on the r = 2 blurred planes; the better
 * wins (zero on ties), the prediction keeping its full float value; the
 * motion confidence of the winner.  pdy, pdx, out_* are th x tw tiles. */
void OR_FN(verify_motion)(const R *y_ref, const R *y_cur, int h, int w, const R *pdy, const R *pdx, int tile,
                          R *out_dy, R *out_dx, R *out_conf) {
  const size_t n = (size_t)h * w;
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw;
  R *A = (R *)xmalloc(sizeof(R) * n), *B = (R *)xmalloc(sizeof(R) * n);
  memcpy(A, y_cur, sizeof(R) * n);
  memcpy(B, y_ref, sizeof(R) * n);
  OR_FN(box_blur_prefix)(A, h, w, 2);
  OR_FN(box_blur_prefix)(B, h, w, 2);
  R *pyi = (R *)xmalloc(sizeof(R) * (size_t)nt), *pxi = (R *)xmalloc(sizeof(R) * (size_t)nt);
  for (int i = 0; i < nt; ++i) {
    pyi[i] = (R)lround((double)pdy[i]);
    pxi[i] = (R)lround((double)pdx[i]);
  }
  R *jy = (R *)xmalloc(sizeof(R) * (size_t)nt), *jx = (R *)xmalloc(sizeof(R) * (size_t)nt);
  R *bp = (R *)xmalloc(sizeof(R) * (size_t)nt), *bz = (R *)xmalloc(sizeof(R) * (size_t)nt);
  R *best = (R *)xmalloc(sizeof(R) * (size_t)nt);
  int zero = 0;
  bm_pass(A, B, h, w, tile, &zero, 1, &zero, 1, pyi, pxi, jy, jx, bp, NULL);
  bm_pass(A, B, h, w, tile, &zero, 1, &zero, 1, NULL, NULL, jy, jx, bz, NULL);
  for (int i = 0; i < nt; ++i) {
    if (bz[i] <= bp[i]) {
      out_dy[i] = 0; out_dx[i] = 0; best[i] = bz[i];
    } else {
      out_dy[i] = pdy[i]; out_dx[i] = pdx[i]; best[i] = bp[i];
    }
  }
  motion_confidence(A, h, w, best, tile, th, tw, out_conf);
  free(A); free(B); free(pyi); free(pxi); free(jy); free(jx); free(bp); free(bz); free(best);
}

/* motion.cpp:527-562 FitTrajectory: per tile the numpy median over the
 * measurements (offset != 0, confidence > 0.15) of (dy_k - h_k) / k and
 * dx_k / k; 0 where none qualifies.  dy, dx, conf: m planes of nt tiles. */
void OR_FN(fit_trajectory)(int m, const int *offsets, const R *pshift, const R *dy, const R *dx, const R *conf,
                           int nt, R *vy, R *vx) {
  R *sy = (R *)xmalloc(sizeof(R) * (size_t)(m + 1)), *sx = (R *)xmalloc(sizeof(R) * (size_t)(m + 1));
  for (int i = 0; i < nt; ++i) {
    int k = 0;
    for (int j = 0; j < m; ++j) {
      if (offsets[j] == 0) continue;
      if (conf[(size_t)j * nt + i] <= (R)0.15) continue;
      const R o = (R)offsets[j];
      sy[k] = (dy[(size_t)j * nt + i] - pshift[j]) / o;
      sx[k] = dx[(size_t)j * nt + i] / o;
      ++k;
    }
    vy[i] = k ? median_np(sy, k) : 0;
    vx[i] = k ? median_np(sx, k) : 0;
  }
  free(sy); free(sx);
}

/* motion.cpp:564-601 TrajectorySnap: where >= 3 measurements lie within
 * 1.5 px (|ddy| + |ddx|) of the fitted k v + h_k, those measurements move
 * onto it.  In place; fewer than two measurements: no-op. */
void OR_FN(trajectory_snap)(int m, const int *offsets, const R *pshift, R *dy, R *dx, const R *conf, int nt) {
  if (m < 2) return;
  R *vy = (R *)xmalloc(sizeof(R) * (size_t)nt), *vx = (R *)xmalloc(sizeof(R) * (size_t)nt);
  R *agree = (R *)xmalloc(sizeof(R) * (size_t)nt);
  OR_FN(fit_trajectory)(m, offsets, pshift, dy, dx, conf, nt, vy, vx);
  const R tol = (R)1.5;
  for (int j = 0; j < m; ++j) {
    const R o = (R)offsets[j];
    for (int i = 0; i < nt; ++i) {
      const R py = o * vy[i] + pshift[j], px = o * vx[i];
      const R dev = (R)fabs((double)(dy[(size_t)j * nt + i] - py)) + (R)fabs((double)(dx[(size_t)j * nt + i] - px));
      if (dev <= tol) agree[i] += 1;
    }
  }
  for (int j = 0; j < m; ++j) {
    const R o = (R)offsets[j];
    for (int i = 0; i < nt; ++i) {
      if (agree[i] < (R)3) continue;
      const R py = o * vy[i] + pshift[j], px = o * vx[i];
      const R dev = (R)fabs((double)(dy[(size_t)j * nt + i] - py)) + (R)fabs((double)(dx[(size_t)j * nt + i] - px));
      if (dev <= tol) {
        dy[(size_t)j * nt + i] = py;
        dx[(size_t)j * nt + i] = px;
      }
    }
  }
  free(vy); free(vx); free(agree);
}

/* variational.cpp:558-891 VariationalRefineJoint: Y and chi both free; the
 * exact data term (S - Y - Re[chi c])^2, per neighbour nu wt (S_w - Y -
 * Re[chi c_w])^2 with wt lagged per stage, the anchor nu_a w_a (|Y - Y_a|^2 +
 * |chi - chi_a|^2), Charbonnier priors on grad Y and the chroma priors of
 * the 2-D solve (the oriented ones too), the diffusivities scaled by
 * 1 + nu sum wt.  y, chi: the start on entry, the result on return.
 * Neighbour planes as in variational_refine_nbr; a_* may be NULL (no anchor).
 * stage_iters (optional) receives the CG iterations per stage. */
R OR_FN(joint_refine)(const R *s, const R *carrier_r, int h, int w, const or_cfg *cfg, R *y, R *chi_r, int nn,
                      const R *sw, const R *cw_r, const R *conf, R nu_in, R eps_t, const R *a_luma,
                      const R *a_chroma_r, const R *a_weight, R nu_a_in, int *stage_iters) {
  const size_t n = (size_t)h * w;
  const C *car = (const C *)carrier_r, *cw = (const C *)cw_r, *a_chroma = (const C *)a_chroma_r;
  C *chi = (C *)chi_r;
  const R eps = (R)cfg->charbonnier_eps, eps_c = (R)cfg->chroma_eps, k = (R)cfg->structure_coupling;
  const R lambda_c = (R)cfg->lambda_c;
  const R aniso = OR_FN(resolve_chroma_aniso)(chi_r, h, w, cfg->is_pal, (R)cfg->chroma_aniso);
  R mu_h = lambda_c * aniso, mu_v = lambda_c, mu_d = lambda_c * (R)cfg->diag_prior * (R)0.5;
  {
    const R tot0 = lambda_c * (1 + aniso);
    const R sc = mu_d > 0 ? tot0 / (tot0 + 2 * mu_d) : 1;
    mu_h *= sc; mu_v *= sc; mu_d *= sc;
  }
  if (nn > 16) nn = 16;
  const R nu = nn > 0 ? nu_in : 0;
  const R nu_a = a_weight ? nu_a_in : 0;
  const int irls_outer = cfg->irls_outer > 1 ? cfg->irls_outer : 1;
  const int cg_total = cfg->fast ? cfg->cg_iterations * 2 / 3 : cfg->cg_iterations;
  const int n_inner = cg_total / irls_outer > 1 ? cg_total / irls_outer : 1;
  const R cg_tol = cfg->cg_tol > 0 ? (R)cfg->cg_tol : (cfg->fast ? (R)0.10 : (R)0.02);
  const R e2 = eps * eps, ec2 = eps_c * eps_c, et2 = eps_t * eps_t;

  R *rp[14];
  for (int i = 0; i < 14; ++i) rp[i] = (R *)xmalloc(sizeof(R) * n);
  R *wx = rp[0], *wy = rp[1], *wcx = rp[2], *wcy = rp[3], *wd1 = rp[4], *wd2 = rp[5];
  R *gxY = rp[6], *gyY = rp[7], *t1 = rp[8], *t2 = rp[9], *gY = rp[10], *gYn = rp[11], *dY = rp[12], *tr = rp[13];
  C *cp[8];
  for (int i = 0; i < 8; ++i) cp[i] = (C *)xmalloc(sizeof(C) * n);
  C *c1 = cp[0], *c2 = cp[1], *c3 = cp[2], *gC = cp[3], *gCn = cp[4], *dC = cp[5], *c4 = cp[6], *c5 = cp[7];
  R *wt = (R *)xmalloc(sizeof(R) * n * (size_t)(nn > 0 ? nn : 1));

  /* :615-683 grad(Y, chi) */
#define JGRAD(YY, CC, GY, GC)                                                                          \
  do {                                                                                                 \
    for (size_t i = 0; i < n; ++i) {                                                                   \
      const R r0 = s[i] - (YY)[i] - re_prod((CC)[i], car[i]);                                          \
      const R f = (R)-2 * r0;                                                                          \
      (GY)[i] = f;                                                                                     \
      (GC)[i].re = f * car[i].re;                                                                      \
      (GC)[i].im = f * -car[i].im;                                                                     \
    }                                                                                                  \
    for (int j = 0; j < nn; ++j) {                                                                     \
      const R *swj = sw + (size_t)j * n, *wtj = wt + (size_t)j * n;                                    \
      const C *cwj = cw + (size_t)j * n;                                                               \
      for (size_t i = 0; i < n; ++i) {                                                                 \
        const R rk = swj[i] - (YY)[i] - re_prod((CC)[i], cwj[i]);                                      \
        (GY)[i] += (R)-2 * nu * wtj[i] * rk;                                                           \
        const R f = (R)-2 * nu * wtj[i] * rk;                                                          \
        (GC)[i].re += f * cwj[i].re;                                                                   \
        (GC)[i].im += f * -cwj[i].im;                                                                  \
      }                                                                                                \
    }                                                                                                  \
    if (nu_a > 0) {                                                                                    \
      for (size_t i = 0; i < n; ++i) {                                                                 \
        const R f = (R)2 * nu_a * a_weight[i];                                                         \
        (GY)[i] += f * ((YY)[i] - a_luma[i]);                                                          \
        (GC)[i].re += f * ((CC)[i].re - a_chroma[i].re);                                               \
        (GC)[i].im += f * ((CC)[i].im - a_chroma[i].im);                                               \
      }                                                                                                \
    }                                                                                                  \
    OR_FN(dx)((YY), h, w, t1);                                                                         \
    for (size_t i = 0; i < n; ++i) t1[i] *= wx[i];                                                     \
    OR_FN(dxt)(t1, h, w, t2);                                                                          \
    for (size_t i = 0; i < n; ++i) (GY)[i] += (R)2 * t2[i];                                            \
    OR_FN(dy)((YY), h, w, t1);                                                                         \
    for (size_t i = 0; i < n; ++i) t1[i] *= wy[i];                                                     \
    OR_FN(dyt)(t1, h, w, t2);                                                                          \
    for (size_t i = 0; i < n; ++i) (GY)[i] += (R)2 * t2[i];                                            \
    cdx((CC), h, w, c1);                                                                               \
    for (size_t i = 0; i < n; ++i) { c1[i].re *= wcx[i]; c1[i].im *= wcx[i]; }                          \
    cdxt(c1, h, w, c3);                                                                                \
    for (size_t i = 0; i < n; ++i) { (GC)[i].re += (R)2 * mu_h * c3[i].re; (GC)[i].im += (R)2 * mu_h * c3[i].im; } \
    cdy((CC), h, w, c2);                                                                               \
    for (size_t i = 0; i < n; ++i) { c2[i].re *= wcy[i]; c2[i].im *= wcy[i]; }                          \
    cdyt(c2, h, w, c3);                                                                                \
    for (size_t i = 0; i < n; ++i) { (GC)[i].re += (R)2 * mu_v * c3[i].re; (GC)[i].im += (R)2 * mu_v * c3[i].im; } \
    if (mu_d > 0) {                                                                                    \
      cd1((CC), h, w, c1);                                                                             \
      for (size_t i = 0; i < n; ++i) { c1[i].re *= wd1[i]; c1[i].im *= wd1[i]; }                        \
      cd1t(c1, h, w, c3);                                                                              \
      for (size_t i = 0; i < n; ++i) { (GC)[i].re += (R)2 * mu_d * c3[i].re; (GC)[i].im += (R)2 * mu_d * c3[i].im; } \
      cd2((CC), h, w, c1);                                                                             \
      for (size_t i = 0; i < n; ++i) { c1[i].re *= wd2[i]; c1[i].im *= wd2[i]; }                        \
      cd2t(c1, h, w, c3);                                                                              \
      for (size_t i = 0; i < n; ++i) { (GC)[i].re += (R)2 * mu_d * c3[i].re; (GC)[i].im += (R)2 * mu_d * c3[i].im; } \
    }                                                                                                  \
  } while (0)

  for (int outer = 0; outer < irls_outer; ++outer) {
    /* :773-858 the stage's weights, lagged on (Y, chi) */
    OR_FN(dx)(y, h, w, gxY);
    OR_FN(dy)(y, h, w, gyY);
    cdx(chi, h, w, c1);
    cdy(chi, h, w, c2);
    for (size_t i = 0; i < n; ++i) {
      wx[i] = eps / (R)sqrt((double)(gxY[i] * gxY[i] + e2));
      wy[i] = eps / (R)sqrt((double)(gyY[i] * gyY[i] + e2));
      const R mx = cabsr(c1[i]), my = cabsr(c2[i]);
      wcx[i] = eps_c / (R)sqrt((double)(mx * mx + k * gxY[i] * gxY[i] + ec2));
      wcy[i] = eps_c / (R)sqrt((double)(my * my + k * gyY[i] * gyY[i] + ec2));
    }
    if (mu_d > 0) {
      cd1(chi, h, w, c1);
      cd2(chi, h, w, c2);
      for (size_t i = 0; i < n; ++i) {
        const R a1 = cabsr(c1[i]) / eps_c, a2 = cabsr(c2[i]) / eps_c;
        wd1[i] = (R)1 / (R)sqrt((double)(1 + a1 * a1));
        wd2[i] = (R)1 / (R)sqrt((double)(1 + a2 * a2));
      }
    }
    if (nu > 0) {
      for (int j = 0; j < nn; ++j) {
        const R *swj = sw + (size_t)j * n, *cfj = conf + (size_t)j * n;
        const C *cwj = cw + (size_t)j * n;
        R *wtj = wt + (size_t)j * n;
        for (size_t i = 0; i < n; ++i) {
          const R rk = swj[i] - y[i] - re_prod(chi[i], cwj[i]);
          wtj[i] = cfj[i] * et2 / (rk * rk + et2);
        }
      }
      for (size_t i = 0; i < n; ++i) tr[i] = 1;
      for (int j = 0; j < nn; ++j)
        for (size_t i = 0; i < n; ++i) tr[i] += nu * wt[(size_t)j * n + i];
      for (size_t i = 0; i < n; ++i) {
        const R ki = tr[i];
        wx[i] *= ki; wy[i] *= ki; wcx[i] *= ki; wcy[i] *= ki;
        if (mu_d > 0) { wd1[i] *= ki; wd2[i] *= ki; }
      }
    }
    JGRAD(y, chi, gY, gC);
    for (size_t i = 0; i < n; ++i) { dY[i] = -gY[i]; dC[i].re = -gC[i].re; dC[i].im = -gC[i].im; }
    double gg = 0;
    { double a = 0, b = 0; for (size_t i = 0; i < n; ++i) a += (double)gY[i] * gY[i]; b = sum_abs2(gC, n); gg = a + b; }
    const double gg0 = gg;
    int it = 0;
    for (; it < n_inner; ++it) {
      /* :687-755 curv along (dY, dC) */
      double hh = 0;
      for (size_t i = 0; i < n; ++i) { const R v = dY[i] + re_prod(dC[i], car[i]); hh += (double)v * v; }
      for (int j = 0; j < nn; ++j) {
        const R *wtj = wt + (size_t)j * n;
        const C *cwj = cw + (size_t)j * n;
        double acc = 0;
        for (size_t i = 0; i < n; ++i) {
          const R dk = dY[i] + re_prod(dC[i], cwj[i]);
          acc += (double)wtj[i] * dk * dk;
        }
        hh += (double)nu * acc;
      }
      if (nu_a > 0) {
        double acc = 0;
        for (size_t i = 0; i < n; ++i) acc += (double)a_weight[i] * ((double)dY[i] * dY[i] + (double)cabs2(dC[i]));
        hh += (double)nu_a * acc;
      }
      OR_FN(dx)(dY, h, w, t1);
      OR_FN(dy)(dY, h, w, t2);
      cdx(dC, h, w, c4);
      cdy(dC, h, w, c5);
      double spatial = 0;
      for (size_t i = 0; i < n; ++i) {
        spatial += (double)wx[i] * t1[i] * t1[i];
        spatial += (double)wy[i] * t2[i] * t2[i];
        spatial += (double)mu_h * wcx[i] * cabs2(c4[i]);
        spatial += (double)mu_v * wcy[i] * cabs2(c5[i]);
      }
      if (mu_d > 0) {
        cd1(dC, h, w, c4);
        cd2(dC, h, w, c5);
        double ad = 0;
        for (size_t i = 0; i < n; ++i) {
          ad += (double)wd1[i] * cabs2(c4[i]);
          ad += (double)wd2[i] * cabs2(c5[i]);
        }
        spatial += mu_d * ad;
      }
      hh += spatial;
      if (hh <= 1e-12) break;
      double gd = 0;
      { double a = 0; for (size_t i = 0; i < n; ++i) a += (double)gY[i] * dY[i]; gd = a + dot_real(gC, dC, n); }
      const R alpha = (R)(-0.5 * gd / hh);
      for (size_t i = 0; i < n; ++i) {
        y[i] += alpha * dY[i];
        chi[i].re += alpha * dC[i].re;
        chi[i].im += alpha * dC[i].im;
      }
      JGRAD(y, chi, gYn, gCn);
      double ggn = 0;
      { double a = 0; for (size_t i = 0; i < n; ++i) a += (double)gYn[i] * gYn[i]; ggn = a + sum_abs2(gCn, n); }
      const R beta = (R)(ggn / (gg > 1e-30 ? gg : 1e-30));
      for (size_t i = 0; i < n; ++i) {
        dY[i] = -gYn[i] + beta * dY[i];
        dC[i].re = -gCn[i].re + beta * dC[i].re;
        dC[i].im = -gCn[i].im + beta * dC[i].im;
      }
      memcpy(gY, gYn, sizeof(R) * n);
      memcpy(gC, gCn, sizeof(C) * n);
      gg = ggn;
      if (gg < (double)cg_tol * cg_tol * gg0) { ++it; break; }
      if (gg < 1e-10 * (double)n) { ++it; break; }
    }
    if (stage_iters) stage_iters[outer] = it;
  }
#undef JGRAD
  for (int i = 0; i < 14; ++i) free(rp[i]);
  for (int i = 0; i < 8; ++i) free(cp[i]);
  free(wt);
  return aniso;
}
