// motion.hip -- integer SAD block matching over uint16 fields.
//
// hvd-core's motion estimation (engine/motion.cpp:105-260 BmPass) matches
// float SSD on 32 px tiles; this operator is the integer form the batched
// pipeline uses: bw x bh blocks, a full +/-sy x +/-sx search, SAD in int32,
// the reference sample for candidate (dy, dx) at (y - dy, x - dx) with edge
// clamping (the BmPass convention, motion.cpp:124-131), and the winner
// chosen as the least (SAD, |dy| + |dx|, raster index) -- a total order, so
// the result is bit-exact and independent of evaluation order.
//
// One block = 4 wavefronts = 4 horizontally adjacent macroblocks sharing one
// LDS copy of the reference search window; each lane evaluates candidates
// lane, lane + 64, ... of its wavefront's macroblock and the (cost, |v|,
// index) keys are min-reduced across the wave.
#include <algorithm>

#include "hvd_device.h"
#include "hvd_kernels.h"

namespace hvdk {

constexpr int kSadMb = 4;          // macroblocks per block
constexpr int kSadMaxBw = 32, kSadMaxBh = 16, kSadMaxSx = 32, kSadMaxSy = 16;
constexpr int kSadWinW = kSadMb * kSadMaxBw + 2 * kSadMaxSx;
constexpr int kSadWinH = kSadMaxBh + 2 * kSadMaxSy;

__global__ void __launch_bounds__(256) k_sad_bm(const SadDesc* D, const int* blk_prefix, int nf) {
  __shared__ uint16_t win[kSadWinH][kSadWinW];
  __shared__ uint16_t cur[kSadMaxBh][kSadMb * kSadMaxBw];
  const int fi = field_of_tile(blk_prefix, nf, blockIdx.x);
  const SadDesc& d = D[fi];
  const int lb = blockIdx.x - blk_prefix[fi];
  const int groups_x = (d.nbx + kSadMb - 1) / kSadMb;
  const int by = lb / groups_x, bxg = (lb % groups_x) * kSadMb;
  const int y0 = by * d.bh, x0 = bxg * d.bw;
  const int ww = kSadMb * d.bw + 2 * d.sx, wh = d.bh + 2 * d.sy;
  for (int e = threadIdx.x; e < wh * ww; e += blockDim.x) {
    const int ry = e / ww, rx = e % ww;
    int y = y0 - d.sy + ry, x = x0 - d.sx + rx;
    y = y < 0 ? 0 : (y >= d.h ? d.h - 1 : y);
    x = x < 0 ? 0 : (x >= d.w ? d.w - 1 : x);
    win[ry][rx] = d.ref[(size_t)y * d.w + x];
  }
  for (int e = threadIdx.x; e < d.bh * kSadMb * d.bw; e += blockDim.x) {
    const int ry = e / (kSadMb * d.bw), rx = e % (kSadMb * d.bw);
    const int y = y0 + ry, x = x0 + rx;
    cur[ry][rx] = (y < d.h && x < d.w) ? d.cur[(size_t)y * d.w + x] : (uint16_t)0;
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bx = bxg + wv;
  if (bx >= d.nbx) return;
  const int mx0 = wv * d.bw;  // this macroblock's column offset inside cur/win
  const int yl = min(d.bh, d.h - y0), xl = min(d.bw, d.w - (x0 + mx0));
  const int ncx = 2 * d.sx + 1, ncand = ncx * (2 * d.sy + 1);
  unsigned long long best = ~0ull;
  if (xl == d.bw && (d.bw & 1) == 0) {
    // Full macroblock of even width: a lane takes four neighbouring
    // candidates (dy, dx0 .. dx0 + 3), dx0 = -sx + 4 g, which share one read
    // of their window rows as 32-bit pairs.  Their columns start at
    // ox_j = ox0 - j with ox0 = mx0 + 2 sx - 4 g even, so candidate j's sample
    // pairs are the dwords from u16 index ox0 - 4 shifted by 4 - j samples:
    // whole dwords for even j, byte-aligned across two dwords for odd j.  The
    // block's rows are read as dwords every lane of the wave shares.
    const int ngx = (ncx + 3) / 4, ntask = ngx * (2 * d.sy + 1), nw = d.bw / 2;
    for (int task = lane; task < ntask; task += 64) {
      const int iy = task / ngx, g = task % ngx;
      const int dy = iy - d.sy, dx0 = 4 * g - d.sx;
      const int oy = d.sy - dy, s = mx0 + 2 * d.sx - 4 * g - 4;  // even, >= 0 once g <= sx / 2
      unsigned acc[4] = {0u, 0u, 0u, 0u};
      for (int yy = 0; yy < yl; ++yy) {
        const uint16_t* wr = &win[yy + oy][0];
        const uint32_t* cw = reinterpret_cast<const uint32_t*>(&cur[yy][mx0]);
#pragma unroll 4
        for (int k = 0; k < nw; ++k) {
          const uint32_t cv = cw[k];
          // window dwords k .. k + 2 from u16 index s: i0 + 5 <= ox0 + bw - 1 <= ww - 1;
          // s >= -4, and an index below 0 belongs only to the candidates
          // dx > sx of the last group, which are dropped (read as 0)
          const int i0 = s + 2 * k;
          const uint32_t w0 = i0 >= 0 ? *reinterpret_cast<const uint32_t*>(wr + i0) : 0u;
          const uint32_t w1 = i0 + 2 >= 0 ? *reinterpret_cast<const uint32_t*>(wr + i0 + 2) : 0u;
          const uint32_t w2 = *reinterpret_cast<const uint32_t*>(wr + i0 + 4);
          const uint32_t p[4] = {w2, __builtin_amdgcn_alignbyte(w2, w1, 2), w1,
                                 __builtin_amdgcn_alignbyte(w1, w0, 2)};
          // v_sad_u16 on gfx950: |a[15:0] - b[15:0]| + |a[31:16] - b[31:16]|
          // (unsigned) + acc, both samples of a pair in one instruction
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[j] = __builtin_amdgcn_sad_u16(cv, p[j], acc[j]);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int dx = dx0 + j;
        if (dx > d.sx) break;
        const int c = iy * ncx + dx + d.sx;
        const unsigned m = (unsigned)(abs(dy) + abs(dx));
        const unsigned long long key =
            ((unsigned long long)acc[j] << 32) | ((unsigned long long)m << 16) | (unsigned)c;
        best = key < best ? key : best;
      }
    }
  } else
  for (int c = lane; c < ncand; c += 64) {
    const int dy = c / ncx - d.sy, dx = c % ncx - d.sx;
    // ref(y - dy, x - dx) lives at win[(y - y0) + sy - dy][(x - x0) + sx - dx]
    const int oy = d.sy - dy, ox = mx0 + d.sx - dx;
    unsigned acc = 0;
    for (int yy = 0; yy < yl; ++yy) {
      const uint16_t* wr = &win[yy + oy][ox];
      const uint16_t* cr = &cur[yy][mx0];
      for (int xx = 0; xx < xl; ++xx) acc += (unsigned)abs((int)cr[xx] - (int)wr[xx]);
    }
    const unsigned m = (unsigned)(abs(dy) + abs(dx));
    const unsigned long long key = ((unsigned long long)acc << 32) | ((unsigned long long)m << 16) | (unsigned)c;
    best = key < best ? key : best;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_xor(best, o, 64);
    best = other < best ? other : best;
  }
  if (lane == 0) {
    const int c = (int)(best & 0xFFFFu);
    const int idx = by * d.nbx + bx;
    d.vy[idx] = (int16_t)(c / ncx - d.sy);
    d.vx[idx] = (int16_t)(c % ncx - d.sx);
    d.sad[idx] = (int32_t)(best >> 32);
  }
}

}  // namespace hvdk

// ===========================================================================
// hvd-core EstimateMotion (engine/motion.cpp:397-488) on the device, for a
// batch of (ref, cur) plane pairs of one shape.  Every float is produced in
// the reference's order of operations so the vectors and confidences are
// bit-exact with it: box blur by FP64 prefix sums (:21-55), 4x decimation
// summed in (dy, dx) order (:59-78), SSD per (tile, candidate) accumulated
// sequentially in raster order without fused multiply-add (:171-231, the
// partial-distortion early-out is exact and omitted), candidates selected in
// centre-out order with a strict '<' (:233-252, :286-298).
// ===========================================================================

// hipcc contracts a * b + c into one FMA by default; hvd-core's float code
// rounds the product first, so contraction is switched off from here on.
#pragma clang fp contract(off)

namespace hvdk {

namespace {

__device__ inline const float* mplane(const MotionDesc& d, int which) {
  return which == 0 ? d.ref : d.cur;
}

// The window sums of one line (a column or a row: n samples at stride `st`
// in a, outputs at stride `ost` in o) for positions [i0, i1), the first of
// them given: with SEQ, hvd-core's two FP64 running sums (hi, the window's
// far edge, gains a[i + r]; lo, its near edge, gains a[i - r - 1]; samples
// past the ends add 0, which leaves the sums as they are) and out = (hi - lo)
// / (2r + 1); without, one window sum (hi) that gains a[i + r] and loses
// a[i - r - 1], exact by the line's certificate.  The next kBlurPre
// positions' samples are loaded while the current ones are added, so the
// chain of adds waits on no load.
constexpr int kBlurPre = 16;
template <bool SEQ>
__device__ __forceinline__ void blur_line(const float* a, float* o, int n, size_t st, size_t ost, int r, double norm,
                                          int i0, int i1, double hi, double lo) {
#pragma clang fp contract(off)
  o[(size_t)i0 * ost] = (float)((SEQ ? hi - lo : hi) * norm);
  float vi[kBlurPre], vo[kBlurPre];
  auto fetch = [&](int b, float* pi, float* po) {
#pragma unroll
    for (int u = 0; u < kBlurPre; ++u) {
      const int i = b + u, ki = i + r, ko = i - r - 1;
      pi[u] = (i < i1 && ki < n) ? a[(size_t)ki * st] : 0.0f;
      po[u] = (i < i1 && ko >= 0) ? a[(size_t)ko * st] : 0.0f;
    }
  };
  fetch(i0 + 1, vi, vo);
  for (int b = i0 + 1; b < i1; b += kBlurPre) {
    float ni[kBlurPre], no[kBlurPre];
    if (b + kBlurPre < i1) fetch(b + kBlurPre, ni, no);
#pragma unroll
    for (int u = 0; u < kBlurPre; ++u) {
      if (b + u >= i1) break;
      if (SEQ) {
        hi += (double)vi[u];
        lo += (double)vo[u];
      } else {
        hi += (double)vi[u];
        hi -= (double)vo[u];
      }
      o[(size_t)(b + u) * ost] = (float)((SEQ ? hi - lo : hi) * norm);
    }
#pragma unroll
    for (int u = 0; u < kBlurPre; ++u) {
      vi[u] = ni[u];
      vo[u] = no[u];
    }
  }
}

}  // namespace

// Box blur pass 1 (motion.cpp:21-36, BoxBlur1DRows): along y for every
// column, FP64 running sums c_hi (the window's far edge) and c_lo (its near
// edge) added in hvd-core's order, out = (c_hi - c_lo) * (1 / (2r + 1)).
//
// When no partial sum of the column can round, c_hi and c_lo are exact, so
// c_hi - c_lo at row y is the exact sum of the window around y, which any
// order of adds forms with the same bits.  That holds when every sample is a
// multiple of 2^e (e the exponent of the smallest ulp among the column's
// nonzero floats) and h max|a| < 2^(53 + e) (the certificate of the row pass
// below).  The block's waves take the row segments of 64 columns: each
// checks its segment, the column's verdict is combined through LDS, and a
// certified column's segments are formed in parallel, each from its own
// running window sum; a column that fails it (a nonzero sample below about
// 2^-20 of the column's largest, in practice) is summed top to bottom by the
// first wave, as hvd-core does.
// Block: 64 consecutive columns of one plane x (blockDim / 64) segments.
// Grid: (ceil(w / 64), planes).
constexpr int kBlurSegs = 8;
__global__ void __launch_bounds__(64 * kBlurSegs) k_me_blur_cols(const MotionDesc* D, int npairs, int h, int w, int r,
                                                                 float* tmp) {
  __shared__ float smx[kBlurSegs][64];
  __shared__ int sem[kBlurSegs][64];
  const int t = threadIdx.x & 63, seg = threadIdx.x >> 6, nseg = blockDim.x >> 6;
  const int x = blockIdx.x * 64 + t, pp = blockIdx.y;
  const bool col = x < w;
  const float* a = mplane(D[pp >> 1], pp & 1);
  float* o = tmp + (size_t)pp * h * w;
  const float norm = 1.0f / (float)(2 * r + 1);
  const int ya = (int)((long long)seg * h / nseg), yb = (int)((long long)(seg + 1) * h / nseg);
  float mx = 0.0f;
  int emin = 1 << 20;
  for (int y0 = ya; col && y0 < yb; y0 += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = y0 + u < yb ? a[(size_t)(y0 + u) * w + x] : 0.0f;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      mx = fmaxf(mx, fabsf(v[u]));
      if (v[u] != 0.0f) emin = min(emin, max((int)((__float_as_uint(v[u]) >> 23) & 255u), 1) - 150);
    }
  }
  smx[seg][t] = mx;
  sem[seg][t] = emin;
  __syncthreads();
  if (!col) return;  // no barriers below
  for (int k = 0; k < nseg; ++k) {
    mx = fmaxf(mx, smx[k][t]);
    emin = min(emin, sem[k][t]);
  }
  const bool exact = emin == (1 << 20) || (double)h * (double)mx < ldexp(1.0, 53 + emin);
  if (exact) {
    if (ya >= yb) return;
    double s = 0.0;  // the window sum of row ya (a subset sum: exact)
    for (int k = max(ya - r, 0); k <= min(ya + r, h - 1); ++k) s += (double)a[(size_t)k * w + x];
    blur_line<false>(a + x, o + x, h, w, w, r, (double)norm, ya, yb, s, 0.0);
    return;
  }
  if (seg != 0) return;
  // rows beyond the plane add 0, which leaves the sums (never -0) as they are
  double c_hi = 0.0;
  const int y1 = min(r + 1, h);
  for (int i = 0; i < y1; ++i) c_hi += (double)a[(size_t)i * w + x];
  blur_line<true>(a + x, o + x, h, w, w, r, (double)norm, 0, h, c_hi, 0.0);
}

// Box blur pass 2 (motion.cpp:38-55, BoxBlur1DCols): along x for every row,
// out[x] = (c[x1] - c[x0]) * (1 / (2r + 1)) with c the FP64 running sum of
// the row.  One wave per row.
//
// hvd-core adds the row left to right.  When no partial sum of the row can
// round, any order of the adds gives the same c, and the wave forms c as a
// parallel prefix sum.  That holds when every input is a multiple of 2^e
// (e the exponent of the smallest ulp among the row's nonzero floats) and
// w max|a| < 2^(53 + e): every partial sum, of any subset, is then an
// integer multiple of 2^e below 2^53 of them, which a double holds exactly.
// The wave checks this on the row first; a row that fails it (a nonzero
// sample below about 2^-12 of the row's largest, in practice) is summed left
// to right by lane 0, as hvd-core does, from a copy of the row in LDS.
// Grid: rows = npairs * 2 * h blocks.
__global__ void __launch_bounds__(64) k_me_blur_rows(int npairs, int h, int w, int r, const float* tmp, float* out) {
  __shared__ double pc[kBlurMaxW + 1];
  const int lane = threadIdx.x;
  const size_t row = blockIdx.x;
  if (row >= (size_t)npairs * 2 * h) return;
  const float* a = tmp + row * w;
  float* o = out + row * w;
  const float norm = 1.0f / (float)(2 * r + 1);
  float mx = 0.0f;
  int emin = 1 << 20;
  for (int x = lane; x < w; x += 64) {
    const float v = a[x];
    mx = fmaxf(mx, fabsf(v));
    if (v != 0.0f) emin = min(emin, max((int)((__float_as_uint(v) >> 23) & 255u), 1) - 150);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    emin = min(emin, __shfl_xor(emin, off, 64));
  }
  const bool exact = w <= kBlurMaxW && (emin == (1 << 20) || (double)w * (double)mx < ldexp(1.0, 53 + emin));
  if (exact) {
    // pc[k] = a[0] + ... + a[k - 1]: lanes take 8 consecutive samples, a
    // scan of the lanes' sums, a carry between chunks of 512
    double carry = 0.0;
    if (lane == 0) pc[0] = 0.0;
    for (int c0 = 0; c0 < w; c0 += 512) {
      double loc[8], sacc = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int x = c0 + 8 * lane + k;
        sacc += x < w ? (double)a[x] : 0.0;
        loc[k] = sacc;
      }
      double incl = sacc;
      for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
      }
      const double base = carry + (incl - sacc);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int x = c0 + 8 * lane + k;
        if (x < w) pc[x + 1] = base + loc[k];
      }
      carry += __shfl(incl, 63, 64);
    }
    __syncthreads();
    for (int x = lane; x < w; x += 64) {
      const int x1 = min(x + r + 1, w), x0 = max(x - r, 0);
      o[x] = (float)((pc[x1] - pc[x0]) * (double)norm);
    }
  } else if (12 * w + 8 <= 8 * (kBlurMaxW + 1)) {
    // hvd-core's prefix sums c[k + 1] = c[k] + a[k] by lane 0 (the row's
    // samples staged after them in LDS, the next kBlurPre loaded while the
    // current ones are added), then every lane forms (c[x1] - c[x0]) / (2r + 1)
    float* sa = reinterpret_cast<float*>(pc + w + 1);
    for (int x = lane; x < w; x += 64) sa[x] = a[x];
    __syncthreads();
    if (lane == 0) {
      double c = 0.0;
      pc[0] = 0.0;
      float v[kBlurPre];
#pragma unroll
      for (int u = 0; u < kBlurPre; ++u) v[u] = u < w ? sa[u] : 0.0f;
      for (int b = 0; b < w; b += kBlurPre) {
        float nv[kBlurPre];
#pragma unroll
        for (int u = 0; u < kBlurPre; ++u) nv[u] = b + kBlurPre + u < w ? sa[b + kBlurPre + u] : 0.0f;
#pragma unroll
        for (int u = 0; u < kBlurPre; ++u) {
          if (b + u < w) {
            c += (double)v[u];
            pc[b + u + 1] = c;
          }
        }
#pragma unroll
        for (int u = 0; u < kBlurPre; ++u) v[u] = nv[u];
      }
    }
    __syncthreads();
    for (int x = lane; x < w; x += 64) {
      const int x1 = min(x + r + 1, w), x0 = max(x - r, 0);
      o[x] = (float)((pc[x1] - pc[x0]) * (double)norm);
    }
  } else if (w <= kBlurMaxW) {
    float* sa = reinterpret_cast<float*>(pc);  // the row, then its result
    float* so = sa + kBlurMaxW;
    for (int x = lane; x < w; x += 64) sa[x] = a[x];
    __syncthreads();
    if (lane == 0) {
      double c_hi = 0.0;
      for (int i = 0; i < min(r + 1, w); ++i) c_hi += (double)sa[i];
      blur_line<true>(sa, so, w, 1, 1, r, (double)norm, 0, w, c_hi, 0.0);
    }
    __syncthreads();
    for (int x = lane; x < w; x += 64) o[x] = so[x];
  } else if (lane == 0) {
    double c_lo = 0.0, c_hi = 0.0;
    int i_lo = 0, i_hi = 0;
    for (int x = 0; x < w; ++x) {
      const int x0 = max(x - r, 0), x1 = min(x + r + 1, w);
      while (i_hi < x1) c_hi += (double)a[i_hi++];
      while (i_lo < x0) c_lo += (double)a[i_lo++];
      o[x] = (float)((c_hi - c_lo) * (double)norm);
    }
  }
}

void launch_me_confidence(hipStream_t st, int n, int h, int w, int tile, const float* planes, const float* best,
                          float* const* conf_out, float* scratch) {
  const size_t lds = sizeof(float) * (size_t)tile * w;
  const bool band = lds <= kConfBandMaxLds;
  if (band) {
    const int th = (h + tile - 1) / tile;
    hipLaunchKernelGGL(k_me_conf_tiles, dim3(th, n), dim3(256), lds, st, h, w, tile, planes, best, scratch);
  }
  hipLaunchKernelGGL(k_me_confidence, dim3(n), dim3(256), 0, st, h, w, tile, planes, best, conf_out, scratch,
                     band ? 1 : 0);
}

void launch_me_blur(hipStream_t st, const MotionDesc* D, int npairs, int h, int w, int r, float* tmp, float* out) {
  launch_me_blur_planes(st, D, 2 * npairs, h, w, r, tmp, out);
}

// The first `nplanes` planes of D's pairs (plane p: pair p / 2, ref then cur)
// into out (nplanes planes after one another), tmp holding as many.
void launch_me_blur_planes(hipStream_t st, const MotionDesc* D, int nplanes, int h, int w, int r, float* tmp,
                           float* out) {
  const int npairs = (nplanes + 1) / 2;
  const int nseg = std::min(kBlurSegs, std::max(1, h / 32));  // row segments of a certified column
  hipLaunchKernelGGL(k_me_blur_cols, dim3((w + 63) / 64, nplanes), dim3(64 * nseg), 0, st, D, npairs, h, w, r, tmp);
  const size_t rows = (size_t)nplanes * h;
  hipLaunchKernelGGL(k_me_blur_rows, dim3((unsigned)rows), dim3(64), 0, st, npairs, h, w, r, (const float*)tmp, out);
}

// f x f block average, remainder cropped (motion.cpp:59-78).
__global__ void k_me_decimate(int npairs, int h, int w, int f, const float* in, float* out) {
  const int hh = h / f, ww = w / f;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npairs * 2 * hh * ww) return;
  const int x = idx % ww, y = (idx / ww) % hh, pp = idx / (hh * ww);
  const float* a = in + (size_t)pp * h * w;
  const float norm = 1.0f / (float)(f * f);
  float acc = 0.0f;
  for (int dy = 0; dy < f; ++dy)
    for (int dx = 0; dx < f; ++dx) acc = (acc + a[(size_t)(y * f + dy) * w + x * f + dx]);
  out[(size_t)pp * hh * ww + (size_t)y * ww + x] = (acc * norm);
}

// SSD of one candidate for one tile (motion.cpp:171-231 without the early
// out).  planes: [pair][2][h][w] (A = cur = plane 1, B = ref = plane 0).
// base_dy/base_dx (per tile, may be null) make the candidate relative.
// se: [pair][tile][cand].  Thread per (pair, tile, cand).
__global__ void k_me_ssd(int npairs, int h, int w, int tile, const float* planes, const int* dys, const int* dxs,
                         int ndy, int ndx, const float* base_dy, const float* base_dx, float* se) {
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile;
  const int nc = ndy * ndx, nt = th * tw;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npairs * nt * nc) return;
  const int c = idx % nc, t = (idx / nc) % nt, pi = idx / (nc * nt);
  const int ty = t / tw, tx = t % tw;
  int sy = dys[c / ndx], sx = dxs[c % ndx];
  if (base_dy) {
    sy += (int)lroundf(base_dy[(size_t)pi * nt + t]);
    sx += (int)lroundf(base_dx[(size_t)pi * nt + t]);
  }
  const float* B = planes + ((size_t)pi * 2 + 0) * h * w;
  const float* A = planes + ((size_t)pi * 2 + 1) * h * w;
  const int y_end = min((ty + 1) * tile, h), x_end = min((tx + 1) * tile, w);
  float acc = 0.0f;
  for (int y = ty * tile; y < y_end; ++y) {
    const int yy = min(max(y - sy, 0), h - 1);
    for (int x = tx * tile; x < x_end; ++x) {
      const int xx = min(max(x - sx, 0), w - 1);
      const float diff = (A[(size_t)y * w + x] - B[(size_t)yy * w + xx]);
      acc = (acc + (diff * diff));
    }
  }
  se[idx] = acc;
}

// The same SSDs with one wave per (pair, tile, candidate), for full-size
// tiles: the lanes form the squared differences of kMeChunk pixels at a time
// into LDS, in raster order, and lane 0 adds them one after another, so the
// sum has the bits of the sequential loop above while the loads and products
// run in parallel.  Four waves per block, each with its own LDS row.
constexpr int kMeChunk = 1024;
__global__ void __launch_bounds__(256) k_me_ssd_wave(int npairs, int h, int w, int tile, const float* planes,
                                                     const int* dys, const int* dxs, int ndy, int ndx,
                                                     const float* base_dy, const float* base_dx, float* se) {
  __shared__ float sq[4][kMeChunk];
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile;
  const int nc = ndy * ndx, nt = th * tw;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long idx = (long long)blockIdx.x * 4 + wv;
  if (idx >= (long long)npairs * nt * nc) return;  // whole waves leave; no block barrier below
  const int c = (int)(idx % nc), t = (int)((idx / nc) % nt), pi = (int)(idx / ((long long)nc * nt));
  const int ty = t / tw, tx = t % tw;
  int sy = dys[c / ndx], sx = dxs[c % ndx];
  if (base_dy) {
    sy += (int)lroundf(base_dy[(size_t)pi * nt + t]);
    sx += (int)lroundf(base_dx[(size_t)pi * nt + t]);
  }
  const float* B = planes + ((size_t)pi * 2 + 0) * h * w;
  const float* A = planes + ((size_t)pi * 2 + 1) * h * w;
  const int y0 = ty * tile, x0 = tx * tile;
  const int y_end = min(y0 + tile, h), x_end = min(x0 + tile, w);
  const int cw = x_end - x0, rows = max(1, kMeChunk / cw);
  float* q = sq[wv];
  float acc = 0.0f;
  for (int ya = y0; ya < y_end; ya += rows) {
    const int nr = min(rows, y_end - ya), n = nr * cw;
    for (int e = lane; e < n; e += 64) {
      const int y = ya + e / cw, x = x0 + e % cw;
      const int yy = min(max(y - sy, 0), h - 1), xx = min(max(x - sx, 0), w - 1);
      const float diff = (A[(size_t)y * w + x] - B[(size_t)yy * w + xx]);
      q[e] = (diff * diff);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane == 0)
      for (int e = 0; e < n; ++e) acc = (acc + q[e]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (lane == 0) se[idx] = acc;
}

// The SSDs of full-size tiles with many candidates (the refinement pass),
// one wave per (pair, tile): the tile of A and the window of B that every
// candidate reaches (rows and columns clamped to the plane, as BmPass clamps
// them) are staged in LDS, then lane c sums candidate c's squared
// differences in the tile's raster order (hvd-core's adds, motion.cpp:
// 105-260), the candidates' chains side by side.  Candidate offsets span a
// contiguous range (RangeFromCentre): the window is the tile grown by the
// span; a window beyond kSsdWin is summed from global memory instead.
// Grid: npairs * tiles blocks of 64.
__global__ void __launch_bounds__(64) k_me_ssd_tile(int npairs, int h, int w, int tile, const float* planes,
                                                    const int* dys, const int* dxs, int ndy, int ndx,
                                                    const float* base_dy, const float* base_dx, float* se) {
  __shared__ float sa[kSsdTile * kSsdTile];
  __shared__ float sb[kSsdWin * kSsdWin];
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw, nc = ndy * ndx;
  const int t = (int)(blockIdx.x % nt), pi = (int)(blockIdx.x / nt), lane = threadIdx.x;
  if (pi >= npairs) return;
  const int ty = t / tw, tx = t % tw;
  const int y0 = ty * tile, x0 = tx * tile;
  const int ch = min(y0 + tile, h) - y0, cw = min(x0 + tile, w) - x0;
  int by = 0, bx = 0;
  if (base_dy) {
    by = (int)lroundf(base_dy[(size_t)pi * nt + t]);
    bx = (int)lroundf(base_dx[(size_t)pi * nt + t]);
  }
  int dymin = dys[0], dymax = dys[0], dxmin = dxs[0], dxmax = dxs[0];
  for (int i = 1; i < ndy; ++i) {
    dymin = min(dymin, dys[i]);
    dymax = max(dymax, dys[i]);
  }
  for (int i = 1; i < ndx; ++i) {
    dxmin = min(dxmin, dxs[i]);
    dxmax = max(dxmax, dxs[i]);
  }
  const int symax = dymax + by, sxmax = dxmax + bx;
  const int wh = ch + dymax - dymin, ww = cw + dxmax - dxmin;
  const float* B = planes + ((size_t)pi * 2 + 0) * h * w;
  const float* A = planes + ((size_t)pi * 2 + 1) * h * w;
  const bool staged = ch * cw <= kSsdTile * kSsdTile && wh <= kSsdWin && ww <= kSsdWin;
  if (staged) {
    for (int e = lane; e < ch * cw; e += 64) sa[e] = A[(size_t)(y0 + e / cw) * w + x0 + e % cw];
    for (int e = lane; e < wh * ww; e += 64) {
      const int yy = min(max(y0 - symax + e / ww, 0), h - 1), xx = min(max(x0 - sxmax + e % ww, 0), w - 1);
      sb[e] = B[(size_t)yy * w + xx];
    }
  }
  __syncthreads();
  for (int c = lane; c < nc; c += 64) {
    const int sy = dys[c / ndx] + by, sx = dxs[c % ndx] + bx;
    float acc = 0.0f;
    if (staged) {
      const int ro = symax - sy, co = sxmax - sx;
      for (int y = 0; y < ch; ++y) {
        const float* ar = sa + y * cw;
        const float* br = sb + (y + ro) * ww + co;
#pragma unroll 8
        for (int x = 0; x < cw; ++x) {
          const float diff = (ar[x] - br[x]);
          acc = (acc + (diff * diff));
        }
      }
    } else {
      for (int y = y0; y < y0 + ch; ++y)
        for (int x = x0; x < x0 + cw; ++x) {
          const int yy = min(max(y - sy, 0), h - 1), xx = min(max(x - sx, 0), w - 1);
          const float diff = (A[(size_t)y * w + x] - B[(size_t)yy * w + xx]);
          acc = (acc + (diff * diff));
        }
    }
    se[((size_t)pi * nt + t) * nc + c] = acc;
  }
}

// In-order candidate selection (motion.cpp:233-252): cost = se * bias,
// strict '<'.  Thread per (pair, tile).  cost0 (may be null) receives the
// SSD of the (0, 0) candidate.
__global__ void k_me_select(int npairs, int nt, const int* dys, const int* dxs, int ndy, int ndx, const float* se,
                            const float* base_dy, const float* base_dx, float* mdy, float* mdx, float* best,
                            float* cost0) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npairs * nt) return;
  const int nc = ndy * ndx;
  float b = __int_as_float(0x7f800000);  // +inf
  float vy = 0.0f, vx = 0.0f;
  for (int c = 0; c < nc; ++c) {
    const int dy = dys[c / ndx], dx = dxs[c % ndx];
    const float bias = (1.0f + (0.02f * (float)(abs(dy) + abs(dx))));
    const float s = se[(size_t)idx * nc + c];
    if (cost0 && dy == 0 && dx == 0) cost0[idx] = s;
    const float cost = (s * bias);
    if (cost < b) {
      b = cost;
      if (base_dy) {
        vy = (base_dy[idx] + (float)dy);
        vx = (base_dx[idx] + (float)dx);
      } else {
        vy = (float)dy;
        vx = (float)dx;
      }
    }
  }
  mdy[idx] = vy;
  mdx[idx] = vx;
  best[idx] = b;
}

// Coarse tile grid -> full-res tile grid, nearest, x 4 (motion.cpp:416-428).
__global__ void k_me_upscale(int npairs, int th, int tw, int hh, int ww, int f, const float* cdy, const float* cdx,
                             float* bdy, float* bdx) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npairs * th * tw) return;
  const int t = idx % (th * tw), pi = idx / (th * tw);
  const int ty = t / tw, tx = t % tw;
  const int gy = min((ty * hh) / th, hh - 1), gx = min((tx * ww) / tw, ww - 1);
  bdy[idx] = (cdy[(size_t)pi * hh * ww + gy * ww + gx] * (float)f);
  bdx[idx] = (cdx[(size_t)pi * hh * ww + gy * ww + gx] * (float)f);
}

// base +/- 1 in y or x for the four parabolic-fit costs (motion.cpp:309-315).
__global__ void k_me_offset(int n, const float* mdy, const float* mdx, float ody, float odx, float* tdy, float* tdx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  tdy[i] = (mdy[i] + ody);
  tdx[i] = (mdx[i] + odx);
}

// Parabolic half-pel fit, prefer-zero margin rule, vector sum
// (motion.cpp:441-485 without the confidence).  Thread per (pair, tile).
__global__ void k_me_finish(int n, const float* cym, const float* cyp, const float* cxm, const float* cxp,
                            const float* best0, const float* cost0, float* mdy, float* mdx, float* best, float* out_dy,
                            float* out_dx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto parabolic = [](float cm, float c0, float cp) {
    const float den = ((cm - (2.0f * c0)) + cp);
    float sub = 0.0f;
    if (den > 1e-9f) sub = ((0.5f * (cm - cp)) / fmaxf(den, 1e-9f));
    return fminf(fmaxf(sub, -0.5f), 0.5f);
  };
  float sdy = parabolic(cym[i], best[i], cyp[i]);
  float sdx = parabolic(cxm[i], best[i], cxp[i]);
  float vy = mdy[i], vx = mdx[i], b = best[i];
  if (best0[i] < b) { vy = 0.0f; vx = 0.0f; b = best0[i]; }
  if (b > (0.85f * cost0[i])) { vy = 0.0f; vx = 0.0f; b = cost0[i]; sdy = 0.0f; sdx = 0.0f; }
  best[i] = b;
  out_dy[i] = (vy + sdy);
  out_dx[i] = (vx + sdx);
}

// MotionConfidence's per-tile moments for one tile row of one pair: the
// band of `tile` sample rows is staged in LDS with coalesced loads, then one
// thread per tile adds its samples in the reference's raster order (the
// same adds as k_me_confidence's own loop, motion.cpp:330-350).  Grid: (tile
// rows, pairs); dynamic LDS: tile * w floats.
__global__ void __launch_bounds__(256) k_me_conf_tiles(int h, int w, int tile, const float* planes, const float* best,
                                                       float* scratch) {
  extern __shared__ float band[];
  const int ty = blockIdx.x, pi = blockIdx.y;
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw;
  const float* A = planes + ((size_t)pi * 2 + 1) * h * w;
  const int y0 = ty * tile, nr = min(y0 + tile, h) - y0;
  for (int i = threadIdx.x; i < nr * w; i += 256) band[i] = A[(size_t)y0 * w + i];
  __syncthreads();
  float* ratio = scratch + (size_t)pi * nt * 2;
  float* conf = ratio + nt;
  for (int tx = threadIdx.x; tx < tw; tx += 256) {
    const int x0 = tx * tile, x1 = min(x0 + tile, w);
    float sum = 0.0f, sum2 = 0.0f, np = 0.0f;
    for (int y = 0; y < nr; ++y) {
      const float* r = band + (size_t)y * w;
#pragma unroll 8
      for (int x = x0; x < x1; ++x) {
        const float v = r[x];
        sum = (sum + v);
        sum2 = (sum2 + (v * v));
        np = (np + 1.0f);
      }
    }
    const int t = ty * tw + tx;
    const float b = best[(size_t)pi * nt + t];
    const float npc = fmaxf(1.0f, np);
    const float norm = (1.0f / npc);
    const float mean = (sum * norm);
    const float var = ((sum2 * norm) - (mean * mean));
    const float energy = ((var + 1.0f) * npc);
    const float c = (1.0f - (b / energy));
    conf[t] = fminf(fmaxf(c, 0.0f), 1.0f);
    ratio[t] = (b / npc);
  }
}

// MotionConfidence (motion.cpp:321-387).  One block per pair; per-tile sums
// in the reference's raster order (unless k_me_conf_tiles formed them:
// have_tiles), the numpy median of best / npix, then the outlier and
// global-validity scalings.  scratch: 2 * nt floats per pair.
__global__ void __launch_bounds__(256) k_me_confidence(int h, int w, int tile, const float* planes, const float* best,
                                                       float* const* conf_out, float* scratch, int have_tiles) {
  const int pi = blockIdx.x;
  const int th = (h + tile - 1) / tile, tw = (w + tile - 1) / tile, nt = th * tw;
  const float* A = planes + ((size_t)pi * 2 + 1) * h * w;
  const float* bst = best + (size_t)pi * nt;
  float* ratio = scratch + (size_t)pi * nt * 2;
  float* conf = ratio + nt;
  __shared__ float s_lo, s_hi;
  for (int t = have_tiles ? nt : (int)threadIdx.x; t < nt; t += blockDim.x) {
    const int ty = t / tw, tx = t % tw;
    const int y0 = ty * tile, y1 = min((ty + 1) * tile, h);
    const int x0 = tx * tile, x1 = min((tx + 1) * tile, w);
    float sum = 0.0f, sum2 = 0.0f, np = 0.0f;
    if (x1 - x0 == 32 && (w & 3) == 0) {
      // full 32-sample tile rows as eight 16-byte loads, the next row's in
      // flight while this one is added (the same adds in the same order)
      const float4* row = reinterpret_cast<const float4*>(A + (size_t)y0 * w + x0);
      const size_t pitch = (size_t)w / 4;
      float4 cur[8], nxt[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) cur[k] = row[k];
      for (int y = y0; y < y1; ++y) {
        if (y + 1 < y1) {
#pragma unroll
          for (int k = 0; k < 8; ++k) nxt[k] = row[(size_t)(y + 1 - y0) * pitch + k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float q[4] = {cur[k].x, cur[k].y, cur[k].z, cur[k].w};
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            sum = (sum + q[c]);
            sum2 = (sum2 + (q[c] * q[c]));
            np = (np + 1.0f);
          }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) cur[k] = nxt[k];
      }
    } else {
      for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
          const float v = A[(size_t)y * w + x];
          sum = (sum + v);
          sum2 = (sum2 + (v * v));
          np = (np + 1.0f);
        }
    }
    const float npc = fmaxf(1.0f, np);
    const float norm = (1.0f / npc);
    const float mean = (sum * norm);
    const float var = ((sum2 * norm) - (mean * mean));
    const float energy = ((var + 1.0f) * npc);
    const float c = (1.0f - (bst[t] / energy));
    conf[t] = fminf(fmaxf(c, 0.0f), 1.0f);
    ratio[t] = (bst[t] / npc);
  }
  __syncthreads();
  // numpy median: the element of rank nt/2, averaged with rank nt/2 - 1 when nt is even
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    const float vi = ratio[t];
    int rank = 0;
    for (int j = 0; j < nt; ++j) {
      const float vj = ratio[j];
      rank += (vj < vi) || (vj == vi && j < t);
    }
    if (rank == nt / 2) s_hi = vi;
    if (nt % 2 == 0 && rank == nt / 2 - 1) s_lo = vi;
  }
  __syncthreads();
  const float med = nt % 2 == 0 ? (0.5f * (s_hi + s_lo)) : s_hi;
  float* out = conf_out[pi];
  for (int t = threadIdx.x; t < nt; t += blockDim.x) {
    const float excess = fmaxf(0.0f, (ratio[t] - med));
    const float m1 = (med + 1.0f);
    float c = (conf[t] * (m1 / (m1 + excess)));
    c = (c * (6.0f / (6.0f + med)));
    out[t] = c;
  }
}

}  // namespace hvdk
