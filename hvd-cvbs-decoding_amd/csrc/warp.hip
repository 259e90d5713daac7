// warp.hip -- the temporal neighbours' warps (temporal.cpp:19-287): per-pixel
// motion vectors from a tile motion field, the integer-pel warp of composite
// and carrier, the bilinear warp of baseband planes, the confidence upsample
// and the complex-coherence gate.  One thread per output value; every plane
// is read once (the warps gather, so the reads are as coalesced as the
// motion is smooth).  Float arithmetic is written in hvd-core's order with
// contraction off, so the results are hvd-core's float bits.
#include "hvd_kernels.h"

namespace hvdk {

namespace {

__device__ inline float tile_at(const float* v, int tw, int y, int x) { return v[(size_t)y * tw + x]; }

// :73-87 the tile-centre bracket of a pixel coordinate.
__device__ inline void tile_bracket(int coord, int n_tiles, int tile, int& lo, float& frac) {
#pragma clang fp contract(off)
  if (n_tiles > 1) {
    const float rel = (float)coord / (float)tile - 0.5f;
    lo = min(max((int)floorf(rel), 0), n_tiles - 2);
    const float centre = ((float)lo + 0.5f) * (float)tile;
    frac = fminf(fmaxf(((float)coord - centre) / (float)tile, 0.f), 1.f);
  } else {
    lo = 0;
    frac = 0.f;
  }
}

// Median of nine by a fixed exchange network (Paeth's 19 compare-swaps).
__device__ inline float median9(float* p) {
#define HVD_SW(a, b)                 \
  {                                  \
    const float lo_ = fminf(p[a], p[b]); \
    p[b] = fmaxf(p[a], p[b]);        \
    p[a] = lo_;                      \
  }
  HVD_SW(1, 2) HVD_SW(4, 5) HVD_SW(7, 8) HVD_SW(0, 1) HVD_SW(3, 4) HVD_SW(6, 7) HVD_SW(1, 2) HVD_SW(4, 5)
  HVD_SW(7, 8) HVD_SW(0, 3) HVD_SW(5, 8) HVD_SW(4, 7) HVD_SW(3, 6) HVD_SW(1, 4) HVD_SW(2, 5) HVD_SW(4, 7)
  HVD_SW(4, 2) HVD_SW(6, 4) HVD_SW(4, 2)
#undef HVD_SW
  return p[4];
}

}  // namespace

// :44-66 outlier snap: a tile vector more than 3 px (|dy - my| + |dx - mx|)
// from its 3x3 edge-clamped median takes the median; grids under 3x3 pass.
__global__ void k_vpp_snap(int th, int tw, const float* mdy, const float* mdx, float* sy, float* sx) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= th * tw) return;
  const int y = i / tw, x = i % tw;
  const float vy = mdy[i], vx = mdx[i];
  if (th < 3 || tw < 3) {
    sy[i] = vy;
    sx[i] = vx;
    return;
  }
  float a[9], b[9];
  int k = 0;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int yy = min(max(y + dy, 0), th - 1), xx = min(max(x + dx, 0), tw - 1);
      a[k] = tile_at(mdy, tw, yy, xx);
      b[k] = tile_at(mdx, tw, yy, xx);
      ++k;
    }
  const float my = median9(a), mx = median9(b);
  const bool snap = fabsf(vy - my) + fabsf(vx - mx) > 3.0f;
  sy[i] = snap ? my : vy;
  sx[i] = snap ? mx : vx;
}

// :68-113 bilinear interpolation of the snapped tile vectors between tile
// centres onto the out_h x out_w grid.
__global__ void k_vpp_interp(int th, int tw, int tile, int out_h, int out_w, const float* sy, const float* sx,
                             float* dy, float* dx) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)out_h * out_w) return;
  const int y = (int)(i / out_w), x = (int)(i % out_w);
  int iy, ix;
  float fy, fx;
  tile_bracket(y, th, tile, iy, fy);
  tile_bracket(x, tw, tile, ix, fx);
  const int iy1 = min(iy + 1, th - 1), ix1 = min(ix + 1, tw - 1);
  const float* v[2] = {sy, sx};
  float r[2];
  for (int c = 0; c < 2; ++c) {
    const float v00 = tile_at(v[c], tw, iy, ix), v10 = tile_at(v[c], tw, iy1, ix);
    const float v01 = tile_at(v[c], tw, iy, ix1), v11 = tile_at(v[c], tw, iy1, ix1);
    r[c] = v00 * (1 - fy) * (1 - fx) + v10 * fy * (1 - fx) + v01 * (1 - fy) * fx + v11 * fy * fx;
  }
  dy[i] = r[0];
  if (dx) dx[i] = r[1];
}

// :117-132, :198-215 integer-pel warp: out(y, x) = a(clamp(y - lround(dy)),
// clamp(x - lround(dx))), `ch` (1 or 2) floats per pixel.
__global__ void k_warp_tiles(int h, int w, int ch, const float* a, const float* dy, const float* dx, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)h * w) return;
  const int y = (int)(i / w), x = (int)(i % w);
  const int sy = min(max(y - (int)lroundf(dy[i]), 0), h - 1);
  const int sx = min(max(x - (int)lroundf(dx[i]), 0), w - 1);
  const size_t j = (size_t)sy * w + sx;
  if (ch == 2) reinterpret_cast<float2*>(out)[i] = reinterpret_cast<const float2*>(a)[j];
  else out[i] = a[j];
}

// :140-166 bilinear warp of a baseband plane (hs x ws) onto out_h x out_w,
// the source clipped to [0, hs - 1.001] x [0, ws - 1.001].
__global__ void k_warp_bilinear(int hs, int ws, const float* a, const float* dy, const float* dx, float row_offset,
                                int out_h, int out_w, float* out) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)out_h * out_w) return;
  const int y = (int)(i / out_w), x = (int)(i % out_w);
  float sy = (float)y - dy[i] + row_offset, sx = (float)x - dx[i];
  const float hy = (float)hs - 1.001f, hx = (float)ws - 1.001f;
  sy = sy < 0.f ? 0.f : (hy < sy ? hy : sy);
  sx = sx < 0.f ? 0.f : (hx < sx ? hx : sx);
  const int y0 = (int)sy, x0 = (int)sx;
  const float fy = sy - (float)y0, fx = sx - (float)x0;
  const int y1 = min(y0 + 1, hs - 1), x1 = min(x0 + 1, ws - 1);
  out[i] = a[(size_t)y0 * ws + x0] * (1 - fy) * (1 - fx) + a[(size_t)y1 * ws + x0] * fy * (1 - fx) +
           a[(size_t)y0 * ws + x1] * (1 - fy) * fx + a[(size_t)y1 * ws + x1] * fy * fx;
}

// :217-233 nearest-tile confidence, squared (the radius-8 blur follows).
__global__ void k_conf_nearest(int th, int tw, int tile, int out_h, int out_w, const float* conf, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)out_h * out_w) return;
  const int y = (int)(i / out_w), x = (int)(i % out_w);
  const float c = tile_at(conf, tw, min(y / tile, th - 1), min(x / tile, tw - 1));
  out[i] = c * c;
}

// :235-249 the fast form's squared tile map.
__global__ void k_square(int n, const float* a, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * a[i];
}

// :169-182 the four planes ComplexCoherence averages: z1 conj(z2) and
// |z1|^2, |z2|^2.
__global__ void k_coh_terms(size_t n, const float2* z1, const float2* z2, float* cre, float* cim, float* m1, float* m2) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 a = z1[i], b = z2[i];
  cre[i] = a.x * b.x + a.y * b.y;
  cim[i] = a.y * b.x - a.x * b.y;
  m1[i] = a.x * a.x + a.y * a.y;
  m2[i] = b.x * b.x + b.y * b.y;
}

// :184-195 gamma = |<cross>| / (sqrt(max(0, <m1><m2>)) + 1e-6), in [0, 1].
__global__ void k_coh_out(size_t n, const float* bre, const float* bim, const float* b1, const float* b2, float* out) {
#pragma clang fp contract(off)
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float num = sqrtf(bre[i] * bre[i] + bim[i] * bim[i]);
  const float den = sqrtf(fmaxf(0.0f, b1[i] * b2[i])) + 1e-6f;
  out[i] = fminf(fmaxf(num / den, 0.0f), 1.0f);
}

}  // namespace hvdk
