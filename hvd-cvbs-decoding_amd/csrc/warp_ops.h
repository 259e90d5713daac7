// warp_ops.h -- per-element arithmetic of the temporal neighbours' warps
// (hvd-core engine/temporal.cpp:19-287), shared by the single-plane kernels
// (warp.hip) and the window driver's batched ones (window.hip), so both give
// the same float bits.  Contraction is off in every function: the order of
// operations is hvd-core's.
#pragma once

#include <hip/hip_runtime.h>

namespace hvdk {
namespace wops {

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
#define HVD_SW(a, b)                     \
  {                                      \
    const float lo_ = fminf(p[a], p[b]); \
    p[b] = fmaxf(p[a], p[b]);            \
    p[a] = lo_;                          \
  }
  HVD_SW(1, 2) HVD_SW(4, 5) HVD_SW(7, 8) HVD_SW(0, 1) HVD_SW(3, 4) HVD_SW(6, 7) HVD_SW(1, 2) HVD_SW(4, 5)
  HVD_SW(7, 8) HVD_SW(0, 3) HVD_SW(5, 8) HVD_SW(4, 7) HVD_SW(3, 6) HVD_SW(1, 4) HVD_SW(2, 5) HVD_SW(4, 7)
  HVD_SW(4, 2) HVD_SW(6, 4) HVD_SW(4, 2)
#undef HVD_SW
  return p[4];
}

// :44-66 outlier snap of tile i: a vector more than 3 px (|dy - my| + |dx -
// mx|) from its 3x3 edge-clamped median takes the median; grids under 3x3
// pass.
__device__ inline void vpp_snap(int th, int tw, const float* mdy, const float* mdx, int i, float& oy, float& ox) {
#pragma clang fp contract(off)
  const int y = i / tw, x = i % tw;
  const float vy = mdy[i], vx = mdx[i];
  if (th < 3 || tw < 3) {
    oy = vy;
    ox = vx;
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
  oy = snap ? my : vy;
  ox = snap ? mx : vx;
}

// :68-113 bilinear interpolation of one snapped tile map between tile
// centres at pixel (y, x).
__device__ inline float vpp_interp(int th, int tw, int tile, const float* sv, int y, int x) {
#pragma clang fp contract(off)
  int iy, ix;
  float fy, fx;
  tile_bracket(y, th, tile, iy, fy);
  tile_bracket(x, tw, tile, ix, fx);
  const int iy1 = min(iy + 1, th - 1), ix1 = min(ix + 1, tw - 1);
  const float v00 = tile_at(sv, tw, iy, ix), v10 = tile_at(sv, tw, iy1, ix);
  const float v01 = tile_at(sv, tw, iy, ix1), v11 = tile_at(sv, tw, iy1, ix1);
  return v00 * (1 - fy) * (1 - fx) + v10 * fy * (1 - fx) + v01 * (1 - fy) * fx + v11 * fy * fx;
}

// :117-132 the integer-pel warp's source index of pixel (y, x).
__device__ inline size_t warp_tiles_src(int h, int w, int y, int x, float dy, float dx) {
  const int sy = min(max(y - (int)lroundf(dy), 0), h - 1);
  const int sx = min(max(x - (int)lroundf(dx), 0), w - 1);
  return (size_t)sy * w + sx;
}

// :140-166 the bilinear warp of a baseband plane (hs x ws) at output pixel
// (y, x), the source clipped to [0, hs - 1.001] x [0, ws - 1.001].
__device__ inline float warp_bilinear(int hs, int ws, const float* a, float dy, float dx, float row_offset, int y,
                                      int x) {
#pragma clang fp contract(off)
  float sy = (float)y - dy + row_offset, sx = (float)x - dx;
  const float hy = (float)hs - 1.001f, hx = (float)ws - 1.001f;
  sy = sy < 0.f ? 0.f : (hy < sy ? hy : sy);
  sx = sx < 0.f ? 0.f : (hx < sx ? hx : sx);
  const int y0 = (int)sy, x0 = (int)sx;
  const float fy = sy - (float)y0, fx = sx - (float)x0;
  const int y1 = min(y0 + 1, hs - 1), x1 = min(x0 + 1, ws - 1);
  return a[(size_t)y0 * ws + x0] * (1 - fy) * (1 - fx) + a[(size_t)y1 * ws + x0] * fy * (1 - fx) +
         a[(size_t)y0 * ws + x1] * (1 - fy) * fx + a[(size_t)y1 * ws + x1] * fy * fx;
}

// The same on an interleaved complex plane, one component (c = 0 real, 1
// imaginary).
__device__ inline float warp_bilinear_c(int hs, int ws, const float2* a, int c, float dy, float dx, float row_offset,
                                        int y, int x) {
#pragma clang fp contract(off)
  float sy = (float)y - dy + row_offset, sx = (float)x - dx;
  const float hy = (float)hs - 1.001f, hx = (float)ws - 1.001f;
  sy = sy < 0.f ? 0.f : (hy < sy ? hy : sy);
  sx = sx < 0.f ? 0.f : (hx < sx ? hx : sx);
  const int y0 = (int)sy, x0 = (int)sx;
  const float fy = sy - (float)y0, fx = sx - (float)x0;
  const int y1 = min(y0 + 1, hs - 1), x1 = min(x0 + 1, ws - 1);
  const float* f = reinterpret_cast<const float*>(a) + c;
  return f[2 * ((size_t)y0 * ws + x0)] * (1 - fy) * (1 - fx) + f[2 * ((size_t)y1 * ws + x0)] * fy * (1 - fx) +
         f[2 * ((size_t)y0 * ws + x1)] * (1 - fy) * fx + f[2 * ((size_t)y1 * ws + x1)] * fy * fx;
}

// :169-182 the four planes ComplexCoherence averages: z1 conj(z2), |z1|^2, |z2|^2.
__device__ inline void coh_terms(float2 a, float2 b, float& cre, float& cim, float& m1, float& m2) {
#pragma clang fp contract(off)
  cre = a.x * b.x + a.y * b.y;
  cim = a.y * b.x - a.x * b.y;
  m1 = a.x * a.x + a.y * a.y;
  m2 = b.x * b.x + b.y * b.y;
}

// :184-195 gamma = |<cross>| / (sqrt(max(0, <m1><m2>)) + 1e-6), in [0, 1].
__device__ inline float coh_out(float bre, float bim, float b1, float b2) {
#pragma clang fp contract(off)
  const float num = sqrtf(bre * bre + bim * bim);
  const float den = sqrtf(fmaxf(0.0f, b1 * b2)) + 1e-6f;
  return fminf(fmaxf(num / den, 0.0f), 1.0f);
}

}  // namespace wops
}  // namespace hvdk
