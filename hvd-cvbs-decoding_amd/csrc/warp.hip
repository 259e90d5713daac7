// warp.hip -- the temporal neighbours' warps (temporal.cpp:19-287): per-pixel
// motion vectors from a tile motion field, the integer-pel warp of composite
// and carrier, the bilinear warp of baseband planes, the confidence upsample
// and the complex-coherence gate.  One thread per output value; every plane
// is read once (the warps gather, so the reads are as coalesced as the
// motion is smooth).  Float arithmetic is written in hvd-core's order with
// contraction off, so the results are hvd-core's float bits.
#include "hvd_kernels.h"
#include "warp_ops.h"

namespace hvdk {

using namespace wops;

// :44-66 outlier snap: a tile vector more than 3 px (|dy - my| + |dx - mx|)
// from its 3x3 edge-clamped median takes the median; grids under 3x3 pass.
__global__ void k_vpp_snap(int th, int tw, const float* mdy, const float* mdx, float* sy, float* sx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= th * tw) return;
  float oy, ox;
  vpp_snap(th, tw, mdy, mdx, i, oy, ox);
  sy[i] = oy;
  sx[i] = ox;
}

// :68-113 bilinear interpolation of the snapped tile vectors between tile
// centres onto the out_h x out_w grid.
__global__ void k_vpp_interp(int th, int tw, int tile, int out_h, int out_w, const float* sy, const float* sx,
                             float* dy, float* dx) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)out_h * out_w) return;
  const int y = (int)(i / out_w), x = (int)(i % out_w);
  dy[i] = vpp_interp(th, tw, tile, sy, y, x);
  if (dx) dx[i] = vpp_interp(th, tw, tile, sx, y, x);
}

// :117-132, :198-215 integer-pel warp: out(y, x) = a(clamp(y - lround(dy)),
// clamp(x - lround(dx))), `ch` (1 or 2) floats per pixel.
__global__ void k_warp_tiles(int h, int w, int ch, const float* a, const float* dy, const float* dx, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)h * w) return;
  const size_t j = warp_tiles_src(h, w, (int)(i / w), (int)(i % w), dy[i], dx[i]);
  if (ch == 2) reinterpret_cast<float2*>(out)[i] = reinterpret_cast<const float2*>(a)[j];
  else out[i] = a[j];
}

// :140-166 bilinear warp of a baseband plane (hs x ws) onto out_h x out_w,
// the source clipped to [0, hs - 1.001] x [0, ws - 1.001].
__global__ void k_warp_bilinear(int hs, int ws, const float* a, const float* dy, const float* dx, float row_offset,
                                int out_h, int out_w, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)out_h * out_w) return;
  out[i] = warp_bilinear(hs, ws, a, dy[i], dx[i], row_offset, (int)(i / out_w), (int)(i % out_w));
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
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  coh_terms(z1[i], z2[i], cre[i], cim[i], m1[i], m2[i]);
}

// :184-195 gamma = |<cross>| / (sqrt(max(0, <m1><m2>)) + 1e-6), in [0, 1].
__global__ void k_coh_out(size_t n, const float* bre, const float* bim, const float* b1, const float* b2, float* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = coh_out(bre[i], bim[i], b1[i], b2[i]);
}

}  // namespace hvdk
