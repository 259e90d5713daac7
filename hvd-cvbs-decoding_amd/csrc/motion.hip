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
