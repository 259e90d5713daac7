// selective.hip -- hvd-core's selective 3-D window (frame_bridge.cpp:385-720):
// the window decoded in 2-D, the boxes of the first field's most Y/C-
// ambiguous tiles (AmbiguousBoxes, :399-621), the window cropped to each
// box and decoded again with the full 3-D machinery, and each crop's result
// feather-blended into the 2-D fields (DecodeWindowSelective, :638-720).
//
// The tile scores are formed on the device in hvd-core's order of
// operations (count-normalised box means summed in window order in float,
// tile energies summed in raster order in double, |chi| as glibc's hypotf:
// the double square root of the double sum of squares, rounded to float),
// so they are hvd-core's bits; the decisions on the few hundred tile scores
// (median, threshold, density filter, bands, runs, area cap) are taken on
// the host.  The crops' inits are the library's holographic init and their
// decodes the window driver itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "hvd_gpu.h"
#include "window.h"

namespace hvdk {

namespace {

constexpr int kSelTile = 16;      // tile side (field rows and samples)
constexpr int kSelHalf = kSelTile / 2;  // box-mean window length
constexpr int kSelHaloX = 48;     // crop margin across
constexpr int kSelHaloY = 16;     // crop margin down (covers the motion search)
constexpr int kSelMaxBands = 4;
constexpr int kSelMaxFields = 64;

__device__ inline float sel_abs(float2 z) {
  const double re = z.x, im = z.y;
  return (float)sqrt(re * re + im * im);
}

// box2d's pass along y (frame_bridge.cpp:415-429): window d in [-4, 4),
// edge-truncated, float sum in window order divided by the count; for the
// luma and for |chi| of one pixel.
__global__ void __launch_bounds__(256) k_sel_vmean(int h, int w, const float* Y, const float2* chi, float* vY,
                                                   float* vC) {
#pragma clang fp contract(off)
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= w) return;
  float aY = 0.f, aC = 0.f;
  int n = 0;
  for (int d = -kSelHalf / 2; d < kSelHalf - kSelHalf / 2; ++d) {
    const int yy = y + d;
    if (yy < 0 || yy >= h) continue;
    const size_t i = (size_t)yy * w + x;
    aY += Y[i];
    aC += sel_abs(chi[i]);
    ++n;
  }
  const float c = (float)max(n, 1);
  vY[(size_t)y * w + x] = aY / c;
  vC[(size_t)y * w + x] = aC / c;
}

// One block per tile: the pass along x (:430-443) at each of the tile's
// pixels, the squared deviations of Y and |chi| from their box means in
// double (:465-478), and thread 0 adds them in the tile's raster order.
__global__ void __launch_bounds__(kSelTile* kSelTile) k_sel_tiles(int h, int w, const float* Y, const float2* chi,
                                                                 const float* vY, const float* vC, double* texY,
                                                                 double* texC, int* cnt) {
#pragma clang fp contract(off)
  __shared__ double sY[kSelTile * kSelTile], sC[kSelTile * kSelTile];
  __shared__ int ok[kSelTile * kSelTile];
  const int r = threadIdx.x / kSelTile, c = threadIdx.x % kSelTile;
  const int y = blockIdx.y * kSelTile + r, x = blockIdx.x * kSelTile + c;
  const bool in = y < h && x < w;
  ok[threadIdx.x] = in ? 1 : 0;
  if (in) {
    float aY = 0.f, aC = 0.f;
    int n = 0;
    const size_t row = (size_t)y * w;
    for (int d = -kSelHalf / 2; d < kSelHalf - kSelHalf / 2; ++d) {
      const int xx = x + d;
      if (xx < 0 || xx >= w) continue;
      aY += vY[row + xx];
      aC += vC[row + xx];
      ++n;
    }
    const float cn = (float)max(n, 1);
    const double dy = (double)(Y[row + x] - aY / cn);
    const double dc = (double)(sel_abs(chi[row + x]) - aC / cn);
    sY[threadIdx.x] = dy * dy;
    sC[threadIdx.x] = dc * dc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    int n = 0;
    for (int i = 0; i < kSelTile * kSelTile; ++i)
      if (ok[i]) {
        a += sY[i];
        b += sC[i];
        ++n;
      }
    const int t = blockIdx.y * gridDim.x + blockIdx.x;
    texY[t] = a;
    texC[t] = b;
    cnt[t] = n;
  }
}

// CropObs (:623-636) of every field of the window onto contiguous planes.
struct SelCrop {
  const float* s[kSelMaxFields];
  const float2* car[kSelMaxFields];
};

__global__ void __launch_bounds__(256) k_sel_crop(SelCrop src, int w, int fy0, int x0, int bh, int bw, float* cs,
                                                  float2* ccar) {
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, j = blockIdx.z;
  if (x >= bw) return;
  const size_t si = (size_t)(fy0 + y) * w + x0 + x, di = ((size_t)j * bh + y) * bw + x;
  cs[di] = src.s[j][si];
  ccar[di] = src.car[j][si];
}

// The feather blend of one box (:693-718) into every field: weights ramp
// linearly over min(16, bh / 3) rows and min(48, bw / 3) columns from the
// box's edges; Y = Y (1 - m) + Y3 m, chi likewise per component.
struct SelBlend {
  float* y[kSelMaxFields];
  float2* chi[kSelMaxFields];
};

__global__ void __launch_bounds__(256) k_sel_blend(SelBlend dst, int w, int fy0, int x0, int bh, int bw,
                                                   const float* y3, const float2* c3) {
#pragma clang fp contract(off)
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, j = blockIdx.z;
  if (x >= bw) return;
  const int hy = min(16, bh / 3), hx = min(48, bw / 3);
  float ry = 1.f, rx = 1.f;
  if (y < hy)
    ry = (float)y / (float)hy;
  else if (y >= bh - hy)
    ry = (float)(bh - 1 - y) / (float)hy;
  if (x < hx)
    rx = (float)x / (float)hx;
  else if (x >= bw - hx)
    rx = (float)(bw - 1 - x) / (float)hx;
  const float m = ry * rx, om = 1.0f - m;
  const size_t di = (size_t)(fy0 + y) * w + x0 + x, si = ((size_t)j * bh + y) * bw + x;
  float* Y = dst.y[j];
  float2* Cd = dst.chi[j];
  const float2 a = Cd[di], b = c3[si];
  Y[di] = Y[di] * om + y3[si] * m;
  Cd[di] = make_float2(a.x * om + b.x * m, a.y * om + b.y * m);
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

}  // namespace hvdk

namespace hvd_window {

using namespace hvdk;

size_t sel_scratch_bytes(int h, int w) {
  const size_t np = (size_t)h * w, nt = (size_t)((h + kSelTile - 1) / kSelTile) * ((w + kSelTile - 1) / kSelTile);
  return align256(sizeof(float) * 2 * np) + align256(sizeof(double) * 2 * nt) + align256(sizeof(int) * nt);
}

int ambiguous_boxes(hvd_gpu_ctx* ctx, const float* y, const float* chi, int h, int w, float max_area, char* scratch,
                    hipStream_t st, std::vector<SelBox>* out, bool* everywhere, std::vector<float>* scores_out) {
  out->clear();
  *everywhere = false;
  if (!y || !chi || h < 1 || w < 1) return fail(ctx, HVD_GPU_ERR_INVALID, "ambiguous boxes: bad plane");
  const size_t np = (size_t)h * w;
  const int th = (h + kSelTile - 1) / kSelTile, tw = (w + kSelTile - 1) / kSelTile, nt = th * tw;
  float* vY = (float*)scratch;
  float* vC = vY + np;
  double* texY = (double*)(scratch + align256(sizeof(float) * 2 * np));
  double* texC = texY + nt;
  int* cnt = (int*)((char*)texY + align256(sizeof(double) * 2 * nt));
  hipLaunchKernelGGL(k_sel_vmean, dim3((w + 255) / 256, h), dim3(256), 0, st, h, w, y, (const float2*)chi, vY, vC);
  hipLaunchKernelGGL(k_sel_tiles, dim3(tw, th), dim3(kSelTile * kSelTile), 0, st, h, w, y, (const float2*)chi,
                     (const float*)vY, (const float*)vC, texY, texC, cnt);
  if (hipGetLastError() != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "ambiguous boxes: launch failed");
  std::vector<double> hY(nt), hC(nt);
  std::vector<int> hn(nt);
  if (hipMemcpyAsync(hY.data(), texY, sizeof(double) * nt, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hC.data(), texC, sizeof(double) * nt, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hn.data(), cnt, sizeof(int) * nt, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(ctx, HVD_GPU_ERR_HIP, "ambiguous boxes: readback failed");
  // the tile scores (:472-477)
  std::vector<float> score(nt);
  for (int t = 0; t < nt; ++t) {
    const double n = (double)std::max(hn[t], 1);
    score[t] = (float)std::sqrt((hY[t] / n) * (hC[t] / n));
  }
  if (scores_out) *scores_out = score;
  // threshold: 4 x the median tile (:479-483)
  std::vector<float> sorted = score;
  const size_t mid = sorted.size() / 2;
  std::nth_element(sorted.begin(), sorted.begin() + mid, sorted.end());
  const float med = sorted[mid];
  const float thr = 4.0F * med;
  std::vector<char> flag(nt), dense(nt, 0);
  for (int t = 0; t < nt; ++t) flag[t] = score[t] >= thr;
  // density: at least 2 flagged of the 8 neighbours, outer ring excluded (:485-507)
  for (int ty = 1; ty + 1 < th; ++ty)
    for (int tx = 1; tx + 1 < tw; ++tx) {
      if (!flag[ty * tw + tx]) continue;
      int nb = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
          if (dy || dx) nb += flag[(ty + dy) * tw + tx + dx];
      dense[ty * tw + tx] = nb >= 2;
    }
  // bands of adjacent tile rows holding >= 2 dense tiles (:508-531)
  std::vector<std::pair<int, int>> bands;
  for (int ty = 0; ty < th; ++ty) {
    int k = 0;
    for (int tx = 0; tx < tw; ++tx) k += dense[ty * tw + tx];
    if (k < 2) continue;
    if (!bands.empty() && ty - bands.back().second <= 1)
      bands.back().second = ty;
    else
      bands.push_back({ty, ty});
  }
  if (bands.empty()) {
    // no localized peaks: a globally ambiguous field (median above 3 IRE^2)
    // goes to full 3-D, a clean one stays 2-D (:532-543)
    *everywhere = med > 3.0F;
    return HVD_GPU_OK;
  }
  // at most 4 bands: merge across the smallest gap, first on ties (:544-555)
  while ((int)bands.size() > kSelMaxBands) {
    size_t bi = 0;
    int best = 0;
    for (size_t i = 0; i + 1 < bands.size(); ++i) {
      const int gap = bands[i + 1].first - bands[i].second;
      if (i == 0 || gap < best) {
        best = gap;
        bi = i;
      }
    }
    bands[bi].second = bands[bi + 1].second;
    bands.erase(bands.begin() + bi + 1);
  }
  // one box per run of flagged columns (gaps <= 2 tiles) of each band,
  // single-tile runs dropped, runs whose margins overlap merged (:556-610)
  for (const auto& [tr0, tr1] : bands) {
    std::vector<char> colf(tw, 0);
    for (int ty = tr0; ty <= tr1; ++ty)
      for (int tx = 0; tx < tw; ++tx) colf[tx] |= dense[ty * tw + tx];
    const int y0raw = std::max(0, tr0 * kSelTile - kSelHaloY);
    const int fy0 = y0raw - y0raw % 2, fy1 = std::min(h, (tr1 + 1) * kSelTile + kSelHaloY);
    const size_t first = out->size();
    int c0 = -1, c1 = -1;
    auto close_run = [&]() {
      if (c0 < 0 || c1 - c0 < 1) return;
      const int x0 = std::max(0, c0 * kSelTile - kSelHaloX), x1 = std::min(w, (c1 + 1) * kSelTile + kSelHaloX);
      if (out->size() > first && x0 <= out->back().x1)
        out->back().x1 = x1;
      else
        out->push_back(SelBox{fy0, fy1, x0, x1});
    };
    for (int tx = 0; tx < tw; ++tx) {
      if (!colf[tx]) continue;
      if (c0 >= 0 && tx - c1 <= 2) {
        c1 = tx;
        continue;
      }
      close_run();
      c0 = c1 = tx;
    }
    close_run();
  }
  // the area cap: boxes larger in total than max_area of the field mean
  // ambiguity everywhere (:611-619)
  double total = 0.0;
  for (const SelBox& b : *out) total += (double)(b.fy1 - b.fy0) * (b.x1 - b.x0);
  if (total > (double)max_area * h * w) {
    out->clear();
    *everywhere = true;
  }
  return HVD_GPU_OK;
}

int drive_selective(hvd_gpu_ctx* ctx, const hvd_gpu_plane* F, int nf, const int* par, int is_pal,
                    const int* y_given, const hvd_solver_params* p, const hvd_window_params* wp,
                    hvd_window_diag* diag, hipStream_t st) {
  if (!F || !par || !p || !wp || nf < 1) return fail(ctx, HVD_GPU_ERR_INVALID, "null argument or empty window");
  if (nf > kSelMaxFields) return fail(ctx, HVD_GPU_ERR_INVALID, "selective 3-D takes at most 64 fields");
  const int h = F[0].lines, w = F[0].width;
  for (int j = 0; j < nf; ++j)
    if (F[j].lines != h || F[j].width != w || !F[j].s || !F[j].carrier || !F[j].chi || !F[j].y)
      return fail(ctx, HVD_GPU_ERR_INVALID, "all fields of a window share one shape, with s, carrier, chi and y");
  const size_t np = (size_t)h * w;
  // device memory: the inits (for the full 3-D fallback), the score
  // scratch, and crops of every field up to the whole field
  const size_t o_chi0 = 0, o_y0 = o_chi0 + align256(sizeof(float2) * nf * np);
  const size_t o_scr = o_y0 + align256(sizeof(float) * nf * np);
  const size_t o_crop = o_scr + align256(sel_scratch_bytes(h, w));
  const size_t crop_bytes = align256(sizeof(float) * nf * np) * 2 + align256(sizeof(float2) * nf * np) * 2;
  char* base = sel_pool(ctx, o_crop + crop_bytes, st);
  if (!base) return HVD_GPU_ERR_NOMEM;
  float2* chi0 = (float2*)(base + o_chi0);
  float* y0 = (float*)(base + o_y0);
  for (int j = 0; j < nf; ++j) {
    if (hipMemcpyAsync(chi0 + (size_t)j * np, F[j].chi, sizeof(float2) * np, hipMemcpyDeviceToDevice, st) !=
            hipSuccess ||
        (y_given && y_given[j] &&
         hipMemcpyAsync(y0 + (size_t)j * np, F[j].y, sizeof(float) * np, hipMemcpyDeviceToDevice, st) != hipSuccess))
      return fail(ctx, HVD_GPU_ERR_HIP, "selective 3-D: init copy failed");
  }
  // (1) the 2-D window (:645-651): strength < 0 turns the 3-D machinery off
  hvd_window_params c2 = *wp;
  c2.temporal_strength = -1.0;
  c2.selective_3d = 0;
  int rc = drive(ctx, F, nf, par, is_pal, y_given, p, &c2, diag, st);
  if (rc != HVD_GPU_OK) return rc;
  // (2) the boxes of the first field
  std::vector<SelBox> boxes;
  bool everywhere = false;
  rc = ambiguous_boxes(ctx, F[0].y, F[0].chi, h, w, (float)wp->selective_max_area, base + o_scr, st, &boxes,
                       &everywhere, nullptr);
  if (rc != HVD_GPU_OK) return rc;
  if (diag) diag->selective_boxes = everywhere ? -1 : (int)boxes.size();
  hvd_window_params c3 = *wp;
  c3.selective_3d = 0;
  if (everywhere) {
    // (3a) ambiguity everywhere: the whole window in 3-D from its inits
    for (int j = 0; j < nf; ++j) {
      if (hipMemcpyAsync(F[j].chi, chi0 + (size_t)j * np, sizeof(float2) * np, hipMemcpyDeviceToDevice, st) !=
              hipSuccess ||
          (y_given && y_given[j] &&
           hipMemcpyAsync(F[j].y, y0 + (size_t)j * np, sizeof(float) * np, hipMemcpyDeviceToDevice, st) !=
               hipSuccess))
        return fail(ctx, HVD_GPU_ERR_HIP, "selective 3-D: init copy failed");
    }
    return drive(ctx, F, nf, par, is_pal, y_given, p, &c3, nullptr, st);
  }
  // (3b) each box: crop, init, 3-D decode, blend (:652-718)
  float* cs = (float*)(base + o_crop);
  float* cy = (float*)((char*)cs + align256(sizeof(float) * nf * np));
  float2* ccar = (float2*)((char*)cy + align256(sizeof(float) * nf * np));
  float2* cchi = (float2*)((char*)ccar + align256(sizeof(float2) * nf * np));
  SelCrop src{};
  SelBlend dst{};
  for (int j = 0; j < nf; ++j) {
    src.s[j] = F[j].s;
    src.car[j] = (const float2*)F[j].carrier;
    dst.y[j] = F[j].y;
    dst.chi[j] = (float2*)F[j].chi;
  }
  std::vector<int> ones(nf, 1);
  for (const SelBox& b : boxes) {
    const int bh = b.fy1 - b.fy0, bw = b.x1 - b.x0;
    const size_t nb = (size_t)bh * bw;
    hipLaunchKernelGGL(k_sel_crop, dim3((bw + 255) / 256, bh, nf), dim3(256), 0, st, src, w, b.fy0, b.x0, bh, bw, cs,
                       ccar);
    if (hipGetLastError() != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "selective 3-D: crop launch failed");
    std::vector<hvd_gpu_plane> cp(nf);
    for (int j = 0; j < nf; ++j) {
      hvd_gpu_plane& q = cp[j];
      q = hvd_gpu_plane{};
      q.lines = bh;
      q.width = bw;
      q.s = cs + (size_t)j * nb;
      q.carrier = (const float*)(ccar + (size_t)j * nb);
      q.active_video_start = F[j].active_video_start + b.x0;
      q.phase_per_sample = F[j].phase_per_sample;
      q.sample_rate = F[j].sample_rate;
      q.chi = (float*)(cchi + (size_t)j * nb);
      q.y = cy + (size_t)j * nb;
    }
    rc = hvd_gpu_holographic_init(ctx, cp.data(), nf, 0, st);
    if (rc != HVD_GPU_OK) return rc;
    rc = drive(ctx, cp.data(), nf, par, is_pal, ones.data(), p, &c3, nullptr, st);
    if (rc != HVD_GPU_OK) return rc;
    hipLaunchKernelGGL(k_sel_blend, dim3((bw + 255) / 256, bh, nf), dim3(256), 0, st, dst, w, b.fy0, b.x0, bh, bw,
                       (const float*)cy, (const float2*)cchi);
    if (hipGetLastError() != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "selective 3-D: blend launch failed");
  }
  return HVD_GPU_OK;
}

}  // namespace hvd_window
