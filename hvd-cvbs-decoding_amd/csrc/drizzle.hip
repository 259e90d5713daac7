// drizzle.hip -- DrizzleFrame (hvd-core engine/sequence.cpp:283-421): the
// vertical super-resolution of two opposite fields by scatter accumulation
// of their window neighbours onto a fine (2 lines x scale) grid, robustly
// weighted against the target field's own decode, blended with the woven
// linear interpolation where coverage is thin.
//
// hvd-core deposits sequentially, in the order (target field jt, source field
// k, source row y, column x, upper / lower fine row), so every fine pixel's
// float sums have one order of adds.  Here every deposit is formed in
// parallel with its sequence number; the deposits are bucketed per fine
// pixel (a count, an exclusive scan, a scatter), each bucket is put back in
// sequence order and summed in that order: hvd-core's float bits, whatever
// order the GPU formed them in.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "hvd_kernels.h"
#include "warp_ops.h"
#include "window.h"

namespace hvdk {

// One (jt, k) pair: field k's samples deposited for target field jt.
struct DzJob {
  const float* yk;     // L x W
  const float2* ck;
  const float* yt;     // the target's decode (the robust agreement)
  const float2* ct;
  const float* sy;     // snapped tile vectors and confidence of EstimateMotion(Y_k, Y_jt); null: k == jt
  const float* sx;
  const float* conf;
  int pk;              // parity of k
};

// (1) every deposit: its fine pixel and its four products (ww Y_k, ww Re
// chi_k, ww Im chi_k, ww) at index ((job L + y) W + x) 2 + off; counts per
// fine pixel.
__global__ void k_dz_deposit(const DzJob* J, int L, int W, int th, int tw, int tile, int scale, float eps_b,
                             int* target, float4* prod, int* count) {
#pragma clang fp contract(off)
  const DzJob& jb = J[blockIdx.y];
  const size_t np = (size_t)L * W;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int y = (int)(i / W), x = (int)(i % W);
  const int HF = 2 * L * scale;
  float cpx = 1.0f, vy = 0.0f, vx = 0.0f;
  if (jb.sy) {
    cpx = wops::tile_at(jb.conf, tw, min(y / tile, th - 1), min(x / tile, tw - 1));
    vy = wops::vpp_interp(th, tw, tile, jb.sy, y, x);
    vx = wops::vpp_interp(th, tw, tile, jb.sx, y, x);
  }
  const int syi = min(max((int)lroundf((float)y - vy), 0), L - 1);
  const int sxi = min(max((int)lroundf((float)x - vx), 0), W - 1);
  const size_t si = (size_t)syi * W + sxi;
  const float dY = jb.yk[i] - jb.yt[si];
  const float2 dc = make_float2(jb.ck[i].x - jb.ct[si].x, jb.ck[i].y - jb.ct[si].y);
  const float d2 = dY * dY + (dc.x * dc.x + dc.y * dc.y);
  const float w = cpx * eps_b * eps_b / (d2 + eps_b * eps_b);
  const float yf = (2.0f * ((float)y + vy) + (float)jb.pk) * (float)scale;
  const int xs = min(max((int)lroundf((float)x + vx), 0), W - 1);
  const int y0 = (int)floorf(yf);
  const float fy = yf - (float)y0;
  const float2 ck = jb.ck[i];
  const float sk = jb.yk[i];
  const size_t base = ((size_t)blockIdx.y * np + i) * 2;
  for (int off = 0; off < 2; ++off) {
    const int yt = min(max(y0 + off, 0), HF - 1);
    const float ww = w * (off == 0 ? (1.0f - fy) : fy);
    const int t = yt * W + xs;
    target[base + off] = t;
    prod[base + off] = make_float4(ww * sk, ww * ck.x, ww * ck.y, ww);
    atomicAdd(&count[t], 1);
  }
}

// (2) every deposit's seat in its fine pixel's bucket (any order).
__global__ void k_dz_bucket(size_t nd, const int* target, const int* start, int* fill, int* seat) {
  const size_t d = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nd) return;
  const int t = target[d];
  seat[start[t] + atomicAdd(&fill[t], 1)] = (int)d;
}

// (3) per fine pixel: the bucket sorted back into deposit order, the sums in
// that order, then the blend with the woven fallback (sequence.cpp:371-421).
__global__ void k_dz_finish(int L, int W, int scale, int p0, int p1, const float* y0, const float2* c0,
                            const float* y1, const float2* c1, const int* start, const int* count, int* seat,
                            const float4* prod, float* out_y, float2* out_c) {
#pragma clang fp contract(off)
  const int HF = 2 * L * scale;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)HF * W) return;
  const int r = (int)(i / W), x = (int)(i % W);
  int* s = seat + start[i];
  const int m = count[i];
  for (int a = 1; a < m; ++a)
    for (int b = a; b > 0 && s[b] < s[b - 1]; --b) {
      const int z = s[b];
      s[b] = s[b - 1];
      s[b - 1] = z;
    }
  float aY = 0.f, aCr = 0.f, aCi = 0.f, aW = 0.f;
  for (int a = 0; a < m; ++a) {
    const float4 q = prod[s[a]];
    aY += q.x;
    aCr += q.y;
    aCi += q.z;
    aW += q.w;
  }
  // the woven rows 2 r' + parity of the two fields, the second written last
  auto wy = [&](int f, float* Y, float* Cr, float* Ci) {
    const int rr = f >> 1;
    const bool second = (f & 1) == p1;
    const bool first = (f & 1) == p0;
    // the weave writes field j0's rows then field j0 + 1's, so a row both
    // parities claim holds field j0 + 1's sample
    if (second) {
      *Y = y1[(size_t)rr * W + x];
      *Cr = c1[(size_t)rr * W + x].x;
      *Ci = c1[(size_t)rr * W + x].y;
    } else if (first) {
      *Y = y0[(size_t)rr * W + x];
      *Cr = c0[(size_t)rr * W + x].x;
      *Ci = c0[(size_t)rr * W + x].y;
    } else {
      *Y = 0.f;
      *Cr = 0.f;
      *Ci = 0.f;
    }
  };
  const float fine = (float)r / (float)scale;
  const int f0 = min(max((int)floorf(fine), 0), 2 * L - 2);
  const float ff = fine - (float)f0;
  float ya, cra, cia, yb, crb, cib;
  wy(f0, &ya, &cra, &cia);
  wy(f0 + 1, &yb, &crb, &cib);
  const float baseY = ya * (1.0f - ff) + yb * ff;
  const float bCr = cra * (1.0f - ff) + crb * ff;
  const float bCi = cia * (1.0f - ff) + cib * ff;
  const float kLam = 0.35f;
  const float mix = aW / (aW + kLam);
  const float dzY = aW > 0.0f ? aY / fmaxf(aW, 1e-9f) : 0.0f;
  const float dzCr = aW > 0.0f ? aCr / fmaxf(aW, 1e-9f) : 0.0f;
  const float dzCi = aW > 0.0f ? aCi / fmaxf(aW, 1e-9f) : 0.0f;
  out_y[i] = mix * dzY + (1.0f - mix) * baseY;
  out_c[i] = make_float2(mix * dzCr + (1.0f - mix) * bCr, mix * dzCi + (1.0f - mix) * bCi);
}

__global__ void k_dz_zero(size_t n, int* a) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0;
}

}  // namespace hvdk

namespace hvd_window {

using namespace hvdk;

int drizzle(hvd_gpu_ctx* ctx, int j0, int n, const float* const* Ys, const float* const* chis, const int* par,
            int L, int W, int tile, int search, int nr_radius, float nr_eps, int scale, float* out_luma,
            float* out_chroma, hipStream_t st) {
  if (!Ys || !chis || !par || !out_luma || !out_chroma) return fail(ctx, HVD_GPU_ERR_INVALID, "null argument");
  if (j0 < 0 || j0 + 1 >= n) return fail(ctx, HVD_GPU_ERR_INVALID, "j0 and j0 + 1 must be fields of the window");
  if (L < 2 || W < 1 || scale < 1 || tile < 1 || tile > 256 || search < 0 || search > 64 || nr_radius < 0)
    return fail(ctx, HVD_GPU_ERR_INVALID, "bad drizzle geometry or motion parameters");
  for (int k = 0; k < n; ++k)
    if (!Ys[k] || !chis[k]) return fail(ctx, HVD_GPU_ERR_INVALID, "null field plane");
  const float eps_b = nr_eps > 0.0f ? nr_eps : 3.0f;
  const int th = (L + tile - 1) / tile, tw = (W + tile - 1) / tile, nt = th * tw;
  const size_t np = (size_t)L * W, HF = (size_t)2 * L * scale, nfine = HF * W;
  // the (jt, k) jobs in hvd-core's order (sequence.cpp:302-310)
  struct Pair {
    int jt, k;
  };
  std::vector<Pair> pairs;
  for (int jt = j0; jt <= j0 + 1; ++jt)
    for (int k = std::max(0, j0 - 2 * nr_radius); k < std::min(n, j0 + 2 + 2 * nr_radius); ++k) {
      if ((k == j0 || k == j0 + 1) && k != jt) continue;
      pairs.push_back({jt, k});
    }
  const int nj = (int)pairs.size();
  std::vector<int> mi(nj, -1);  // motion slot of each job (-1: the target itself)
  std::vector<const float*> refs, curs;
  for (int q = 0; q < nj; ++q)
    if (pairs[q].k != pairs[q].jt) {
      mi[q] = (int)refs.size();
      refs.push_back(Ys[pairs[q].k]);
      curs.push_back(Ys[pairs[q].jt]);
    }
  const int nm = (int)refs.size();
  const size_t nd = (size_t)nj * np * 2;
  size_t scan_bytes = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int*)nullptr, (int*)nullptr, (int)nfine, st) !=
      hipSuccess)
    return fail(ctx, HVD_GPU_ERR_HIP, "scan sizing failed");
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += (b + 255) & ~size_t(255);
    return o;
  };
  const size_t o_mot = take(sizeof(float) * 5 * nt * (size_t)std::max(nm, 1));  // dy dx conf sy sx
  const size_t o_jobs = take(sizeof(DzJob) * nj);
  const size_t o_tgt = take(sizeof(int) * nd), o_prod = take(sizeof(float4) * nd), o_seat = take(sizeof(int) * nd);
  const size_t o_cnt = take(sizeof(int) * nfine), o_start = take(sizeof(int) * nfine),
               o_fill = take(sizeof(int) * nfine), o_scan = take(scan_bytes);
  char* base = pool(ctx, off, st);
  if (!base) return HVD_GPU_ERR_NOMEM;
  float* mot = (float*)(base + o_mot);
  auto mfield = [&](int s, int c) { return mot + ((size_t)s * 5 + c) * nt; };
  if (nm > 0) {
    std::vector<float*> ody(nm), odx(nm), ocf(nm);
    for (int s = 0; s < nm; ++s) {
      ody[s] = mfield(s, 0);
      odx[s] = mfield(s, 1);
      ocf[s] = mfield(s, 2);
    }
    const int rc = hvd_gpu_estimate_motion(ctx, refs.data(), curs.data(), nm, L, W, tile, search, ody.data(),
                                           odx.data(), ocf.data(), st);
    if (rc != HVD_GPU_OK) return rc;
  }
  std::vector<DzJob> jobs(nj);
  for (int q = 0; q < nj; ++q) {
    DzJob& jb = jobs[q];
    jb.yk = Ys[pairs[q].k];
    jb.ck = (const float2*)chis[pairs[q].k];
    jb.yt = Ys[pairs[q].jt];
    jb.ct = (const float2*)chis[pairs[q].jt];
    jb.pk = par[pairs[q].k];
    if (mi[q] >= 0) {
      jb.sy = mfield(mi[q], 3);
      jb.sx = mfield(mi[q], 4);
      jb.conf = mfield(mi[q], 2);
    }
  }
  // VectorsPerPixel's snap of every motion field (temporal.cpp:44-66)
  for (int s = 0; s < nm; ++s)
    hipLaunchKernelGGL(k_vpp_snap, dim3((nt + 255) / 256), dim3(256), 0, st, th, tw, (const float*)mfield(s, 0),
                       (const float*)mfield(s, 1), mfield(s, 3), mfield(s, 4));
  DzJob* dj = (DzJob*)(base + o_jobs);
  int* tgt = (int*)(base + o_tgt);
  float4* prod = (float4*)(base + o_prod);
  int* seat = (int*)(base + o_seat);
  int* cnt = (int*)(base + o_cnt);
  int* start = (int*)(base + o_start);
  int* fill = (int*)(base + o_fill);
  if (hipMemcpyAsync(dj, jobs.data(), sizeof(DzJob) * nj, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(ctx, HVD_GPU_ERR_HIP, "job upload failed");
  const unsigned gz = (unsigned)std::min<size_t>((nfine + 255) / 256, 8192);
  hipLaunchKernelGGL(k_dz_zero, dim3(gz), dim3(256), 0, st, nfine, cnt);
  hipLaunchKernelGGL(k_dz_zero, dim3(gz), dim3(256), 0, st, nfine, fill);
  hipLaunchKernelGGL(k_dz_deposit, dim3((unsigned)((np + 255) / 256), nj), dim3(256), 0, st, (const DzJob*)dj, L, W,
                     th, tw, tile, scale, eps_b, tgt, prod, cnt);
  if (hipGetLastError() != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "deposit launch failed");
  if (hipcub::DeviceScan::ExclusiveSum(base + o_scan, scan_bytes, cnt, start, (int)nfine, st) != hipSuccess)
    return fail(ctx, HVD_GPU_ERR_HIP, "bucket scan failed");
  hipLaunchKernelGGL(k_dz_bucket, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, nd, (const int*)tgt,
                     (const int*)start, fill, seat);
  hipLaunchKernelGGL(k_dz_finish, dim3((unsigned)((nfine + 255) / 256)), dim3(256), 0, st, L, W, scale, par[j0],
                     par[j0 + 1], Ys[j0], (const float2*)chis[j0], Ys[j0 + 1], (const float2*)chis[j0 + 1],
                     (const int*)start, (const int*)cnt, seat, (const float4*)prod, out_luma, (float2*)out_chroma);
  if (hipGetLastError() != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "drizzle launch failed");
  if (hipStreamSynchronize(st) != hipSuccess) return fail(ctx, HVD_GPU_ERR_HIP, "drizzle sync failed");
  return HVD_GPU_OK;
}

}  // namespace hvd_window
