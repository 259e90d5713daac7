// bridge.hip -- the frame_bridge side of the field path on the device:
// field-sequential int16 frames in, field-sequential Y/C frames out.
//
// Restates hvd-core frame_bridge.cpp for the 2-D sequence window
// (DecodeFrameSequenceWindow :722-947): SplitFrameFields (:61-99) with PAL's
// non-orthogonal line offsets (FrameLineOffset, frame_bridge.h:106-112),
// per-field BurstAmplitudeIre -> window median -> AccGain (:776-795), the
// window majority field-order vote over DetectFieldParity (:102-149,
// :759-765), and the parity weave with 10-bit requantisation (:867-945,
// ClampU10 / ClampS10 :25-39).
#include "hvd_device.h"
#include "hvd_kernels.h"

// Requantisation must round exactly as hvd-core's float code does: no fused
// multiply-add.
#pragma clang fp contract(off)

namespace hvdk {

namespace {

__device__ inline long long line_offset(const BridgeFrameParams& fp, int line) {
  long long off = (long long)line * fp.frame_width;
  for (int i = 0; i < fp.extra_count; ++i)
    if (fp.extra_lines[i] < line) ++off;
  return off;
}

__device__ inline int16_t clamp_u10(float v) {
  const long r = lroundf(v);
  return (int16_t)(r < 0 ? 0 : (r > 1023 ? 1023 : r));
}
__device__ inline int16_t clamp_s10(float v) {
  const long r = lroundf(v);
  return (int16_t)(r < -1023 ? -1023 : (r > 1023 ? 1023 : r));
}

}  // namespace

// SplitFrameFields: frame t -> fields[2t] (block 1) and fields[2t + 1]
// (block 2), each field1_lines x frame_width, in IRE relative to the code
// `zero` (blanking for composite; chroma_dc for a separate chroma channel,
// frame_bridge.cpp:276-297); rows past a block's end read `zero`.
// Grid: (ceil(fw / 256), field1_lines, nframes).
__global__ void k_split_fields(const BridgeFrameParams* FP, const int16_t* const* frames, float zero,
                               float* const* fields) {
  const BridgeFrameParams& fp = *FP;
  const int t = blockIdx.z, y = blockIdx.y;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int fw = fp.frame_width, f1 = fp.field1_lines, f2 = fp.frame_height - f1;
  if (x >= fw) return;
  const int16_t* fr = frames[t];
  const float st = y < f1 ? (float)fr[line_offset(fp, y) + x] : zero;
  const float sb = y < f2 ? (float)fr[line_offset(fp, f1 + y) + x] : zero;
  const float cpi = (fp.white_level - fp.blanking_level) / 100.0f;
  fields[2 * t][(size_t)y * fw + x] = (st - zero) / cpi;
  fields[2 * t + 1][(size_t)y * fw + x] = (sb - zero) / cpi;
}

// DetectFieldParity for every frame, then the window majority vote and the
// window ACC gain (median burst amplitude, AccGain with the standard's
// nominal burst).  One block; out[0] = f1_bottom, out_gain[0] = gain,
// out_vote[0] = the summed per-frame votes.  `amps` may be null when acc is 0.
__global__ void __launch_bounds__(256) k_window_decisions(const BridgeFrameParams* FP, const float* const* fields,
                                                          int nframes, const float* amps, int acc, float chroma_gain,
                                                          float nominal_burst, int field_order, int* out_f1_bottom,
                                                          float* out_gain, int* out_vote) {
  const BridgeFrameParams& fp = *FP;
  const int fw = fp.frame_width, f1 = fp.field1_lines;
  const int fal = fp.first_active_frame_line / 2;
  const int lal = fp.last_active_frame_line != 0 ? fp.last_active_frame_line / 2 : f1;
  const int a0 = fp.active_video_start, a1 = fp.active_video_end;
  const int lines = lal - fal, width = a1 - a0;
  __shared__ double sp[256], sm[256];
  __shared__ int s_vote;
  __shared__ float s_sel;
  if (threadIdx.x == 0) s_vote = 0;
  __syncthreads();
  const int ystep = max(1, lines / 120);
  const int nx = (a1 - a0 >= 4) ? ((a1 - 4 - a0) / 16 + 1) : 0;  // x = a0, a0 + 16, ..., x + 4 <= a1
  const int nr = (lal - 1 - (fal + 1) > 0) ? ((lal - 2 - (fal + 1)) / ystep + 1) : 0;
  for (int t = 0; t < nframes && field_order == 0; ++t) {
    const float* b1 = fields[2 * t];
    const float* b2 = fields[2 * t + 1];
    double ep = 0.0, em = 0.0;
    if (lines >= 8 && width >= 32) {
      for (int k = threadIdx.x; k < nr * nx; k += blockDim.x) {
        const int r = fal + 1 + (k / nx) * ystep, x = a0 + (k % nx) * 16;
        auto lp = [&](const float* p, int row) {
          const float* q = p + (size_t)row * fw + x;
          return 0.25f * (((q[0] + q[1]) + q[2]) + q[3]);
        };
        const float b = lp(b2, r), ac = lp(b1, r), au = lp(b1, r - 1), ad = lp(b1, r + 1);
        ep += fabs((double)(b - 0.5f * (ac + ad)));
        em += fabs((double)(b - 0.5f * (ac + au)));
      }
    }
    sp[threadIdx.x] = ep;
    sm[threadIdx.x] = em;
    __syncthreads();
    if (threadIdx.x == 0 && lines >= 8 && width >= 32 && nr * nx > 0) {
      double tp = 0.0, tm = 0.0;
      for (int i = 0; i < (int)blockDim.x; ++i) {
        tp += sp[i];
        tm += sm[i];
      }
      tp /= (double)(nr * nx);
      tm /= (double)(nr * nx);
      const double lo = fmin(tp, tm), hi = fmax(tp, tm);
      if (!(hi <= 1e-9 || (hi - lo) / hi < 0.02)) s_vote += tp < tm ? 1 : -1;
    }
    __syncthreads();
  }
  // ACC: the element of rank n/2 of the 2 * nframes burst amplitudes
  const int na = acc ? 2 * nframes : 0;
  for (int i = threadIdx.x; i < na; i += blockDim.x) {
    int rank = 0;
    for (int j = 0; j < na; ++j) rank += (amps[j] < amps[i]) || (amps[j] == amps[i] && j < i);
    if (rank == na / 2) s_sel = amps[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out_f1_bottom[0] = field_order == 2 ? 1 : (field_order == 1 ? 0 : (s_vote < 0 ? 1 : 0));
    float gain = 1.0f;
    if (acc) gain = fminf(fmaxf(nominal_burst / fmaxf(s_sel, 1.0f), 0.5f), 2.0f);
    out_gain[0] = chroma_gain * gain;
    out_vote[0] = s_vote;
  }
}

// Parity weave + requantisation of core frames (frame_bridge.cpp:867-945).
// Grid: (ceil(fw / 256), frame_height, ncore).  Outside the active picture:
// composite passthrough, zero chroma.
__global__ void k_weave_requant(const BridgeFrameParams* FP, const int16_t* const* frames, int core_begin,
                                const BridgeField* dec, const int* f1_bottom, const float* gain_p, int monochrome,
                                const BridgeOut* out) {
  const BridgeFrameParams& fp = *FP;
  const int k = blockIdx.z, t = core_begin + k;
  const int line = blockIdx.y;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int fw = fp.frame_width, f1 = fp.field1_lines;
  if (x >= fw) return;
  const BridgeOut& o = out[k];
  const size_t idx = (size_t)line * fw + x;
  int16_t lum = frames[t][line_offset(fp, line) + x];
  int16_t chr = 0;
  double u = 0.0, v = 0.0;
  const int fal = fp.first_active_frame_line / 2;
  const int a0 = fp.active_video_start;
  const BridgeField& d0 = dec[2 * t];
  const int lines = d0.lines, aw = d0.width;
  // which woven active row (if any) lands on this flat line
  int r = -1;
  if (line >= fal && line < fal + lines) r = 2 * (line - fal);
  else if (line >= f1 + fal && line < f1 + fal + lines) r = 2 * (line - f1 - fal) + 1;
  const int c = x - a0;
  if (r >= 0 && c >= 0 && c < aw) {
    const int p0 = f1_bottom[0] ? 1 : 0;
    const bool use_first = (r % 2) == p0;
    const BridgeField& d = use_first ? d0 : dec[2 * t + 1];
    const int fl = r / 2;
    if (fl < d.lines) {
      const size_t i = (size_t)fl * d.width + c;
      const float y_ire = d.y[i];
      const float2 chi = monochrome ? make_float2(0.f, 0.f) : d.chi[i];
      float2 car;
      if (d.carrier) {
        car = d.carrier[i];
      } else {
        car = cmul(d.rowph[fl], d.colph[c]);
        if (d.vsw && d.vsw[fl] < 0) car.x = -car.x;
      }
      const float scale = (fp.white_level - fp.blanking_level) / 100.0f;
      lum = clamp_u10((fp.blanking_level + (y_ire * scale)));
      const float re = ((chi.x * car.x) - (chi.y * car.y));
      chr = clamp_s10((re * scale));
      const float g = gain_p[0];
      v = (double)g * (double)chi.x * (double)scale;
      u = (double)g * (double)(-chi.y) * (double)scale;
    }
  }
  o.luma[idx] = lum;
  o.chroma[idx] = chr;
  o.u[idx] = u;
  o.v[idx] = v;
}

// Parity weave of a WOVEN decode (DecodeFrameBuffer :161-255 and
// DecodeYcFrameBuffer :263-377): woven row r belongs to flat block r % 2 at
// field line first_active + r / 2.  Every line first takes `src` (the
// composite, or the separate luma channel) through FrameLineOffset; inside
// the active picture luma becomes ClampU10(blank + Y scale) when `y` is given
// (left as passed through otherwise), chroma ClampS10(C scale), and
// v / u = gain (Re chi, -Im chi) scale in double.  Grid: (ceil(fw / 256), fh).
__global__ void k_weave_woven(const BridgeFrameParams* FP, const int16_t* src, int rows, int width, const float* y,
                              const float* chroma, const float2* chi, float gain, BridgeOut o) {
  const BridgeFrameParams& fp = *FP;
  const int line = blockIdx.y;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int fw = fp.frame_width, f1 = fp.field1_lines;
  if (x >= fw) return;
  const size_t idx = (size_t)line * fw + x;
  int16_t lum = src[line_offset(fp, line) + x];
  int16_t chr = 0;
  double u = 0.0, v = 0.0;
  const int fal = fp.first_active_frame_line / 2;
  const int field = line < f1 ? 0 : 1;
  const int r = 2 * ((field == 0 ? line : line - f1) - fal) + field;
  const int c = x - fp.active_video_start;
  if (r >= 0 && r < rows && c >= 0 && c < width) {
    const size_t i = (size_t)r * width + c;
    const float scale = (fp.white_level - fp.blanking_level) / 100.0f;
    if (y) lum = clamp_u10((fp.blanking_level + (y[i] * scale)));
    chr = clamp_s10((chroma[i] * scale));
    const float2 ph = chi[i];
    v = (double)gain * (double)ph.x * (double)scale;
    u = (double)gain * (double)(-ph.y) * (double)scale;
  }
  o.luma[idx] = lum;
  o.chroma[idx] = chr;
  o.u[idx] = u;
  o.v[idx] = v;
}

// MakeCarrier / MakeCarrierPal (ntsc_geometry.cpp:29-47, :75-93):
// c(y, x) = polar(1, theta[y] + w0 (a0 + x)) in float, -conj(c) on lines
// whose vswitch < 0.  Grid: (ceil(width / 256), lines).
__global__ void k_make_carrier(const float* theta, const int8_t* vswitch, int width, int a0, float w0,
                               float2* carrier) {
#pragma clang fp contract(off)
  const int yy = blockIdx.y;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= width) return;
  const float phase = theta[yy] + w0 * (float)(a0 + x);
  float sn, cs;
  sincosf(phase, &sn, &cs);
  if (vswitch && vswitch[yy] < 0) cs = -cs;
  carrier[(size_t)yy * width + x] = make_float2(cs, sn);
}

// FrameYc's chroma plane (engine.cpp:181-183, :221-224): mode 0 C = S - Y,
// mode 1 C = Re[chi c].
__global__ void k_chroma_plane(int mode, const float* s, const float* yl, const float2* chi, const float2* car,
                               size_t n, float* out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (mode == 0) {
      out[i] = s[i] - yl[i];
    } else {
      const float2 a = chi[i], b = car[i];
      out[i] = (a.x * b.x) - (a.y * b.y);
    }
  }
}

// The woven active picture of a field-sequential YcFrameS16 and its RGB24
// rendering: ReorderToWoven (hvd_chroma_decoder_stage.cpp:316-413; raster
// rows alternate fields, or one field's lines in the single-field view) and
// YuvToRgb8 (:79-88) over the luma excursion black..white
// (WovenToPreviewImage / WriteWovenAsRgb24, :500-638), in double as there.
// One thread per output pixel; grid (ceil(width / 256), height).
__global__ void k_woven_rgb(WovenDesc D, const int16_t* luma, const double* u, const double* v, double* yo,
                            double* uo, double* vo, uint8_t* rgb) {
#pragma clang fp contract(off)
  const int row = blockIdx.y;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= D.width || row >= D.height) return;
  int flat_line;
  if (D.field == 0 || D.field == 1) {
    const int fl = D.fl0 + row;
    flat_line = D.field == 0 ? fl : D.f1 + fl;
  } else {
    const int frame_line = D.y0 + row;
    const int field_line = frame_line / 2;
    flat_line = (frame_line % 2 == 0) ? field_line : D.f1 + field_line;
  }
  const int flat_col = D.a0 + col;
  double yv = 0.0, uv = 0.0, vv = 0.0;
  if (flat_line >= 0 && flat_line < D.fh && flat_col >= 0 && flat_col < D.fw) {
    const size_t src = (size_t)flat_line * D.fw + flat_col;
    yv = (double)luma[src];
    uv = u[src];
    vv = v[src];
  }
  const size_t dst = (size_t)row * D.width + col;
  if (yo) yo[dst] = yv;
  if (uo) uo[dst] = uv;
  if (vo) vo[dst] = vv;
  if (rgb) {
    const double ny = (yv - D.black) / D.range;
    const double nu = uv / D.range;
    const double nv = vv / D.range;
    const double c3[3] = {ny + 1.13983 * nv, ny - 0.39465 * nu - 0.58060 * nv, ny + 2.03211 * nu};
    for (int k = 0; k < 3; ++k) {
      const double cl = fmin(fmax(c3[k], 0.0), 1.0);
      rgb[dst * 3 + k] = (uint8_t)(cl * 255.0 + 0.5);
    }
  }
}

}  // namespace hvdk
