// hvd_kernels.h -- kernel declarations and the descriptor structs the host
// side fills for them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hvd_device.h"

namespace hvdk {

constexpr int kMaxLines = 640;   // lines per field the lock-in handles in LDS
constexpr int kColGroup = 8;     // columns per block in the column FFTs
constexpr int kRowGroup = 2;     // rows per block in the init's row FFTs
constexpr int kBlendW = 32;      // holographic blend tile
constexpr int kBlendH = 8;
constexpr int kMaxRadix = 16;    // stages per 1-D FFT plan

struct LockinGeom {
  int fw, fh, x0, x1;
  int std_phase;       // phase per sample exactly pi/2 (no custom subcarrier)
  int vswitch;         // PAL family: joint phase + V-switch parity
  double w0;           // phase per sample (rad)
  double line_advance; // per-line model advance (rad)
  float nominal_burst; // 20 NTSC / PAL-M, 21.43 PAL
  int pad_;
};

struct PrepDesc {
  const float* field;        // fh x fw IRE
  const float* theta_field;  // fh
  const int8_t* vswitch_field;  // fh or null
  float* s;                  // lines x width
  float2* rowph;             // lines
  int8_t* vsw_active;        // lines or null
  float2* colph;             // width
  int fw, fal, a0, lines, width;
  int std_phase, signed_by_parity;
  float phase_offset;        // -chroma_phase_deg * pi / 180
  double w0;
  float* th_active;          // lines or null: the trimmed float theta of the active lines
};

// A prime stage of radix R >= kRaderMin runs by Rader's algorithm: its R - 1
// non-zero inputs, permuted by powers of a generator g of (Z/R)*, are
// cyclically convolved with the permuted roots by a length-(R - 1) transform
// of small radices.
constexpr int kRaderMin = 37;
constexpr int kMaxRader = 2;       // Rader stages per plan (37 * 41 * 43 > 4096)
constexpr int kMaxInnerRadix = 8;  // stages of the length-(R - 1) transform

struct FftRader {
  int L, nst;                  // L = R - 1 and its stage count
  int radix[kMaxInnerRadix];   // small radices and primes < kRaderMin
  int tab;  // float2 offset in the root buffer of [L roots of L | bhat fwd L | bhat inv L | perm L]
  int pad_;
};

struct FftPlan {
  int n, nst;
  int radix[kMaxRadix];
  int root_off;  // offset of this size's n roots in the shared root buffer
  int nrader;
  signed char rader_of[kMaxRadix];  // -1, or the index into rader[] of stage s
  FftRader rader[kMaxRader];
  // float2 slots of LDS the Rader stages need beyond the two buffers and the
  // n roots: rader_fixed + rader_per_col x (interleaved sequences)
  int rader_fixed, rader_per_col;
};

// LDS (float2 slots) of an in-LDS transform of `ncols` interleaved sequences
// of plan P: two buffers, the n roots, the Rader tables.
__host__ __device__ inline int fft_lds_slots(const FftPlan& P, int ncols) {
  return 2 * P.n * ncols + P.n + P.rader_fixed + P.rader_per_col * ncols;
}

struct InitField {
  int h, w, hp, pad;
  const float* s;
  const float2* rowph;
  const int8_t* vsw;
  const float2* colph;
  const float2* carrier;
  float2* spec;     // hp x w
  float2* var[3];   // h x w each
  float2* chi0;     // out
  float* y0;        // out (may be null)
  int row_plan, col_plan;
  int nvar;
  float sample_rate;
};

struct SadDesc {
  const uint16_t* ref;
  const uint16_t* cur;
  int16_t* vy;
  int16_t* vx;
  int32_t* sad;
  int h, w, bh, bw, sy, sx, nby, nbx;
};

struct MotionDesc {
  const float* ref;
  const float* cur;
};

// hvd::FrameParams (frame_bridge.h:40-101) as the device sees it.
struct BridgeFrameParams {
  int standard, frame_width, frame_height, field1_lines;
  int active_video_start, active_video_end, colour_burst_start, colour_burst_end;
  int first_active_frame_line, last_active_frame_line;
  float black_level, white_level, blanking_level, chroma_dc;
  double sample_rate, subcarrier_hz;
  int extra_lines[8];
  int extra_count;
  int pad_;
};

// One decoded field as the weave reads it.
struct BridgeField {
  int lines, width;
  const float* y;
  const float2* chi;
  const float2* rowph;
  const float2* colph;
  const int8_t* vsw;
  const float2* carrier;  // explicit carrier plane (lines x width) or null: the tables
};

struct BridgeOut {
  int16_t* luma;
  int16_t* chroma;
  double* u;
  double* v;
};

// lockin.hip
__global__ void k_lockin_demod(const LockinGeom* G, const float* const* fields, const int* line_prefix, int nf,
                               double* z);
__global__ void k_lockin(const LockinGeom* G, const double* z, const int* line_prefix, float* const* theta,
                         int8_t* const* vswitch);
__global__ void k_burst_amp(const LockinGeom* G, const float* const* fields, float* out);
__global__ void k_u16_to_ire(const uint16_t* raw, size_t n, float blanking, float cpi, float* out);
struct IreDesc {
  const uint16_t* raw;
  float* out;
  size_t n;
};
__global__ void k_u16_to_ire_batch(const IreDesc* D, float blanking, float cpi);
__global__ void k_prepare(const PrepDesc* P);
__global__ void k_crop(const PrepDesc* P);

// holo.hip
// The init's four FFT passes (0 rows forward, 1 columns forward, 2 columns
// inverse, 3 rows inverse) over the fields listed in `fields` (device array,
// nf entries; `prefix` the running count of blocks, kRowGroup rows or
// kColGroup columns per block) that share the transform class `cls`
// (fft_init_class of their row or column plan).
int fft_init_class(const FftPlan& p, bool rows);
void launch_fft_pass(int pass, int cls, int nblocks, size_t lds, hipStream_t st, const InitField* F,
                     const FftPlan* plans, const float2* roots, const int* prefix, const int* fields, int nf);
hipError_t fft_allow_lds(int bytes);  // dynamic LDS limit of every pass kernel
__global__ void k_fft2_rows(const float2* in, float2* out, int h, int w, FftPlan P, const float2* roots, int sign);
__global__ void k_fft2_cols(float2* io, int h, int w, FftPlan P, const float2* roots, int sign, float scale);
__global__ void k_difference(int op, int nc, const float* a, const float* wt, int h, int w, float* d);
__global__ void k_holo_blend(const InitField* F, const int* tp, int nf);

// solve.hip
__global__ void k_refine_begin(FieldCG* st, const FieldCG* init, int nf);
__global__ void k_aniso_samples(const DevField* F, const FieldCG* st, const int* sample_prefix, int nf, float* gx,
                                float* gy);
constexpr int kSelThreads = 1024;  // threads of the aniso p98 selection block
__global__ void k_aniso_select(FieldCG* st, const int* sample_prefix, const float* gx, const float* gy,
                               float lambda_c, int lds_floats);
__global__ void k_irls_weights(const DevField* F, const FieldCG* st, StripGrid sg, int stage);
__global__ void k_cg_start(const DevField* F, const FieldCG* st, StripGrid sg, double* partials, int adjoint,
                           int stage);
__global__ void k_cg_iter(const DevField* F, FieldCG* st, StripGrid sg, double* partials, int it);
__global__ void k_cg_iter_adj(const DevField* F, FieldCG* st, StripGrid sg, double* partials, int it);
__global__ void k_cg_finish(FieldCG* st, StripGrid sg, const double* partials, int n_inner, int stage);
__global__ void k_refine_finalize(const DevField* F, const FieldCG* st, StripGrid sg);

// The per-pixel solve (pixsolve.hip): the oriented chroma prior and
// motion-compensated temporal neighbours (variational.cpp:198-556 with
// diag_prior > 0 or NeighborTerm, variational.h:54-58), fields of a batch
// each with their own neighbours.
constexpr int kMaxNeighbors = 16;
constexpr int kPxWaves = 4;                            // waves per block
constexpr int kPxPerLane = 2;                          // pixels per lane
constexpr int kPxChunk = 64 * kPxWaves * kPxPerLane;  // pixels per block

// Per-field CG state of the per-pixel solve beside its FieldCG.
struct PxState {
  double gg0;
  int done[2];    // stop flag: k_px_dir reads [0] and writes [1], k_px_step the reverse
  int iters[16];  // CG iterations per IRLS stage
};

struct PxField {
  DevField f;               // s, carrier (plane or tables), chi[0..1], d[0..1], y
  float2* g;                // gradient at chi
  float4* wts;              // wx, wy, wcx, wcy times the prior scale
  float2* wd;               // wd1, wd2 (oriented prior), or null
  float4* nb[kMaxNeighbors];  // per neighbour: S_w - S, dc = c - c_w, robust weight wt
  const float* conf[kMaxNeighbors];
  const float* sw[kMaxNeighbors];   // neighbour composite
  const float2* cw[kMaxNeighbors];  // neighbour carrier
  int nn;
  float nu, et2;
  int nchunks;
  double* part;             // 4 * nchunks: |g|^2 by parity, curv(d), Re<g, d>
  PxState* px;
  const FieldCG* st;        // mu_h, mu_v, mu_d, lam_y, eps, eps_c, k, exits, n_inner
  // the joint solve (VariationalRefineJoint): Y a free unknown beside chi.
  // nb[j] then holds {S_w, c_w, wt} (not the differences), and the anchor
  // (luma, chroma, weight) is optional
  int joint;
  float nu_a;
  float* yv[2];             // Y ping-pong (yv[0] the caller's: Y0 in, Y out)
  float* dyv[2];            // dY ping-pong
  float* gy;                // dE/dY
  const float* a_luma;
  const float2* a_chroma;
  const float* a_weight;
};

__global__ void k_px_prep(const PxField* F, const int* cp, int nf);
__global__ void k_px_weights(const PxField* F, const int* cp, int nf, int cur, int stage);
__global__ void k_px_grad0(const PxField* F, const int* cp, int nf, int cur);
__global__ void k_px_dir(const PxField* F, const int* cp, int nf, int it, int stage);
__global__ void k_px_step(const PxField* F, const int* cp, int nf, int it, int cur, int stage);
__global__ void k_px_finalize(const PxField* F, const int* cp, int nf, int cur);
__global__ void k_pj_weights(const PxField* F, const int* cp, int nf, int cur, int stage);
__global__ void k_pj_grad0(const PxField* F, const int* cp, int nf, int cur);
__global__ void k_pj_dir(const PxField* F, const int* cp, int nf, int it, int stage);
__global__ void k_pj_step(const PxField* F, const int* cp, int nf, int it, int cur, int stage);
__global__ void k_pj_finalize(const PxField* F, const int* cp, int nf, int cur);

// warp.hip
__global__ void k_vpp_snap(int th, int tw, const float* mdy, const float* mdx, float* sy, float* sx);
__global__ void k_vpp_interp(int th, int tw, int tile, int out_h, int out_w, const float* sy, const float* sx,
                             float* dy, float* dx);
__global__ void k_warp_tiles(int h, int w, int ch, const float* a, const float* dy, const float* dx, float* out);
__global__ void k_warp_bilinear(int hs, int ws, const float* a, const float* dy, const float* dx, float row_offset,
                                int out_h, int out_w, float* out);
__global__ void k_conf_nearest(int th, int tw, int tile, int out_h, int out_w, const float* conf, float* out);
__global__ void k_square(int n, const float* a, float* out);
__global__ void k_coh_terms(size_t n, const float2* z1, const float2* z2, float* cre, float* cim, float* m1,
                            float* m2);
__global__ void k_coh_out(size_t n, const float* bre, const float* bim, const float* b1, const float* b2, float* out);

// The woven active picture (hvd_chroma_decoder_stage.cpp:316-413): output
// width x height; columns from a0; rows from frame line y0 (raster view) or,
// in the single-field view (field 0 / 1), that field's lines from fl0.
struct WovenDesc {
  int width, height, a0, y0, field, fl0, f1, fw, fh, pad_;
  double black, range;  // luma excursion black .. white of the RGB rendering
};

// bridge.hip
__global__ void k_woven_rgb(WovenDesc D, const int16_t* luma, const double* u, const double* v, double* yo,
                            double* uo, double* vo, uint8_t* rgb);
__global__ void k_split_fields(const BridgeFrameParams* FP, const int16_t* const* frames, float zero,
                               float* const* fields);
__global__ void k_window_decisions(const BridgeFrameParams* FP, const float* const* fields, int nframes,
                                   const float* amps, int acc, float chroma_gain, float nominal_burst, int field_order,
                                   int* out_f1_bottom, float* out_gain, int* out_vote);
__global__ void k_weave_requant(const BridgeFrameParams* FP, const int16_t* const* frames, int core_begin,
                                const BridgeField* dec, const int* f1_bottom, const float* gain_p, int monochrome,
                                const BridgeOut* out);
__global__ void k_weave_woven(const BridgeFrameParams* FP, const int16_t* src, int rows, int width, const float* y,
                              const float* chroma, const float2* chi, float gain, BridgeOut o);
__global__ void k_make_carrier(const float* theta, const int8_t* vswitch, int width, int a0, float w0,
                               float2* carrier);
__global__ void k_chroma_plane(int mode, const float* s, const float* yl, const float2* chi, const float2* car,
                               size_t n, float* out);
__global__ void k_tridiag(const float* d, const float* a, int n, float lam, double* scratch, float* out);

// One field of the solve's adjoint (adjoint.hip).
struct AdjField {
  int h, w;
  const float* s;
  const float2* rowph;
  const float2* colph;
  const int8_t* vsw;
  const float2* carrier;  // explicit carrier, or null (factored)
  const float2* abar;     // adjoint solution of the current stage
  const float2* chik;     // chi at the end of the stage
  const float2* chip;     // chi at its start (the weights' argument)
  const float4* wts;      // the stage's weights
  float2* gbar;           // dL/dchi: the stage's right-hand side, then dL/dchi_{k-1}
  float* sbar;            // dL/dS accumulator
  float* scratch;         // 6 floats per pixel
  const float2* gchi;     // seeds
  const float* gy;
};

__global__ void k_adj_seed(const AdjField* A);
__global__ void k_adj_local(const AdjField* A, const FieldCG* st, int nbx, double* partials);
__global__ void k_adj_reduce(const FieldCG* st, int nbx, const double* partials, double* glam);
__global__ void k_adj_transpose(const AdjField* A);
__global__ void k_zero2(float2* a, size_t n);
__global__ void k_zero_d(double* a, int n);

// motion.hip
__global__ void k_sad_bm(const SadDesc* D, const int* blk_prefix, int nf);
// The box blur's largest radius.
constexpr int kBlurMaxR = 32;
// Rows up to kBlurMaxW samples take the along-x pass's parallel prefix sum.
constexpr int kBlurMaxW = 2048;
__global__ void k_me_blur_cols(const MotionDesc* D, int npairs, int h, int w, int r, float* tmp);
__global__ void k_me_blur_rows(int npairs, int h, int w, int r, const float* tmp, float* out);
// Both passes of BoxBlur (motion.cpp:21-55) over npairs * 2 planes (D), the
// along-y pass into tmp, the result into out (planes one after another).
void launch_me_blur(hipStream_t st, const MotionDesc* D, int npairs, int h, int w, int r, float* tmp, float* out);
void launch_me_blur_planes(hipStream_t st, const MotionDesc* D, int nplanes, int h, int w, int r, float* tmp,
                           float* out);
__global__ void k_me_decimate(int npairs, int h, int w, int f, const float* in, float* out);
__global__ void k_me_ssd_wave(int npairs, int h, int w, int tile, const float* planes, const int* dys,
                              const int* dxs, int ndy, int ndx, const float* base_dy, const float* base_dx, float* se);
// The staged SSD form (one wave per tile, a lane per candidate): tiles up to
// kSsdTile, windows (tile plus the candidates' span) up to kSsdWin.
constexpr int kSsdTile = 32;
constexpr int kSsdWin = 48;
__global__ void k_me_ssd_tile(int npairs, int h, int w, int tile, const float* planes, const int* dys, const int* dxs,
                              int ndy, int ndx, const float* base_dy, const float* base_dx, float* se);
__global__ void k_me_ssd(int npairs, int h, int w, int tile, const float* planes, const int* dys, const int* dxs,
                         int ndy, int ndx, const float* base_dy, const float* base_dx, float* se);
__global__ void k_me_select(int npairs, int nt, const int* dys, const int* dxs, int ndy, int ndx, const float* se,
                            const float* base_dy, const float* base_dx, float* mdy, float* mdx, float* best,
                            float* cost0);
__global__ void k_me_upscale(int npairs, int th, int tw, int hh, int ww, int f, const float* cdy, const float* cdx,
                             float* bdy, float* bdx);
__global__ void k_me_offset(int n, const float* mdy, const float* mdx, float ody, float odx, float* tdy, float* tdx);
__global__ void k_me_finish(int n, const float* cym, const float* cyp, const float* cxm, const float* cxp,
                            const float* best0, const float* cost0, float* mdy, float* mdx, float* best, float* out_dy,
                            float* out_dx);
__global__ void k_me_confidence(int h, int w, int tile, const float* planes, const float* best, float* const* conf_out,
                                float* scratch, int have_tiles);
__global__ void k_me_conf_tiles(int h, int w, int tile, const float* planes, const float* best, float* scratch);
// MotionConfidence of n pairs: the per-tile moments from LDS-staged tile
// rows when a band of tile x w floats fits kConfBandMaxLds, then the median
// and scalings (motion.hip).
constexpr size_t kConfBandMaxLds = 144 * 1024;
void launch_me_confidence(hipStream_t st, int n, int h, int w, int tile, const float* planes, const float* best,
                          float* const* conf_out, float* scratch);

}  // namespace hvdk
