// window.h -- the 3-D window driver's kernels (window.hip), its host entry
// point and the selective window (selective.hip), called by
// hvd_gpu_decode_window and the frame window in abi.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "hvd_gpu.h"
#include "hvd_kernels.h"

namespace hvdk {

// One neighbour equation of a colour group (window.hip, the batched form of
// MotionCompensatePrev plus the coherence and odd-offset gates).
struct NbJob {
  const float* mdy;     // tile motion (snapped copy), th x tw
  const float* mdx;
  const float* mconf;   // tile confidence, th x tw
  const float* s_k;     // neighbour composite, carrier, chroma (h x w)
  const float2* car_k;
  const float2* chi_k;
  const float2* chi_j;  // the field's own chroma (the coherence gate)
  float* comp;          // outputs: warped composite, carrier, confidence
  float2* car;
  float* conf;
  float* sq;            // scratch: squared tile confidence, snapped vectors and map (th x tw each)
  float* sy;
  float* sx;
  float* sc;
  float* cn;            // slow mode: the nearest tile's squared confidence (h x w)
  float* cb;            //            and its blur
  float* coh;           // the gate: the four coherence terms (4 h x w)
  float* cb4;           //           and their blur
  const float* vgrad;   // the odd-offset envelope or null
  const float* lapsm;
};

// One field's pairwise motion measurements (at most kMaxNeighbors), passed
// to the trajectory kernels by value: no table upload, nothing to wait for.
struct TrajTab {
  float* dy[kMaxNeighbors];
  float* dx[kMaxNeighbors];
  const float* conf[kMaxNeighbors];
  float ho[kMaxNeighbors];  // parity shift of each measurement
  int off[kMaxNeighbors];   // its field offset
};
// The envelope planes of one colour (at most kEnvMax fields), by value.
constexpr int kEnvMax = 64;
struct EnvRows {
  const float* src[2 * kEnvMax];
  float* out[2 * kEnvMax];
};
struct EnvFinish {
  const float* vg_in[kEnvMax];
  float* vgrad[kEnvMax];
  float* lap[kEnvMax];
  float* sample[kEnvMax];
  float* floor2;  // kEnvMax floats: each field's (1.5 median)^2
};
// Tile-map copies (dy and dx, 2 nt floats each) of one colour, by value.
constexpr int kCopyMax = 128;
struct CopyTab {
  const float* src[kCopyMax];
  float* dst[kCopyMax];
};

__global__ void k_verify_pick(int n, const float* pdy, const float* pdx, const float* bp, const float* bz, float* dy,
                              float* dx, float* best);
__global__ void k_noise_sigma(const float* const* s, int h, int w, float* d_all, float* sigma);
__global__ void k_amb_bands(const float* const* s, const float* const* car, int stride, int h, int w, float nr_eps,
                            int nbands, double* acc, int* cnt);
__global__ void k_fidelity(size_t n, const float* s, const float2* chi, const float2* car, float* y);
__global__ void k_traj_predict(int m, int nt, TrajTab T, float* vy, float* vx);
__global__ void k_traj_extrapolate(int nt, const float* vy, const float* vx, float o, float ho, float* pdy,
                                   float* pdx);
__global__ void k_traj_snap(int m, int nt, TrajTab T);
__global__ void k_env_terms(int h, int w, const float* Y, const float2* chi, float* gmag, float* lapmag);
__global__ void k_smooth_rows(int h, int w, EnvRows P);
__global__ void k_env_finish(int n, EnvFinish P);
__global__ void k_env_apply(int n, EnvFinish P);
__global__ void k_copy_tiles(int n, CopyTab P);
__global__ void k_split_cplx(size_t n, const float2* z, float* re, float* im);
__global__ void k_join_cplx(size_t n, const float* re, const float* im, float2* z);
__global__ void k_conf_gates(size_t n, float* conf, const float* gamma, float a, const float* vgrad,
                             const float* lapsm, float e2, float floor_max);
__global__ void k_synth_init(size_t n, const float* Y, const float2* chi, float* aY, float* aCr, float* aCi,
                             float* aW);
__global__ void k_synth_accum(size_t n, const float* Y, const float2* chi, const float* conf, const float* Yw,
                              const float* wre, const float* wim, float eps_b, float* aY, float* aCr, float* aCi,
                              float* aW);
__global__ void k_synth_finish(size_t n, const float* s, const float2* car, const float* aY, const float* aCr,
                               const float* aCi, const float* aW, float* luma, float2* chroma, float* resid);
__global__ void k_synth_weight(size_t n, const float* r, float eps_b, float* weight);
__global__ void k_psi(size_t n, int nn, const float* s, const float2* car, const float* const* sw,
                      const float2* const* cw, const float* const* conf, float2* chi);

__global__ void k_traj_fit(int m, int nt, TrajTab T, float* vy, float* vx);
__global__ void k_nb_square(int nt, const NbJob* J);
__global__ void k_nb_snap(int th, int tw, int fast, const NbJob* J);
__global__ void k_nb_warp(int h, int w, int th, int tw, int tile, int fast, int gate, const NbJob* J);
__global__ void k_nb_gates(int h, int w, int fast, int gate, float a, float e2, float floor_max, const NbJob* J);

}  // namespace hvdk

// Services of abi.hip the driver uses.
namespace hvd_window {
// Grow-only device memory of the driver, separate from the component calls'
// workspace; null (with the context's error set) when it cannot be had.
char* pool(hvd_gpu_ctx* ctx, size_t bytes, hipStream_t st);
// Records a failure message on the context; returns code.
int fail(hvd_gpu_ctx* ctx, int code, const char* msg);
// DrizzleFrame (drizzle.hip): 0 or an HVD_GPU_ERR_* code.
int drizzle(hvd_gpu_ctx* ctx, int j0, int n, const float* const* Ys, const float* const* chis, const int* par,
            int L, int W, int tile, int search, int nr_radius, float nr_eps, int scale, float* out_luma,
            float* out_chroma, hipStream_t st);
// The window driver (window.hip): 0 or an HVD_GPU_ERR_* code.
int drive(hvd_gpu_ctx* ctx, const hvd_gpu_plane* fields, int n, const int* parities, int is_pal, const int* y_given,
          const hvd_solver_params* p, const hvd_window_params* wp, hvd_window_diag* diag, hipStream_t st);
// Grow-only device memory of the selective window, apart from the driver's
// pool and the workspace (both of which the crops' decodes use).
char* sel_pool(hvd_gpu_ctx* ctx, size_t bytes, hipStream_t st);
// One box of the selective window, field coordinates [fy0, fy1) x [x0, x1).
struct SelBox {
  int fy0, fy1, x0, x1;
};
// Scratch bytes of ambiguous_boxes for a lines x width field.
size_t sel_scratch_bytes(int h, int w);
// AmbiguousBoxes (selective.hip) of a decoded field: the boxes, or
// *everywhere (the whole window in 3-D); scores_out optional.  Synchronous.
int ambiguous_boxes(hvd_gpu_ctx* ctx, const float* y, const float* chi, int h, int w, float max_area, char* scratch,
                    hipStream_t st, std::vector<SelBox>* out, bool* everywhere, std::vector<float>* scores_out);
// DecodeWindowSelective (selective.hip) over the driver's fields.
int drive_selective(hvd_gpu_ctx* ctx, const hvd_gpu_plane* fields, int n, const int* parities, int is_pal,
                    const int* y_given, const hvd_solver_params* p, const hvd_window_params* wp,
                    hvd_window_diag* diag, hipStream_t st);
}  // namespace hvd_window
