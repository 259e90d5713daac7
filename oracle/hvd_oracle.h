/*
 * hvd_oracle.h -- CPU restatement of hvd-core's Y/C-separation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP path:
 * only tests, the smoke check and the benchmark's CPU-baseline leg load it.
 * No product code links against it.
 *
 * The same source builds twice:
 *   liboracle64.so  OR_REAL = double  -- the FP64 oracle every GPU result is
 *                                        compared against;
 *   liboracle32.so  OR_REAL = float   -- FP32 per-pixel arithmetic with FP64
 *                                        reductions, the precision hvd-core
 *                                        itself runs at (used to estimate the
 *                                        FP32-vs-FP64 drift of the solve and
 *                                        as the "port" CPU baseline).
 * Every exported symbol carries the suffix of its build (or64_ / or32_) so the
 * two libraries can be loaded side by side.
 *
 * Complex planes are interleaved (re, im) row-major arrays.  Planes are
 * row-major [h][w].  All functions are re-entrant; the FFT keeps no state.
 */
#ifndef HVD_ORACLE_H_
#define HVD_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field geometry, field-sample domain (engine/ntsc_geometry.h:142-267). */
typedef struct {
  int field_width;
  int field_height;
  int active_video_start;
  int active_video_end;
  int colour_burst_start;
  int colour_burst_end;
  int first_active_field_line;
  int last_active_field_line; /* 0 => field_height */
  double sample_rate;
  double subcarrier_hz; /* 0 => standard */
  int standard;         /* 0 NTSC, 1 PAL, 2 PAL-M */
} or_geom;

/* Solver knobs (engine/hvd_config.h:26-102) plus lambda_y, the luma prior
 * weight (1.0 reproduces hvd-core exactly: variational.cpp:302 has an
 * implicit unit weight). */
typedef struct {
  double lambda_c;
  double chroma_aniso; /* 0 => AUTO */
  double charbonnier_eps;
  double chroma_eps;
  double structure_coupling;
  double diag_prior;
  int cg_iterations;
  int irls_outer;
  double cg_tol; /* 0 => auto */
  int fast;
  int is_pal;
  double lambda_y;
} or_cfg;

#ifdef __cplusplus
}
#endif

#endif /* HVD_ORACLE_H_ */
