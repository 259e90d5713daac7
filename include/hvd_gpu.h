/*
 * hvd_gpu.h -- C ABI of the MI355X (gfx950) Y/C-separation engine.
 *
 * The host side of hvd-core (its C++ engine API and frame_bridge) calls
 * these entry points instead of its CPU loops; each declaration names the
 * hvd-core routine it replaces.  Plain C: pointers, sizes, POD structs and
 * int status codes; no C++ types or exceptions cross this boundary.
 *
 * Conventions
 *  - Planes are row-major; complex planes are interleaved (re, im) float
 *    pairs, i.e. std::complex<float> / hvd::ComplexPlane storage.
 *  - Every data pointer is a DEVICE pointer unless the parameter says host.
 *  - Work is ordered on `stream` (a hipStream_t; NULL = the default stream)
 *    and is asynchronous unless stated otherwise.
 *  - A context owns its device workspace; calls on one context must be
 *    issued from one host thread on one stream at a time, like an
 *    hvd::HvdEngine (engine.h:70-76, frame_bridge.h:139-141).
 *  - Status: HVD_GPU_OK (0) or a positive error code; the message is
 *    available from hvd_gpu_last_error().  Invalid arguments are rejected
 *    before any work is queued.
 */
#ifndef HVD_GPU_H_
#define HVD_GPU_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HVD_GPU_OK 0
#define HVD_GPU_ERR_INVALID 1     /* bad argument / size out of range */
#define HVD_GPU_ERR_HIP 2         /* HIP runtime error */
#define HVD_GPU_ERR_UNSUPPORTED 3 /* a configuration this build does not run */
#define HVD_GPU_ERR_NOMEM 4

typedef struct hvd_gpu_ctx hvd_gpu_ctx;

/* hvd::FieldGeometry (engine/ntsc_geometry.h:142-267). */
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
  double subcarrier_hz; /* 0 => the standard's nominal fsc */
  int standard;         /* 0 NTSC, 1 PAL, 2 PAL-M (hvd::VideoStandard) */
} hvd_field_geometry;

/* The solver knobs of hvd::HvdConfig the hot path reads
 * (engine/hvd_config.h:26-102), plus lambda_y: the luma prior weight,
 * implicitly 1 in hvd-core (variational.cpp:302); 1.0 reproduces it. */
typedef struct {
  double lambda_c;
  double chroma_aniso; /* 0 => AUTO (variational.cpp:130-196) */
  double charbonnier_eps;
  double chroma_eps;
  double structure_coupling;
  double diag_prior; /* >= 0; > 0 adds the oriented +-45 deg prior (variational.cpp:205-219),
                        solved by the per-pixel kernels; the adjoint entry points take 0 */
  int cg_iterations;
  int irls_outer;
  double cg_tol; /* 0 => auto (0.02 slow / 0.10 fast) */
  int fast;
  int is_pal; /* derived from the geometry by the field entry points */
  double lambda_y;
} hvd_solver_params;

/* One field's active plane for the init and the solve.  The carrier is
 * either an explicit interleaved plane (`carrier` non-NULL, as hvd-core's
 * ComplexPlane) or, with `carrier` NULL, the factored form
 *     c(y, x) = exp(i (theta[y] + phase_per_sample (active_video_start + x)))
 * with c -> -conj(c) on lines whose vswitch < 0 (MakeCarrier /
 * MakeCarrierPal, ntsc_geometry.cpp:29-93), which costs no memory traffic. */
typedef struct {
  int lines;
  int width;
  const float* s;        /* lines * width, IRE */
  const float* theta;    /* lines (factored carrier) */
  const int8_t* vswitch; /* lines, or NULL */
  const float* carrier;  /* lines * width * 2, or NULL */
  int active_video_start;
  double phase_per_sample;
  double sample_rate;
  float* chi; /* lines * width * 2: refine in/out; init out */
  float* y;   /* lines * width: luma out, may be NULL */
} hvd_gpu_plane;

/* The HvdConfig knobs of hvd-core's window driver (engine/hvd_config.h,
 * sequence.cpp:371-1114) beyond the solver's. */
typedef struct {
  double temporal_strength; /* < 0: 3-D off; 0: adaptive (measured); > 0: fixed */
  double temporal_eps;      /* <= 0: 7 sigma, clamped to [4, 20] */
  double nr_anchor;         /* anchor weight of the passes >= 1; 0: no anchor */
  double nr_eps;            /* <= 0: 3 sigma, clamped to [3, 12] */
  double coherence_gate;
  double odd_gate_floor;
  int passes;
  int nr_radius;
  int mc_tile;
  int mc_search;
  int bidirectional;
  int extended_temporal;
  int trajectory_fit;
  int psi_init;
  int output_fidelity;
  int parallel_across_fields;
  /* DecodeWindowSelective (frame_bridge.cpp:638-720), when non-zero with
   * cg_iterations > 0 and temporal_strength >= 0: the window in 2-D, then the
   * boxes of the first field's ambiguous tiles decoded again in 3-D and
   * feather-blended in.  hvd-core runs it from DecodeFrameSequenceWindow
   * only (HvdConfig::selective_3d, off by default). */
  int selective_3d;
  float selective_max_area; /* boxes above this fraction of the field: full 3-D */
} hvd_window_params;

/* hvd::SequenceDiagnostics (sequence.h:73-79): what the driver decided. */
typedef struct {
  float resolved_strength;
  float ambiguity_ire;
  float sigma_ire;
  float temporal_eps;
  float nr_eps;
  int selective_boxes; /* selective 3-D: boxes decoded in 3-D; 0 none, -1 the whole window */
} hvd_window_diag;

/* hvd::FrameParams (frame_bridge.h:40-101): a field-sequential frame's
 * geometry and levels in the 10-bit sample domain.  extra_sample_lines lists,
 * once per extra sample, the frame-flat lines that carry one more sample
 * than frame_width (PAL: 312 and 624, twice each). */
typedef struct {
  int standard; /* 0 NTSC, 1 PAL, 2 PAL-M */
  int frame_width;
  int frame_height;
  int field1_lines;
  int active_video_start;
  int active_video_end;
  int colour_burst_start;
  int colour_burst_end;
  int first_active_frame_line;
  int last_active_frame_line; /* 0 => from the field height */
  float black_level;
  float white_level;
  float blanking_level;
  float chroma_dc;
  double sample_rate;
  double subcarrier_hz;
  int32_t extra_sample_lines[8];
  int extra_sample_line_count;
} hvd_frame_params;

/* The HvdConfig fields the frame bridge reads besides the solver's. */
typedef struct {
  float chroma_gain;
  int acc;
  int field_order; /* 0 auto (measured), 1 field 1 top, 2 field 1 bottom */
  int monochrome;
  float chroma_phase_deg;
  float temporal_strength; /* < 0: 3-D off */
} hvd_bridge_params;

/* hvd::YcFrameS16 planes (frame_bridge.h:123-131), each frame_width x
 * frame_height, DEVICE pointers. */
typedef struct {
  int16_t* luma;
  int16_t* chroma;
  double* u_plane;
  double* v_plane;
} hvd_yc_frame;

/* ---- context ---------------------------------------------------------- */

/* Creates a context on HIP device `device` (replaces the construction of
 * hvd::HvdEngine, engine.h:70-76, and the Fft2d it owns, fft2d.h:22-58). */
int hvd_gpu_create(int device, hvd_gpu_ctx** out);
void hvd_gpu_destroy(hvd_gpu_ctx* ctx);
/* Message for the last failing call on this context ("" if none). */
const char* hvd_gpu_last_error(const hvd_gpu_ctx* ctx);
/* Blocks until the context's queued work on `stream` is complete. */
int hvd_gpu_synchronize(hvd_gpu_ctx* ctx, void* stream);
/* ABI version (major * 100 + minor). */
int hvd_gpu_abi_version(void);

/* Diagnostics (no hvd-core counterpart): with profiling enabled every
 * refine records HIP events on its stream around each IRLS stage's CG loop
 * and keeps a copy of the per-field iteration counts.  hvd_gpu_cg_timing()
 * synchronizes the device and returns, over all refines since profiling was
 * enabled or last read: the summed CG-loop time (ms), the number of CG
 * iterations launched in those loops (one stencil launch and one reduction
 * each, per solve group), and sum over fields of pixels x CG iterations run;
 * it then clears the record. */
int hvd_gpu_set_profiling(hvd_gpu_ctx* ctx, int enable);
int hvd_gpu_cg_timing(hvd_gpu_ctx* ctx, double* ms, long long* iterations, double* pixel_iterations);

/* ---- solver launch shape (no hvd-core counterpart) ----------------------- */

/* How the IRLS/CG solve is launched: lines per solver strip (2..32; each
 * 64-lane wave walks a strip of 59 columns) and the most pixels per solve
 * group (consecutive fields solved together, sized so their planes stay in
 * the Infinity Cache).  0 selects the built-in choice for that member. */
typedef struct {
  int strip_rows;
  long long group_pixels;
} hvd_gpu_solver_tuning;

/* Fixes the launch shape for every later solve on this context; zeros hand
 * the choice back to the autotuned entries and the built-in policy.  The
 * results of a solve depend on the strip height only through the order of
 * its FP64 sums. */
int hvd_gpu_set_solver_tuning(hvd_gpu_ctx* ctx, const hvd_gpu_solver_tuning* t);

/* The autotuner: times the CG loop of `nfields` synthetic fields of
 * lines x width (PAL: with a V-switched carrier) for every candidate strip
 * height and solve-group size, `iterations` CG iterations each, and records
 * the fastest for that field shape.  Later solves whose fields all share the
 * shape use that entry when their field count matches; otherwise they take
 * the strip height of the entry with the nearest count and the default group
 * size.  A fixed tuning wins over both.
 * Returns the choice and its CG time per pixel-iteration in ns.  Allocates
 * and frees its own synthetic planes, synchronizes the device and clears
 * the profiling record. */
int hvd_gpu_autotune_solver(hvd_gpu_ctx* ctx, int lines, int width, int is_pal, int nfields, int iterations,
                            hvd_gpu_solver_tuning* best, double* ns_per_pixel_iteration);

/* HIP graphs (no hvd-core counterpart): with `enable` (the default) the
 * launches of each forward solve group -- every IRLS stage's start pass, one
 * k_cg_iter launch per CG iteration and the stage's closing launch, and the
 * output pass -- are replayed from one HIP graph built at first use for that
 * launch sequence and workspace, instead of being launched one by one.
 * Results are identical either way.  Disabling drops the cached graphs. */
int hvd_gpu_set_graphs(hvd_gpu_ctx* ctx, int enable);

/* ---- level conversion -------------------------------------------------- */

/* uint16 codes -> IRE, ire = (code - blanking) / ((white - blanking) / 100)
 * (SampleToIre, engine/ntsc_geometry.h:301-304; the split of int16 frames
 * in SplitFrameFields, frame_bridge.cpp:61-99, is the same conversion). */
int hvd_gpu_sample_to_ire_u16(hvd_gpu_ctx* ctx, const uint16_t* raw, size_t n, float blanking, float white,
                              float* out, void* stream);

/* ---- burst lock-in ----------------------------------------------------- */

/* Per-line carrier phase for `n` fields: BurstLockinPhase (NTSC,
 * engine/lockin.cpp:101-186) or, for PAL-family geometries, the joint phase
 * + V-switch estimate BurstLockinPhasePal (lockin.cpp:257-349).
 * fields[i]: field_height x field_width IRE; theta[i]: field_height floats;
 * vswitch[i]: field_height int8 (written for PAL-family fields only; may be
 * NULL for NTSC).  `fields`, `theta`, `vswitch` and `geom` are HOST arrays of
 * n entries (the pointers they hold are device pointers). */
int hvd_gpu_burst_lockin_phase(hvd_gpu_ctx* ctx, const float* const* fields, int n, const hvd_field_geometry* geom,
                               float* const* theta, int8_t* const* vswitch, void* stream);

/* Burst amplitude in IRE for ACC, BurstAmplitudeIre (lockin.cpp:188-255);
 * amp: device array of n floats. */
int hvd_gpu_burst_amplitude_ire(hvd_gpu_ctx* ctx, const float* const* fields, int n, const hvd_field_geometry* geom,
                                float* amp, void* stream);

/* ---- holographic init and variational solve ------------------------------ */

/* HolographicInit (engine/holographic_init.cpp:200-317) for n planes of
 * possibly different sizes: writes chi0 into planes[i].chi and Y0 into
 * planes[i].y.  `planes` is a HOST array. */
int hvd_gpu_holographic_init(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, int symmetry_variant,
                             void* stream);

/* VariationalRefine, 2-D path (engine/variational.cpp:198-556) for n planes
 * in one batch: planes[i].chi holds chi0 on entry and chi on return;
 * planes[i].y receives S - Re[chi c].  p->is_pal selects the AUTO aniso
 * line grouping for every plane.  Optional HOST outputs: stage_iters
 * (n * irls_outer ints, CG iterations run per stage) and aniso (n floats);
 * asking for either makes the call synchronous. */
int hvd_gpu_variational_refine(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, const hvd_solver_params* p,
                               int* stage_iters, float* aniso, void* stream);

/* VerifyMotion (motion.cpp:490-525) for n pairs of h x w planes: per tile
 * the rounded prediction (pdy, pdx: th x tw floats) against the zero
 * vector, one SSD each on the r = 2 blurred planes; the winner (zero on
 * ties; the prediction keeps its float value) and its motion confidence. */
int hvd_gpu_verify_motion(hvd_gpu_ctx* ctx, const float* const* y_ref, const float* const* y_cur, int n, int h,
                          int w, const float* const* pdy, const float* const* pdx, int tile, float* const* dy,
                          float* const* dx, float* const* conf, void* stream);

/* FitTrajectory (motion.h:89-95, motion.cpp:532-562) of m <= 16 pairwise
 * tile motion fields of one field (dy, dx, conf: DEVICE, tiles_h x
 * tiles_w each; offsets and parity_shifts: HOST, m entries): per tile the
 * median over the confident (conf > 0.15) nonzero offsets of (dy - h) / o
 * and dx / o, 0 where none qualifies, into vy, vx (DEVICE). */
int hvd_gpu_trajectory_fit(hvd_gpu_ctx* ctx, int m, const int* offsets, const float* parity_shifts,
                           const float* const* dy, const float* const* dx, const float* const* conf, int tiles_h,
                           int tiles_w, float* vy, float* vx, void* stream);

/* TrajectorySnap (motion.h:97-106, motion.cpp:564-601): the fit above, then
 * on tiles where at least three measurements lie within 1.5 px (L1) of o v
 * + h those measurements snap onto it; dy, dx in place.  m < 2: no change. */
int hvd_gpu_trajectory_snap(hvd_gpu_ctx* ctx, int m, const int* offsets, const float* parity_shifts, float* const* dy,
                            float* const* dx, const float* const* conf, int tiles_h, int tiles_w, void* stream);

/* DrizzleFrame (sequence.h:105-130, sequence.cpp:283-421): fields j0 and
 * j0 + 1 of a window of n decoded fields (Ys: lines x width floats, chis:
 * interleaved complex; DEVICE; parities HOST) stacked onto a (2 lines scale)
 * x width fine grid from their neighbours k in [j0 - 2 nr_radius, j0 + 2 +
 * 2 nr_radius), each matched to the target field by EstimateMotion (tile,
 * search), robustly weighted (scale nr_eps, 3 when <= 0) and blended with the
 * woven interpolation.  out_luma (2 lines scale x width) and out_chroma
 * (interleaved) DEVICE.  The fine pixels' sums are formed in hvd-core's
 * deposit order.  Synchronous. */
int hvd_gpu_drizzle_frame(hvd_gpu_ctx* ctx, int j0, int n, const float* const* ys, const float* const* chis,
                          const int* parities, int lines, int width, int tile, int search, int nr_radius,
                          float nr_eps, int scale, float* out_luma, float* out_chroma, void* stream);

/* hvd-core's window driver, DecodeFieldWindowWithInits
 * (engine/sequence.cpp:371-1114, sequence.h:104-135), over n fields of one
 * shape: fields[j].s, an explicit carrier, chi0 in fields[j].chi and, where
 * y_given[j] is non-zero, Y0 in fields[j].y (else Y0 = S - Re[chi0 c], as
 * for an empty DecodedField::luma; y_given may be null); on return chi and
 * Y (S - Re[chi c] with output_fidelity, else the anchored luma) in the same
 * planes.  parities: n entries (0 top, 1 bottom).  is_pal selects the PAL
 * offsets and defaults (sequence.cpp:452-476) and the aniso grouping.  The
 * noise self-calibration, the adaptive strength, per-pass pair motion with
 * fast mode's prediction + VerifyMotion and the trajectory snap, the
 * motion-compensated neighbours with the coherence and odd-offset gates,
 * the PSI init, SynthReference and the joint pass all run on the device;
 * fields of one colour of the Gauss-Seidel sweep are solved as one batch.
 * diag: optional HOST output.  Synchronous. */
int hvd_gpu_decode_window(hvd_gpu_ctx* ctx, const hvd_gpu_plane* fields, int n, const int* parities, int is_pal,
                          const int* y_given, const hvd_solver_params* p, const hvd_window_params* wp,
                          hvd_window_diag* diag, void* stream);

/* AmbiguousBoxes (frame_bridge.cpp:399-621) of one decoded field (y, chi:
 * DEVICE, lines x width): per 16 x 16 tile the score sqrt(mean (Y -
 * box(Y))^2 x mean (|chi| - box|chi|)^2) with count-normalised 8-sample box
 * means, tiles at or above 4 x the median with >= 2 such neighbours (the
 * outer ring excluded), bands of adjacent tile rows with >= 2 of them (at
 * most 4, merged across the smallest gaps), one box per run of columns
 * (gaps <= 2 tiles, single tiles dropped) with 16-row and 48-column
 * margins.  *nboxes (HOST): 0 no box (the window stays 2-D); -1 ambiguity
 * everywhere, or boxes above max_area of the field in total (the whole
 * window in 3-D); else the count, the first `cap` of them in boxes (HOST,
 * 4 ints each: first row, end row, first column, end column).  scores:
 * optional HOST output, ceil(lines / 16) x ceil(width / 16) floats.
 * Synchronous. */
int hvd_gpu_ambiguous_boxes(hvd_gpu_ctx* ctx, const float* y, const float* chi, int lines, int width,
                            float max_area, int* boxes, int cap, int* nboxes, float* scores, void* stream);

/* Device scratch of hvd_gpu_variational_refine_noalloc: 64 MiB per GPU, held
 * from the library's first context on that GPU until the process ends. */
#define HVD_GPU_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* No counterpart in hvd-core.  hvd_gpu_variational_refine for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same kernels and the same
 * results, with the solver's workspace and descriptors in a fixed per-GPU
 * scratch array instead of the context's workspace.  It calls no
 * hipMalloc/hipFree/hipHostMalloc and builds no HIP graph (the launches are
 * issued one by one).  Limits: at most 256 planes and 63 MiB of solver
 * workspace per call (about 40 B per pixel: six NTSC or PAL fields);
 * larger calls fail with HVD_GPU_ERR_INVALID before any work is queued, as
 * does a context with profiling on.  Every call on one GPU shares that GPU's
 * scratch, from every context: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller. */
int hvd_gpu_variational_refine_noalloc(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                       const hvd_solver_params* p, int* stage_iters, float* aniso, void* stream);

/* One motion-compensated temporal neighbour of a field (hvd::NeighborTerm,
 * variational.h:54-58), on the field's lines x width grid: its composite S_w,
 * its carrier c_w (interleaved complex) and a per-pixel confidence. */
typedef struct {
  const float* composite;  /* lines * width */
  const float* carrier;    /* lines * width * 2 */
  const float* confidence; /* lines * width */
} hvd_gpu_neighbor;

/* VariationalRefine with temporal neighbours (variational.cpp:198-556 with
 * `neighbors` non-empty; engine.h:82-88): one field, the 2-D objective plus
 * per neighbour the robust data term temporal_strength * wt * rt^2 with
 * rt = (S_w - S) + Re[chi (c - c_w)] and wt = conf eps^2 / (rt^2 + eps^2)
 * lagged per IRLS stage, and the smoothness weights scaled by
 * 1 + temporal_strength * sum wt |c - c_w|^2.  nn <= 16.  With nn == 0 or
 * temporal_strength <= 0 every temporal term vanishes and this is
 * hvd_gpu_variational_refine of the one plane.  temporal_eps must be
 * positive (hvd-core's sequence resolves it before the solve,
 * sequence.cpp:445-446).  stage_iters: optional HOST output, irls_outer
 * ints (makes the call synchronous). */
int hvd_gpu_variational_refine_temporal(hvd_gpu_ctx* ctx, const hvd_gpu_plane* plane, const hvd_gpu_neighbor* nbrs,
                                        int nn, const hvd_solver_params* p, float temporal_strength,
                                        float temporal_eps, int* stage_iters, void* stream);

/* VariationalRefine with temporal neighbours (variational.cpp:198-556) for a
 * batch: plane i with nn[i] (<= 16) neighbours taken in order from nbrs
 * (back to back; nn NULL: none).  Fields advance together in one launch per
 * CG step; otherwise as hvd_gpu_variational_refine_temporal of each plane
 * (a plane without neighbours gets the 2-D solve).  stage_iters: optional
 * HOST output, n * irls_outer ints (makes the call synchronous). */
int hvd_gpu_variational_refine_temporal_batch(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                              const hvd_gpu_neighbor* nbrs, const int* nn,
                                              const hvd_solver_params* p, float temporal_strength,
                                              float temporal_eps, int* stage_iters, void* stream);

/* A synthesised reference the joint solve is anchored to (hvd::AnchorTerm,
 * variational.h:60-66): luma, interleaved chroma and per-pixel weight, each
 * lines x width of its plane.  weight NULL: no anchor. */
typedef struct {
  const float* luma;
  const float* chroma;
  const float* weight;
} hvd_gpu_anchor;

/* VariationalRefineJoint (variational.cpp:558-891) for a batch: Y and chi
 * both free.  planes[i].y holds Y0 on entry and Y on return, planes[i].chi
 * chi0 and chi.  Neighbours as in hvd_gpu_variational_refine_temporal_batch
 * (strength temporal_strength, robust scale temporal_eps); anchors (NULL, or
 * n entries) weigh nr_anchor (|Y - Y_a|^2 + |chi - chi_a|^2).  stage_iters:
 * optional HOST output, n * irls_outer ints. */
int hvd_gpu_variational_refine_joint(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                     const hvd_gpu_neighbor* nbrs, const int* nn, const hvd_gpu_anchor* anchors,
                                     const hvd_solver_params* p, float temporal_strength, float temporal_eps,
                                     float nr_anchor, int* stage_iters, void* stream);


/* ---- the temporal neighbours' warps (temporal.h / temporal.cpp) -------- */
/* Motion fields are hvd-core's per-tile MotionField (dy, dx, confidence of
 * tiles_h x tiles_w tiles of `tile` pixels, as hvd_gpu_estimate_motion
 * returns them); per-pixel vector planes are out_h x out_w.  All planes are
 * device memory; the calls are asynchronous on `stream`. */

/* VectorsPerPixel (temporal.h:35-36, temporal.cpp:44-113): 3x3-median
 * outlier snap of the tile vectors, then bilinear interpolation between tile
 * centres. */
int hvd_gpu_vectors_per_pixel(hvd_gpu_ctx* ctx, const float* mdy, const float* mdx, int tiles_h, int tiles_w, int tile,
                              int out_h, int out_w, float* dy, float* dx, void* stream);

/* WarpByTiles / WarpByTilesComplex with precomputed per-pixel vectors
 * (temporal.h:45-50, :83; temporal.cpp:117-132, :205-215): out(y, x) =
 * a(clamp(y - round(dy)), clamp(x - round(dx))); channels 1 (real plane) or
 * 2 (interleaved complex).  out must not alias a. */
int hvd_gpu_warp_by_tiles(hvd_gpu_ctx* ctx, const float* a, int h, int w, int channels, const float* dy,
                          const float* dx, float* out, void* stream);

/* WarpBilinearTiles with precomputed vectors (temporal.h:59-61,
 * temporal.cpp:140-166): sub-pixel warp of a baseband src_h x src_w plane onto
 * out_h x out_w, with a constant extra vertical shift row_offset. */
int hvd_gpu_warp_bilinear(hvd_gpu_ctx* ctx, const float* a, int src_h, int src_w, const float* dy, const float* dx,
                          float row_offset, int out_h, int out_w, float* out, void* stream);

/* UpsampleConfidence (fast == 0: nearest tile, squared, box blur radius 8)
 * or UpsampleConfidenceFast (squared tile map through VectorsPerPixel)
 * (temporal.h:90-101, temporal.cpp:217-249). */
int hvd_gpu_upsample_confidence(hvd_gpu_ctx* ctx, const float* conf, int tiles_h, int tiles_w, int tile, int out_h,
                                int out_w, int fast, float* out, void* stream);

/* ComplexCoherence (temporal.h:68, temporal.cpp:169-196): |<z1 conj z2>| /
 * (sqrt(<|z1|^2><|z2|^2>) + 1e-6) with box averages of radius r (at most
 * 32), in [0, 1]; z1, z2 interleaved complex h x w. */
int hvd_gpu_complex_coherence(hvd_gpu_ctx* ctx, const float* z1, const float* z2, int h, int w, int r, float* out,
                              void* stream);

/* MotionCompensatePrev with a given motion field (temporal.h:139-142,
 * temporal.cpp:253-287): one per-pixel interpolation shared by the
 * integer-pel warps of the neighbour's composite and carrier (interleaved
 * complex), and the confidence upsample.  The results are a
 * hvd_gpu_neighbor's three planes. */
int hvd_gpu_motion_compensate_prev(hvd_gpu_ctx* ctx, const float* composite, const float* carrier, int h, int w,
                                   const float* mdy, const float* mdx, const float* mconf, int tiles_h, int tiles_w,
                                   int tile, int fast, float* out_composite, float* out_carrier, float* out_conf,
                                   void* stream);

/* ---- differentiation of the solve (no hvd-core counterpart) ------------ */

/* hvd_gpu_variational_refine that also saves chi at the start of each IRLS
 * stage: stages[i] holds irls_outer planes of lines * width * 2 floats, stage
 * major.  These, S and the carrier are all the adjoint needs (the stage
 * weights are recomputed from them rather than stored). */
int hvd_gpu_variational_refine_stages(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n,
                                      const hvd_solver_params* p, float* const* stages, void* stream);

/* The adjoint of VariationalRefine's 2-D solve for a loss L of its outputs.
 * Each IRLS stage is taken as solved exactly (implicit differentiation):
 * per stage an adjoint CG with the stage's own operator (adj_iterations
 * iterations at most, relative exit adj_tol), then the lambda, S and weight
 * terms; the weights' dependence on the stage's starting chi and on S is
 * followed back stage by stage.  planes[i]: the forward's S, carrier and, in
 * .chi, the final chi; p: the forward's parameters (chroma_aniso forced or
 * AUTO: the AUTO value is recomputed from chi0 and held constant); stages:
 * as saved by hvd_gpu_variational_refine_stages; gchi[i] = dL/dchi, gy[i] =
 * dL/dY (gy, or its entries, may be NULL).  Outputs: gs[i] = dL/dS (lines *
 * width), gchi0[i] = dL/dchi0 (lines * width * 2), and HOST glam[2 i],
 * glam[2 i + 1] = dL/dlambda_y, dL/dlambda_c.  Synchronous. */
int hvd_gpu_refine_adjoint(hvd_gpu_ctx* ctx, const hvd_gpu_plane* planes, int n, const hvd_solver_params* p,
                           const float* const* stages, const float* const* gchi, const float* const* gy,
                           int adj_iterations, double adj_tol, float* const* gs, float* const* gchi0, double* glam,
                           void* stream);

/* ---- field driver ------------------------------------------------------ */

/* The 2-D field path of DecodeFieldWindow (engine/sequence.cpp:229-281 and
 * :1046, output luma :1102-1114) for n raw fields of possibly different
 * standards and sizes, in one batch: uint16 codes -> IRE -> burst lock-in
 * -> active crop + carrier -> holographic init -> IRLS/CG solve.
 * raw[i]: field_height x field_width uint16 codes; y_out[i]: lines x width
 * floats; chi_out[i]: lines x width x 2 floats; theta_out[i] (optional, may be
 * NULL or hold NULLs): field_height floats.  `raw`, `geom`, `y_out`,
 * `chi_out`, `theta_out` are HOST arrays of n entries.  blanking / white are
 * the code levels of 0 and 100 IRE. */
int hvd_gpu_decode_fields_u16(hvd_gpu_ctx* ctx, const uint16_t* const* raw, int n, const hvd_field_geometry* geom,
                              float blanking, float white, const hvd_solver_params* p, float chroma_phase_deg,
                              float* const* y_out, float* const* chi_out, float* const* theta_out, void* stream);

/* Same, from fields already in IRE (FieldInput::samples, engine.h:48-51). */
int hvd_gpu_decode_fields_ire(hvd_gpu_ctx* ctx, const float* const* field_ire, int n, const hvd_field_geometry* geom,
                              const hvd_solver_params* p, float chroma_phase_deg, float* const* y_out,
                              float* const* chi_out, float* const* theta_out, void* stream);

/* DecodeFrameSequenceWindow (frame_bridge.cpp:722-947): frames[t]
 * (t < nframes) are field-sequential int16 10-bit frames (device); the call
 * splits them into fields (SplitFrameFields :61-99, PAL line offsets
 * FrameLineOffset frame_bridge.h:106-112), measures the window ACC gain
 * (:776-795) and field order (:759-765), decodes all 2 * nframes fields,
 * and weaves and requantises the core frames [core_begin, core_end) into
 * out[0 .. core_end - core_begin) (:867-945).  A 2-D window (temporal
 * strength < 0, or 0 with no more fields than the ambiguity stride, or
 * cg_iterations == 0) runs the field driver's batch; any other runs the
 * window driver (hvd_gpu_decode_window) on the fields' holographic inits
 * with HvdConfig's default window knobs (see the _ex variant).  `frames`,
 * `out` are HOST arrays; acc_gain (HOST, optional) receives chroma_gain x
 * ACC and makes the call synchronous. */
int hvd_gpu_decode_frame_window_s16(hvd_gpu_ctx* ctx, const int16_t* const* frames, int nframes, int core_begin,
                                    int core_end, const hvd_frame_params* fp, const hvd_solver_params* p,
                                    const hvd_bridge_params* b, const hvd_yc_frame* out, float* acc_gain,
                                    void* stream);

/* The same with the window driver's knobs (HvdConfig's 3-D fields; null:
 * their defaults) and its decisions (diag, HOST, optional: zeros for a 2-D
 * window).  The strength is b->temporal_strength; wp->temporal_strength is
 * not read. */
int hvd_gpu_decode_frame_window_ex_s16(hvd_gpu_ctx* ctx, const int16_t* const* frames, int nframes, int core_begin,
                                       int core_end, const hvd_frame_params* fp, const hvd_solver_params* p,
                                       const hvd_bridge_params* b, const hvd_window_params* wp,
                                       const hvd_yc_frame* out, float* acc_gain, hvd_window_diag* diag,
                                       void* stream);

/* ---- motion ------------------------------------------------------------ */

/* EstimateMotion (engine/motion.cpp:397-488) for n (ref, cur) pairs of
 * h x w float planes: box blur r = 2, 4x-decimated exhaustive coarse search
 * over +/-ceil(search / 4), +/-3 full-resolution refinement, parabolic
 * half-pel fit, prefer-zero margin rule, per-tile confidence.  Outputs are
 * ceil(h / tile) x ceil(w / tile) per pair.  Requires h, w >= 4.  The
 * vectors and confidences are produced in hvd-core's order of operations.
 * Pointer arrays are HOST arrays. */
int hvd_gpu_estimate_motion(hvd_gpu_ctx* ctx, const float* const* y_ref, const float* const* y_cur, int n, int h,
                            int w, int tile, int search, float* const* dy, float* const* dx, float* const* conf,
                            void* stream);

/* Integer-SAD block matching over uint16 planes (no hvd-core counterpart:
 * the integer form of BmPass, engine/motion.cpp:105-260).  For each of n
 * (ref, cur) pairs of h x w planes: bw x bh blocks (edge blocks clipped),
 * candidates |dy| <= sy, |dx| <= sx comparing cur(y, x) with the
 * edge-clamped ref(y - dy, x - dx); winner = least (SAD, |dy| + |dx|,
 * raster order).  Outputs ceil(h / bh) x ceil(w / bw) per pair.  Limits:
 * bw <= 32, bh <= 16, sx <= 32, sy <= 16.  Pointer arrays are HOST arrays. */
int hvd_gpu_block_match_sad_u16(hvd_gpu_ctx* ctx, const uint16_t* const* ref, const uint16_t* const* cur, int n,
                                int h, int w, int bh, int bw, int sy, int sx, int16_t* const* vy, int16_t* const* vx,
                                int32_t* const* sad, void* stream);

/* ---- frame bridge and component pieces for the C++ host API ------------- */

/* SplitFrameFields (frame_bridge.cpp:61-99; FrameLineOffset,
 * frame_bridge.h:106-112): the field-sequential int16 frame (device,
 * FrameLineOffset(fp, frame_height) samples) -> block 1 (flat lines
 * [0, field1_lines)) and block 2 (the rest) as field1_lines x frame_width IRE
 * planes, ire = (code - zero_code) / ((white - blanking) / 100); rows past a
 * block's end read zero_code.  zero_code = blanking for composite, chroma_dc
 * for a separate chroma channel (DecodeYcFrameBuffer, :276-297). */
int hvd_gpu_split_frame_fields(hvd_gpu_ctx* ctx, const int16_t* frame, const hvd_frame_params* fp, float zero_code,
                               float* field1, float* field2, void* stream);

/* The woven active picture's size (ReorderToWoven,
 * orc-plugin_hvd/src/hvd_chroma_decoder_stage.cpp:316-413): the active
 * window (or, full_raster != 0, the whole frame) in raster order, or with
 * field 0 / 1 that field's lines only (any other value: both fields
 * woven).  Host only; width = height = 0 when the picture is empty. */
int hvd_gpu_woven_size(const hvd_frame_params* fp, int full_raster, int field, int* width, int* height);

/* ReorderToWoven of a field-sequential hvd::YcFrameS16 (its planes
 * frame_width x frame_height, DEVICE pointers; chroma unused) into width x
 * height doubles y, u, v (hvd_chroma_decoder_stage.cpp:316-413), and its
 * RGB24 rendering rgb24 (width x height x 3 bytes, R G B): YuvToRgb8 (:79-88)
 * on Y and U/V normalised over the luma excursion black..white
 * (WovenToPreviewImage / WriteWovenAsRgb24, :500-638), clamped to [0, 1],
 * x 255 rounded.  Any output may be NULL.  The RGB rendering needs white
 * above black and blanking (the stage returns no image otherwise). */
int hvd_gpu_woven_picture(hvd_gpu_ctx* ctx, const hvd_yc_frame* yc, const hvd_frame_params* fp, int full_raster,
                          int field, double* y, double* u, double* v, uint8_t* rgb24, void* stream);

/* DetectFieldParity (frame_bridge.cpp:102-149) on two field1_lines x
 * frame_width IRE blocks: *vote (HOST) = +1 field 1 top, -1 field 1 bottom,
 * 0 no confident vote.  Synchronous. */
int hvd_gpu_field_parity_vote(hvd_gpu_ctx* ctx, const float* block1, const float* block2, const hvd_frame_params* fp,
                              int* vote, void* stream);

/* MakeCarrier / MakeCarrierPal (engine/ntsc_geometry.cpp:29-47, :75-93): an
 * explicit lines x active_width interleaved carrier plane from per-line
 * theta (and vswitch, NULL for NTSC). */
int hvd_gpu_make_carrier(hvd_gpu_ctx* ctx, const float* theta, const int8_t* vswitch, int lines,
                         const hvd_field_geometry* g, float* carrier, void* stream);

/* TridiagSmooth (engine/lockin.cpp:65-99): x solving
 * (diag(a) + lam L) x = a d with Neumann ends; n floats each. */
int hvd_gpu_tridiag_smooth(hvd_gpu_ctx* ctx, const float* d, const float* a, int n, float lam, float* out,
                           void* stream);

/* The forward differences of engine/gradients.h with Neumann boundaries and
 * their exact adjoints, d = op(a) on an h x w plane: real (is_complex 0) or
 * interleaved complex.  op: HVD_DIFF_* below.  weight (h x w floats, may be
 * NULL; Dx, Dy, D1, D2 only) multiplies the difference: the *WeightedInto
 * forms (gradients.h:290-381).  a, weight, d are DEVICE pointers; d must not
 * alias a. */
enum {
  HVD_DIFF_DX = 0,  /* Dx / DxInto (gradients.h:26-42, :112-123) */
  HVD_DIFF_DY = 1,  /* Dy / DyInto (:44-63, :125-138) */
  HVD_DIFF_DXT = 2, /* DxT / DxTInto (:65-83, :140-159) */
  HVD_DIFF_DYT = 3, /* DyT / DyTInto (:85-104, :161-182) */
  HVD_DIFF_D1 = 4,  /* D1Into (:196-214) */
  HVD_DIFF_D1T = 5, /* D1TInto (:216-236) */
  HVD_DIFF_D2 = 6,  /* D2Into (:238-256) */
  HVD_DIFF_D2T = 7  /* D2TInto (:258-278) */
};
int hvd_gpu_difference(hvd_gpu_ctx* ctx, int op, int is_complex, const float* a, const float* weight, int h, int w,
                       float* d, void* stream);

/* EstimateNoiseIre (ntsc_geometry.h:278-283, ntsc_geometry.cpp:49-71) of an
 * h x w composite plane (DEVICE): the stride-4 second differences' median
 * absolute deviation, as sigma in IRE, into *sigma (HOST).  Synchronous. */
int hvd_gpu_estimate_noise_ire(hvd_gpu_ctx* ctx, const float* s, int h, int w, float* sigma, void* stream);

/* Fft2d::Forward / Inverse (fft2d.h, fft2d.cpp:88-118): the numpy-convention
 * 2-D DFT of an h x w interleaved complex plane (DEVICE), forward with sign
 * -1 unnormalised, inverse with sign +1 divided by h w, into out (may equal
 * in).  h <= 1024, w <= 4096. */
int hvd_gpu_fft2d(hvd_gpu_ctx* ctx, const float* in, int h, int w, int inverse, float* out, void* stream);

/* BoxBlur (engine/motion.cpp:21-55): separable (2r + 1)-box mean with
 * edge-truncated windows and FP64 running sums, h x w floats, r <= 32. */
int hvd_gpu_box_blur(hvd_gpu_ctx* ctx, const float* in, int h, int w, int r, float* out, void* stream);

/* The chroma plane of hvd::FrameYc: mode 0 C = S - Y (DecodeFrame,
 * engine.cpp:181-183), mode 1 C = Re[chi c] with an explicit carrier
 * (DecodeChromaOnly, :221-224); n pixels. */
int hvd_gpu_chroma_plane(hvd_gpu_ctx* ctx, int mode, const float* s, const float* y, const float* chi,
                         const float* carrier, size_t n, float* out, void* stream);

/* The re-weave of a woven-frame decode into a field-sequential YcFrameS16
 * (DecodeFrameBuffer, frame_bridge.cpp:186-254; DecodeYcFrameBuffer
 * :309-377): every line starts as `src` read through FrameLineOffset (the
 * composite, or the separate luma channel); in the active picture woven row r
 * goes to flat block r % 2, field line first_active + r / 2, with luma
 * ClampU10(blanking + y scale) when `y` is non-NULL (else left as passed
 * through), chroma ClampS10(chroma scale) and v / u = gain (Re chi, -Im chi)
 * scale.  y, chroma: rows x width; chi: rows x width x 2. */
int hvd_gpu_weave_frame_s16(hvd_gpu_ctx* ctx, const int16_t* src, const hvd_frame_params* fp, int rows, int width,
                            const float* y, const float* chroma, const float* chi, float gain, const hvd_yc_frame* out,
                            void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HVD_GPU_H_ */
