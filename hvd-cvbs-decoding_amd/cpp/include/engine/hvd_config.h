// engine/hvd_config.h -- the decoder knobs of the hvd-core C++ API
// (hvd-core/src/engine/hvd_config.h:24-380), member for member with hvd-core's
// defaults so a configuration built for hvd-core carries over unchanged.
// The GPU path reads the solver, init, ACC / gain, phase and field-order
// members; the 3-D / temporal members select work this library does not
// run (such calls return empty results, see engine/gpu_status.h).
#ifndef HVD_AMD_ENGINE_HVD_CONFIG_H_
#define HVD_AMD_ENGINE_HVD_CONFIG_H_

#include <string>

namespace hvd {

struct HvdConfig {
  // variational solve
  float lambda_c = 1.0F;
  float chroma_aniso = 0.0F;  // 0 => AUTO
  bool is_pal = false;        // derived from the geometry by the entry points
  float charbonnier_eps = 0.5F;
  float chroma_eps = 1.0F;
  float structure_coupling = 0.25F;
  float diag_prior = 0.0F;
  int cg_iterations = 2;
  int irls_outer = 4;
  float cg_tol = 0.0F;        // 0 => auto (0.02 slow / 0.10 fast)
  bool fast = true;
  bool parallel_across_fields = true;

  // holographic init
  float init_lpf_h_mhz = 1.3F;
  float init_lpf_v_cph = 60.0F;
  bool symmetry_variant = false;

  // colour
  bool acc = true;
  float chroma_gain = 1.0F;
  bool monochrome = false;
  float chroma_phase_deg = 0.0F;
  bool custom_subcarrier = false;
  double subcarrier_hz = 2556818.2;

  // frame / field handling
  bool frame_decode = true;
  int fft_threads = 4;
  std::string debug_dir;
  int field_order = 0;  // 0 auto (measured), 1 field 1 top, 2 field 1 bottom

  // temporal / 3-D (not run by this library)
  bool enable_temporal = false;
  float temporal_strength = 0.0F;  // < 0 off, 0 auto, > 0 on
  bool bidirectional = true;
  int passes = 2;
  float temporal_eps = 0.0F;
  float nr_anchor = 1.0F;
  float nr_eps = 0.0F;
  int nr_radius = 2;
  bool drizzle = false;
  int mc_tile = 32;
  int mc_search = 16;
  bool trajectory_fit = true;
  bool extended_temporal = true;
  float coherence_gate = 0.6F;
  float odd_gate_floor = 0.35F;
  int chunk_frames = 6;
  int chunk_overlap = 1;
  bool selective_3d = false;
  float selective_max_area = 0.45F;
  bool output_fidelity = true;
  bool psi_init = false;
};

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_HVD_CONFIG_H_
