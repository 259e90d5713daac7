// ref_shim.cpp -- extern "C" entry points over hvd-core's own engine code.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// FFT-free hvd-core translation units that sit in the reference checkout
// (engine/ntsc_geometry.cpp, lockin.cpp, variational.cpp, motion.cpp,
// colour.cpp) into oracle/_ref/libhvdref.so.  Nothing from the reference is
// copied here: this file only marshals plain pointers into hvd-core's own
// Plane / FieldGeometry / HvdConfig types and calls its functions, so the
// oracle and the GPU path can be compared with the real engine.
//
// engine/fft2d.cpp (and with it HolographicInit, DecodeFrame and
// DecodeFieldWindow) needs FFTW3, which this image does not have; those
// functions are not exported here.  engine/sequence.cpp is compiled for its
// test seam DecodeFieldWindowWithInits (the whole window driver with the
// inits supplied), which never reaches HolographicInit: that one symbol is
// left unresolved in the library and nothing here calls it.

#include <cstdint>
#include <cstring>
#include <vector>

#include "engine/colour.h"
#include "engine/gradients.h"
#include "engine/lockin.h"
#include "engine/motion.h"
#include "engine/temporal.h"
#include "engine/ntsc_geometry.h"
#include "engine/plane.h"
#include "engine/sequence.h"
#include "engine/variational.h"

namespace {

struct ShimGeom {
  int field_width, field_height, active_video_start, active_video_end;
  int colour_burst_start, colour_burst_end, first_active_field_line, last_active_field_line;
  double sample_rate, subcarrier_hz;
  int standard;
};

struct ShimCfg {
  double lambda_c, chroma_aniso, charbonnier_eps, chroma_eps, structure_coupling, diag_prior;
  int cg_iterations, irls_outer;
  double cg_tol;
  int fast, is_pal;
  double lambda_y;
};

// The HvdConfig fields the window driver reads beyond the solver's.
struct ShimSeqCfg {
  ShimCfg base;
  double temporal_strength, temporal_eps, nr_anchor, nr_eps, coherence_gate, odd_gate_floor;
  int passes, nr_radius, mc_tile, mc_search;
  int bidirectional, extended_temporal, trajectory_fit, psi_init, output_fidelity, parallel_across_fields;
};

hvd::HvdConfig ToCfg(const ShimCfg* sc) {
  hvd::HvdConfig cfg;
  cfg.lambda_c = static_cast<float>(sc->lambda_c);
  cfg.chroma_aniso = static_cast<float>(sc->chroma_aniso);
  cfg.charbonnier_eps = static_cast<float>(sc->charbonnier_eps);
  cfg.chroma_eps = static_cast<float>(sc->chroma_eps);
  cfg.structure_coupling = static_cast<float>(sc->structure_coupling);
  cfg.diag_prior = static_cast<float>(sc->diag_prior);
  cfg.cg_iterations = sc->cg_iterations;
  cfg.irls_outer = sc->irls_outer;
  cfg.cg_tol = static_cast<float>(sc->cg_tol);
  cfg.fast = sc->fast != 0;
  cfg.is_pal = sc->is_pal != 0;
  return cfg;
}

hvd::HvdConfig ToSeqCfg(const ShimSeqCfg* q) {
  hvd::HvdConfig cfg = ToCfg(&q->base);
  cfg.temporal_strength = static_cast<float>(q->temporal_strength);
  cfg.temporal_eps = static_cast<float>(q->temporal_eps);
  cfg.nr_anchor = static_cast<float>(q->nr_anchor);
  cfg.nr_eps = static_cast<float>(q->nr_eps);
  cfg.coherence_gate = static_cast<float>(q->coherence_gate);
  cfg.odd_gate_floor = static_cast<float>(q->odd_gate_floor);
  cfg.passes = q->passes;
  cfg.nr_radius = q->nr_radius;
  cfg.mc_tile = q->mc_tile;
  cfg.mc_search = q->mc_search;
  cfg.bidirectional = q->bidirectional != 0;
  cfg.extended_temporal = q->extended_temporal != 0;
  cfg.trajectory_fit = q->trajectory_fit != 0;
  cfg.psi_init = q->psi_init != 0;
  cfg.output_fidelity = q->output_fidelity != 0;
  cfg.parallel_across_fields = q->parallel_across_fields != 0;
  return cfg;
}

hvd::FieldGeometry ToGeom(const ShimGeom* s) {
  hvd::FieldGeometry g;
  g.field_width = s->field_width;
  g.field_height = s->field_height;
  g.active_video_start = s->active_video_start;
  g.active_video_end = s->active_video_end;
  g.colour_burst_start = s->colour_burst_start;
  g.colour_burst_end = s->colour_burst_end;
  g.first_active_field_line = s->first_active_field_line;
  g.last_active_field_line = s->last_active_field_line;
  g.sample_rate = s->sample_rate;
  g.subcarrier_hz = s->subcarrier_hz;
  g.standard = s->standard == 1   ? hvd::VideoStandard::kPal
               : s->standard == 2 ? hvd::VideoStandard::kPalM
                                  : hvd::VideoStandard::kNtsc;
  return g;
}

hvd::Plane ToPlane(const float* p, int h, int w) {
  hvd::Plane out(h, w);
  std::memcpy(out.data(), p, sizeof(float) * static_cast<size_t>(h) * w);
  return out;
}

hvd::ComplexPlane ToCPlane(const float* p, int h, int w) {
  hvd::ComplexPlane out(h, w);
  std::memcpy(out.data(), p, sizeof(float) * 2 * static_cast<size_t>(h) * w);
  return out;
}

}  // namespace

extern "C" {

void ref_burst_lockin_phase(const float* field, const ShimGeom* sg, float* theta) {
  const hvd::FieldGeometry g = ToGeom(sg);
  const std::vector<float> t = hvd::BurstLockinPhase(ToPlane(field, g.field_height, g.field_width), g);
  std::memcpy(theta, t.data(), sizeof(float) * t.size());
}

void ref_burst_lockin_phase_pal(const float* field, const ShimGeom* sg, float* theta, int8_t* vswitch) {
  const hvd::FieldGeometry g = ToGeom(sg);
  const hvd::PalBurstLockin l = hvd::BurstLockinPhasePal(ToPlane(field, g.field_height, g.field_width), g);
  std::memcpy(theta, l.theta.data(), sizeof(float) * l.theta.size());
  std::memcpy(vswitch, l.vswitch.data(), l.vswitch.size());
}

float ref_burst_amplitude_ire(const float* field, const ShimGeom* sg) {
  const hvd::FieldGeometry g = ToGeom(sg);
  return hvd::BurstAmplitudeIre(ToPlane(field, g.field_height, g.field_width), g);
}

float ref_acc_gain(float measured, float nominal) { return hvd::AccGain(measured, nominal); }

void ref_tridiag_smooth(const float* d, const float* a, int n, float lam, float* x) {
  const std::vector<float> r = hvd::TridiagSmooth(std::vector<float>(d, d + n), std::vector<float>(a, a + n), lam);
  std::memcpy(x, r.data(), sizeof(float) * r.size());
}

void ref_make_carrier(const float* theta, const int8_t* vswitch, int lines, const ShimGeom* sg, float* carrier) {
  const hvd::FieldGeometry g = ToGeom(sg);
  const std::vector<float> t(theta, theta + lines);
  hvd::ComplexPlane c = vswitch ? hvd::MakeCarrierPal(t, std::vector<int8_t>(vswitch, vswitch + lines), g)
                                : hvd::MakeCarrier(t, g);
  std::memcpy(carrier, c.data(), sizeof(float) * 2 * c.size());
}

// op: 0 Dx, 1 Dy, 2 DxT, 3 DyT over a real plane.
void ref_real_op(int op, const float* a, int h, int w, float* d) {
  const hvd::Plane p = ToPlane(a, h, w);
  hvd::Plane r = op == 0 ? hvd::Dx(p) : op == 1 ? hvd::Dy(p) : op == 2 ? hvd::DxT(p) : hvd::DyT(p);
  std::memcpy(d, r.data(), sizeof(float) * r.size());
}

// op: 0 Dx, 1 Dy, 2 DxT, 3 DyT, 4 D1, 5 D1T, 6 D2, 7 D2T over a complex plane.
void ref_cplx_op(int op, const float* a, int h, int w, float* d) {
  const hvd::ComplexPlane p = ToCPlane(a, h, w);
  hvd::ComplexPlane r(h, w);
  switch (op) {
    case 0: hvd::DxInto(p, r); break;
    case 1: hvd::DyInto(p, r); break;
    case 2: hvd::DxTInto(p, r); break;
    case 3: hvd::DyTInto(p, r); break;
    case 4: hvd::D1Into(p, r); break;
    case 5: hvd::D1TInto(p, r); break;
    case 6: hvd::D2Into(p, r); break;
    default: hvd::D2TInto(p, r); break;
  }
  std::memcpy(d, r.data(), sizeof(float) * 2 * r.size());
}

// VariationalRefine, 2-D (no neighbours).  chi: chi0 in, chi out.
void ref_variational_refine(const float* s, const float* carrier, int h, int w, const ShimCfg* sc, float* chi,
                            float* y) {
  hvd::HvdConfig cfg;
  cfg.lambda_c = static_cast<float>(sc->lambda_c);
  cfg.chroma_aniso = static_cast<float>(sc->chroma_aniso);
  cfg.charbonnier_eps = static_cast<float>(sc->charbonnier_eps);
  cfg.chroma_eps = static_cast<float>(sc->chroma_eps);
  cfg.structure_coupling = static_cast<float>(sc->structure_coupling);
  cfg.diag_prior = static_cast<float>(sc->diag_prior);
  cfg.cg_iterations = sc->cg_iterations;
  cfg.irls_outer = sc->irls_outer;
  cfg.cg_tol = static_cast<float>(sc->cg_tol);
  cfg.fast = sc->fast != 0;
  cfg.is_pal = sc->is_pal != 0;
  const hvd::RefineResult r =
      hvd::VariationalRefine(ToPlane(s, h, w), ToCPlane(carrier, h, w), ToCPlane(chi, h, w), cfg);
  std::memcpy(chi, r.chroma.data(), sizeof(float) * 2 * r.chroma.size());
  std::memcpy(y, r.luma.data(), sizeof(float) * r.luma.size());
}

// hvd-core's VariationalRefine with nn temporal neighbours (NeighborTerm:
// composite sw, carrier cw, confidence conf; nn planes each) at
// temporal_strength nu and temporal_eps eps_t.
void ref_variational_refine_nbr(const float* s, const float* carrier, int h, int w, const ShimCfg* sc, int nn,
                                const float* sw, const float* cw, const float* conf, float nu, float eps_t,
                                float* chi, float* y) {
  hvd::HvdConfig cfg;
  cfg.lambda_c = static_cast<float>(sc->lambda_c);
  cfg.chroma_aniso = static_cast<float>(sc->chroma_aniso);
  cfg.charbonnier_eps = static_cast<float>(sc->charbonnier_eps);
  cfg.chroma_eps = static_cast<float>(sc->chroma_eps);
  cfg.structure_coupling = static_cast<float>(sc->structure_coupling);
  cfg.diag_prior = static_cast<float>(sc->diag_prior);
  cfg.cg_iterations = sc->cg_iterations;
  cfg.irls_outer = sc->irls_outer;
  cfg.cg_tol = static_cast<float>(sc->cg_tol);
  cfg.fast = sc->fast != 0;
  cfg.is_pal = sc->is_pal != 0;
  cfg.temporal_strength = nu;
  cfg.temporal_eps = eps_t;
  std::vector<hvd::NeighborTerm> nb(nn);
  const size_t np = static_cast<size_t>(h) * w;
  for (int j = 0; j < nn; ++j) {
    nb[j].composite = ToPlane(sw + j * np, h, w);
    nb[j].carrier = ToCPlane(cw + j * 2 * np, h, w);
    nb[j].confidence = ToPlane(conf + j * np, h, w);
  }
  const hvd::RefineResult r =
      hvd::VariationalRefine(ToPlane(s, h, w), ToCPlane(carrier, h, w), ToCPlane(chi, h, w), cfg, nb);
  std::memcpy(chi, r.chroma.data(), sizeof(float) * 2 * r.chroma.size());
  std::memcpy(y, r.luma.data(), sizeof(float) * r.luma.size());
}

void ref_box_blur(float* a, int h, int w, int r) {
  hvd::Plane p = hvd::BoxBlur(ToPlane(a, h, w), r);
  std::memcpy(a, p.data(), sizeof(float) * p.size());
}

void ref_estimate_motion(const float* y_ref, const float* y_cur, int h, int w, int tile, int search, float* dy,
                         float* dx, float* conf) {
  const hvd::MotionField m = hvd::EstimateMotion(ToPlane(y_ref, h, w), ToPlane(y_cur, h, w), tile, search);
  std::memcpy(dy, m.dy.data(), sizeof(float) * m.dy.size());
  std::memcpy(dx, m.dx.data(), sizeof(float) * m.dx.size());
  std::memcpy(conf, m.confidence.data(), sizeof(float) * m.confidence.size());
}

static hvd::PerPixelMotion ref_vpp(const float* mdy, const float* mdx, int th, int tw, int tile, int h, int w) {
  return hvd::VectorsPerPixel(ToPlane(mdy, th, tw), ToPlane(mdx, th, tw), tile, h, w);
}

void ref_vectors_per_pixel(const float* mdy, const float* mdx, int th, int tw, int tile, int h, int w, float* dy,
                           float* dx) {
  const hvd::PerPixelMotion v = ref_vpp(mdy, mdx, th, tw, tile, h, w);
  std::memcpy(dy, v.dy.data(), sizeof(float) * v.dy.size());
  std::memcpy(dx, v.dx.data(), sizeof(float) * v.dx.size());
}

void ref_warp_by_tiles(const float* a, int h, int w, int ch, const float* mdy, const float* mdx, int th, int tw,
                       int tile, float* out) {
  if (ch == 2) {
    const hvd::ComplexPlane o = hvd::WarpByTilesComplex(ToCPlane(a, h, w), ToPlane(mdy, th, tw),
                                                        ToPlane(mdx, th, tw), tile);
    std::memcpy(out, o.data(), sizeof(float) * 2 * o.size());
  } else {
    const hvd::Plane o = hvd::WarpByTiles(ToPlane(a, h, w), ToPlane(mdy, th, tw), ToPlane(mdx, th, tw), tile);
    std::memcpy(out, o.data(), sizeof(float) * o.size());
  }
}

void ref_warp_bilinear(const float* a, int hs, int ws, const float* mdy, const float* mdx, int th, int tw, int tile,
                       float row_offset, int out_h, int out_w, float* out) {
  const hvd::Plane o = hvd::WarpBilinearTiles(ToPlane(a, hs, ws), ToPlane(mdy, th, tw), ToPlane(mdx, th, tw), tile,
                                              row_offset, out_h, out_w);
  std::memcpy(out, o.data(), sizeof(float) * o.size());
}

void ref_upsample_confidence(const float* conf, int th, int tw, int tile, int h, int w, int fast, float* out) {
  const hvd::Plane o = fast ? hvd::UpsampleConfidenceFast(ToPlane(conf, th, tw), tile, h, w)
                            : hvd::UpsampleConfidence(ToPlane(conf, th, tw), tile, h, w);
  std::memcpy(out, o.data(), sizeof(float) * o.size());
}

void ref_complex_coherence(const float* z1, const float* z2, int h, int w, int r, float* out) {
  const hvd::Plane o = hvd::ComplexCoherence(ToCPlane(z1, h, w), ToCPlane(z2, h, w), r);
  std::memcpy(out, o.data(), sizeof(float) * o.size());
}

void ref_motion_compensate_prev(const float* comp, const float* car, int h, int w, const float* mdy,
                                const float* mdx, const float* mconf, int th, int tw, int tile, int fast,
                                float* out_comp, float* out_car, float* out_conf) {
  hvd::NeighborRawState prev;
  prev.luma = hvd::Plane(h, w);
  prev.composite = ToPlane(comp, h, w);
  prev.carrier = ToCPlane(car, h, w);
  hvd::MotionField m;
  m.dy = ToPlane(mdy, th, tw);
  m.dx = ToPlane(mdx, th, tw);
  m.confidence = ToPlane(mconf, th, tw);
  const hvd::MotionCompensatedResult r =
      hvd::MotionCompensatePrev(prev, hvd::Plane(h, w), tile, 8, &m, fast != 0, nullptr);
  std::memcpy(out_comp, r.composite.data(), sizeof(float) * r.composite.size());
  std::memcpy(out_car, r.carrier.data(), sizeof(float) * 2 * r.carrier.size());
  std::memcpy(out_conf, r.confidence.data(), sizeof(float) * r.confidence.size());
}

float ref_estimate_noise_ire(const float* s, int h, int w) { return hvd::EstimateNoiseIre(ToPlane(s, h, w)); }

void ref_verify_motion(const float* y_ref, const float* y_cur, int h, int w, const float* pdy, const float* pdx,
                       int th, int tw, int tile, float* dy, float* dx, float* conf) {
  const hvd::MotionField m = hvd::VerifyMotion(ToPlane(y_ref, h, w), ToPlane(y_cur, h, w), ToPlane(pdy, th, tw),
                                               ToPlane(pdx, th, tw), tile);
  std::memcpy(dy, m.dy.data(), sizeof(float) * m.dy.size());
  std::memcpy(dx, m.dx.data(), sizeof(float) * m.dx.size());
  std::memcpy(conf, m.confidence.data(), sizeof(float) * m.confidence.size());
}

// TrajectorySnap over m measurements (offset, parity shift, th x tw dy, dx,
// confidence planes back to back); dy, dx in place.
void ref_trajectory_snap(int m, const int* offsets, const float* pshift, float* dy, float* dx, const float* conf,
                         int th, int tw) {
  std::vector<hvd::OffsetMotion> pm(m);
  const size_t nt = static_cast<size_t>(th) * tw;
  for (int j = 0; j < m; ++j) {
    pm[j].offset = offsets[j];
    pm[j].parity_shift = pshift[j];
    pm[j].field.dy = ToPlane(dy + j * nt, th, tw);
    pm[j].field.dx = ToPlane(dx + j * nt, th, tw);
    pm[j].field.confidence = ToPlane(conf + j * nt, th, tw);
  }
  hvd::TrajectorySnap(&pm);
  for (int j = 0; j < m; ++j) {
    std::memcpy(dy + j * nt, pm[j].field.dy.data(), sizeof(float) * nt);
    std::memcpy(dx + j * nt, pm[j].field.dx.data(), sizeof(float) * nt);
  }
}

// VariationalRefineJoint with nn neighbours and an optional anchor (a_*
// null: none).  y, chi: the start in, the result out.
void ref_joint_refine(const float* s, const float* carrier, int h, int w, const ShimCfg* sc, int nn,
                      const float* sw, const float* cw, const float* conf, float nu, float eps_t,
                      const float* a_luma, const float* a_chroma, const float* a_weight, float nu_a, float* y,
                      float* chi) {
  hvd::HvdConfig cfg = ToCfg(sc);
  cfg.temporal_strength = nu;
  cfg.temporal_eps = eps_t;
  cfg.nr_anchor = nu_a;
  std::vector<hvd::NeighborTerm> nb(nn);
  const size_t np = static_cast<size_t>(h) * w;
  for (int j = 0; j < nn; ++j) {
    nb[j].composite = ToPlane(sw + j * np, h, w);
    nb[j].carrier = ToCPlane(cw + j * 2 * np, h, w);
    nb[j].confidence = ToPlane(conf + j * np, h, w);
  }
  hvd::AnchorTerm anchor;
  if (a_weight) {
    anchor.luma = ToPlane(a_luma, h, w);
    anchor.chroma = ToCPlane(a_chroma, h, w);
    anchor.weight = ToPlane(a_weight, h, w);
  }
  const hvd::RefineResult r = hvd::VariationalRefineJoint(ToPlane(s, h, w), ToCPlane(carrier, h, w),
                                                          ToPlane(y, h, w), ToCPlane(chi, h, w), cfg, nb,
                                                          a_weight ? &anchor : nullptr);
  std::memcpy(chi, r.chroma.data(), sizeof(float) * 2 * r.chroma.size());
  std::memcpy(y, r.luma.data(), sizeof(float) * r.luma.size());
}

// DecodeFieldWindowWithInits over nf fields of h x w (s, carrier, chi0 and,
// when init_y is non-null, Y0 planes back to back; parity per field).
// diag receives resolved_strength, ambiguity_ire, sigma_ire, temporal_eps,
// nr_eps.  Returns the number of fields decoded.
int ref_decode_window_with_inits(int nf, int h, int w, const float* s, const float* carrier, const int* parity,
                                 const float* init_chi, const float* init_y, const ShimGeom* sg,
                                 const ShimSeqCfg* q, float* out_y, float* out_chi, float* diag) {
  const size_t np = static_cast<size_t>(h) * w;
  std::vector<hvd::FieldObs> fields(nf);
  std::vector<hvd::DecodedField> inits(nf);
  for (int j = 0; j < nf; ++j) {
    fields[j].s = ToPlane(s + j * np, h, w);
    fields[j].carrier = ToCPlane(carrier + j * 2 * np, h, w);
    fields[j].parity = parity[j];
    inits[j].chroma = ToCPlane(init_chi + j * 2 * np, h, w);
    if (init_y) inits[j].luma = ToPlane(init_y + j * np, h, w);
  }
  hvd::SequenceDiagnostics d;
  const std::vector<hvd::DecodedField> out =
      hvd::DecodeFieldWindowWithInits(fields, inits, ToGeom(sg), ToSeqCfg(q), &d);
  for (size_t j = 0; j < out.size(); ++j) {
    std::memcpy(out_y + j * np, out[j].luma.data(), sizeof(float) * np);
    std::memcpy(out_chi + j * 2 * np, out[j].chroma.data(), sizeof(float) * 2 * np);
  }
  if (diag) {
    diag[0] = d.resolved_strength;
    diag[1] = d.ambiguity_ire;
    diag[2] = d.sigma_ire;
    diag[3] = d.temporal_eps;
    diag[4] = d.nr_eps;
  }
  return static_cast<int>(out.size());
}

// DrizzleFrame of fields j0, j0 + 1 of n decoded fields (Y planes, chi
// planes one after another); out: (2 h scale) x w luma and chroma.
void ref_drizzle_frame(int j0, int n, int h, int w, const float* ys, const float* chis, const int* parity,
                       int mc_tile, int mc_search, int nr_radius, float nr_eps, int scale, float* out_y,
                       float* out_chi) {
  const size_t np = static_cast<size_t>(h) * w;
  std::vector<hvd::Plane> Y(n);
  std::vector<hvd::ComplexPlane> C(n);
  std::vector<int> par(parity, parity + n);
  for (int k = 0; k < n; ++k) {
    Y[k] = ToPlane(ys + k * np, h, w);
    C[k] = ToCPlane(chis + k * 2 * np, h, w);
  }
  hvd::HvdConfig cfg;
  cfg.mc_tile = mc_tile;
  cfg.mc_search = mc_search;
  cfg.nr_radius = nr_radius;
  cfg.nr_eps = nr_eps;
  const hvd::DrizzleResult r = hvd::DrizzleFrame(j0, Y, C, par, cfg, scale);
  std::memcpy(out_y, r.luma.data(), sizeof(float) * r.luma.size());
  std::memcpy(out_chi, r.chroma.data(), sizeof(float) * 2 * r.chroma.size());
}

}  // extern "C"
