// solve.hip -- the IRLS + conjugate-gradient variational solve on gfx950.
//
// Restates hvd-core's VariationalRefine 2-D path (engine/variational.cpp:
// 198-556) as matrix-free stencil kernels over a batch of fields of uneven
// size.  Per CG iteration one stencil kernel runs, k_cg_iter: it forms the
// direction d <- -g + beta d and chi <- chi + alpha d on the fly, writes chi, d
// and the recomputed gradient g <- grad(chi) (:286-347), and leaves the FP64
// partial sums from which the next direction's curvature curv(d) (:352-399)
// and Re<g, d> (:72-117) follow by expansion in beta, so no second pass over
// the planes is needed.  A one-block-per-field kernel then sums that field's
// strip partials in a fixed order and advances its scalar CG state (alpha,
// beta, the exits of :530 and :547-548), so no host round trip happens inside
// the solve and converged fields drop out by themselves (per-field masks).
// chi, g and d are ping-ponged so a block never reads a neighbour's
// already-updated value.
#include "hvd_device.h"
#include "hvd_kernels.h"



namespace hvdk {

namespace {

// A wave (one 64-thread block) owns a strip of kStripCols output columns and
// `rows` output lines of one field.  Lane l sits on column
// x = x0 + l - kStripHalo, so lanes 0, 1 and 61 .. 63 are halo columns whose
// values reach their neighbours through lane shifts; each lane walks its rows
// top to bottom with the rows above and below held in registers.  Every
// per-pixel value is formed by the same expression wherever it is needed (by
// the neighbouring pixel, by the next strip, by the next pass), so a value
// shifted in from a neighbour lane is bit-identical to one recomputed there;
// the helpers that form them do not let the compiler fuse a multiply into an
// add, which it might do differently at different call sites.
// The field descriptor and CG state are copied into registers up front (they
// are restrict-qualified, so the copies are scalar loads that no store in the
// loop can force to be reloaded).
struct StripCtx {
  int f;        // field index
  int lt;       // strip index within the field
  int x;        // this lane's column (-1 or >= w on halo lanes)
  int y0;       // first output row
  bool valid;   // column inside the plane
  bool out;     // lane produces outputs (not a halo lane, inside the plane)
};

__device__ inline int lane_id() { return threadIdx.x & 63; }

__device__ inline StripCtx strip_ctx(const DevField* F, const StripGrid& sg) {
  StripCtx c;
  const int t = xcd_remap(blockIdx.x, sg.nb);
  // equal-sized fields: the owner is a division, not a dependent search
  c.f = sg.spf > 0 ? t / sg.spf : field_of_tile(sg.tp, sg.nf, t);
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  c.lt = sg.spf > 0 ? t - c.f * sg.spf : t - sg.tp[c.f];
  const int lane = lane_id();
  c.x = strip_col(c.lt, f.tiles_x, lane);
  c.y0 = strip_row0(c.lt, f.tiles_x, sg.rows);
  c.valid = c.x >= 0 && c.x < f.w;
  c.out = lane >= kStripHalo && lane < kStripHalo + kStripCols && c.x < f.w;
  return c;
}

// The carrier down this lane's column (ntsc_geometry.cpp:75-93).  The row
// phasors (and V-switch signs) of the strip's rows y0 - 2 ... y0 + rows + 2
// are loaded once, one per lane, and broadcast per row with a lane read; the
// column phasor is loaded once per lane.  Every lane of the wave must have
// run carrier_for before at() is used (the lane read does not look at
// whether the lane it reads is still active).
struct Carrier {
  float2 col;
  float2 rowl;          // row phasor of row ybase + lane
  int vsl;              // its V-switch sign (< 0: switched line)
  int x, ybase, w, h;
  const float2* plane;  // explicit carrier plane, or null
  // Valid for rows and columns inside the plane; elsewhere (halo rows and
  // lanes) the value is unspecified but the access stays in bounds.
  __device__ inline float2 at(int y) const { return plane ? load(y) : factored(y); }
  __device__ inline float2 load(int y) const {
    return gld(plane, (size_t)clamped_index(y, x, h, w));
  }
  __device__ inline float2 factored(int y) const {
#pragma clang fp contract(off)
    const int l = y - ybase;  // uniform across the wave
    const float2 rp = make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(rowl.x), l)),
                                  __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rowl.y), l)));
    // rp * col written out here (not fused into adds at the call site)
    float2 v = make_float2(rp.x * col.x - rp.y * col.y, rp.x * col.y + rp.y * col.x);
    if (__builtin_amdgcn_readlane(vsl, l) < 0) v.x = -v.x;
    return v;
  }
};

__device__ inline Carrier carrier_for(const DevField& f, const StripCtx& c) {
  Carrier k;
  k.x = c.x;
  k.w = f.w;
  k.h = f.h;
  k.ybase = c.y0 - kStripHalo;
  k.plane = f.carrier;
  k.col = make_float2(0.f, 0.f);
  k.rowl = make_float2(0.f, 0.f);
  k.vsl = 1;
  if (!f.carrier) {
    if (c.valid) k.col = gld(f.colph, c.x);
    const int yy = k.ybase + lane_id();
    if (yy >= 0 && yy < f.h) {
      k.rowl = gld(f.rowph, yy);
      if (f.vsw) k.vsl = gld(f.vsw, yy);
    }
  }
  return k;
}

// Element of a two-entry pointer table, selected without indexing (keeps the
// descriptor copy in registers).
template <typename T>
__device__ inline T* pick(T* const (&a)[2], int i) { return i ? a[1] : a[0]; }

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Neighbour exchange across the whole wave with DPP lane shifts (wave_shr:1,
// wave_shl:1; no LDS round trip): shl gives lane l the value of lane l - 1,
// shr the value of lane l + 1.  Lanes 0 (shl) and 63 (shr) get 0 and never
// use it.  Call with every lane active.
__device__ inline float shl(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ inline float shr(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ inline float2 shl(float2 v) { return make_float2(shl(v.x), shl(v.y)); }
__device__ inline float2 shr(float2 v) { return make_float2(shr(v.x), shr(v.y)); }

// chi + alpha d, written once so the centre pixel and its neighbours evaluate
// the identical expression (the value stored must equal the value used).
__device__ inline float2 axpy(float2 c, float alpha, float2 d) {
  return make_float2(__fmaf_rn(alpha, d.x, c.x), __fmaf_rn(alpha, d.y, c.y));
}
// -g + beta d (variational.cpp:541-543).
__device__ inline float2 dir_update(float2 g, float beta, float2 d) {
  return make_float2(__fmaf_rn(beta, d.x, -g.x), __fmaf_rn(beta, d.y, -g.y));
}

inline __device__ bool row_in(const DevField& f, int y, bool valid) { return valid && y >= 0 && y < f.h; }

// Loads of one row of this lane's column.
struct RawRow {
  float2 chi, d;
  float s;
};

__device__ inline RawRow raw_row(const DevField& f, const Carrier& k, const float2* chi, const float2* dd, int y,
                                 bool valid) {
  RawRow r{make_float2(0.f, 0.f), make_float2(0.f, 0.f), 0.f};
  if (row_in(f, y, valid)) {
    const size_t i = (size_t)y * f.w + k.x;
    r.chi = gld(chi, i);
    if (dd) r.d = gld(dd, i);
    r.s = f.s ? gld(f.s, i) : 0.f;
  }
  return r;
}

// S - Re[chi c] in one fixed operation order, so that every place forming a
// row's luma from the same chi (the pass that moves chi, the next pass,
// which recomputes its gradient, and the weights passes) gets the same bits.
__device__ inline float luma_res(float s, float2 chi, float2 car) {
#pragma clang fp contract(off)
  return s - __fmaf_rn(chi.x, car.x, -(chi.y * car.y));
}

struct Sample {
  float yc;    // S - Re[chi c]
  float2 chi;  // chi (+ alpha d)
  float2 car;  // carrier
};

// The sample of row y from its loads; zeros outside the plane.
__device__ inline Sample make_sample(const DevField& f, const Carrier& k, const RawRow& r, bool with_d, float alpha,
                                     int y, bool valid) {
  Sample s{0.f, make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
  if (row_in(f, y, valid)) {
    const float2 c = with_d ? axpy(r.chi, alpha, r.d) : r.chi;
    s.chi = c;
    s.car = k.at(y);
    s.yc = luma_res(r.s, c, s.car);
  }
  return s;
}

// g = -2 lam_y [DxT(wx Dx yc) + DyT(wy Dy yc)] conj(c)
//     + 2 [mu_h DxT(wcx Dx chi) + mu_v DyT(wcy Dy chi)]      (:286-335)
// from the weighted differences ("fluxes") across the pixel's four edges.
// Each edge's flux is formed once, by the pixel on its left / upper side,
// and reaches the other pixel by a lane shift or from the previous line.
// The Neumann boundaries of gradients.h (a missing neighbour contributes a
// zero difference) come from the weights: the stored weight of an edge that
// leaves the plane is zero (k_irls_weights), and the weights of rows and
// columns outside the plane load as zero, so every neighbour value only has
// to be finite.
struct Flux {
  float y;      // luma: w (yc_next - yc)
  float2 c;     // chroma: w (chi_next - chi)
};

__device__ inline Flux flux(float w_luma, float w_chroma, const Sample& a, const Sample& b) {
#pragma clang fp contract(off)
  Flux f;
  f.y = w_luma * (b.yc - a.yc);
  f.c = make_float2(w_chroma * (b.chi.x - a.chi.x), w_chroma * (b.chi.y - a.chi.y));
  return f;
}

// The gradient at a pixel from the fluxes through its left (xl), right (x0),
// upper (yu) and lower (y0) edges; c0 is the pixel's carrier.
__device__ inline float2 grad_from_flux(const FieldCG& st, float2 c0, const Flux& xl, const Flux& x0, const Flux& yu,
                                       const Flux& y0) {
#pragma clang fp contract(off)
  const float img = 2.f * st.lam_y * ((xl.y - x0.y) + (yu.y - y0.y));
  const float2 dxt = make_float2(xl.c.x - x0.c.x, xl.c.y - x0.c.y);
  const float2 dyt = make_float2(yu.c.x - y0.c.x, yu.c.y - y0.c.y);
  const float2 cp = make_float2(2.f * (st.mu_h * dxt.x + st.mu_v * dyt.x), 2.f * (st.mu_h * dxt.y + st.mu_v * dyt.y));
  return make_float2(-img * c0.x + cp.x, img * c0.y + cp.y);
}

// Wave-level FP64 partial of this strip, published to its slot.  The
// per-field totals are formed by the next launch (a kernel boundary orders
// the two; no device-scope fence per wave, which on a multi-XCD part would
// write back the L2 each time).
__device__ inline void publish(double* partials, int tile0, int lt, int nv, int k, double v) {
  v = wave_sum(v);
  if (lane_id() == 0) partials[(size_t)(tile0 + lt) * nv + k] = v;
}

// The four per-field sums of a launch's strip partials (strips t0 .. t0+n-1),
// in one fixed order: lane l adds strips l, l + 64, ..., then a shuffle tree.
// Every wave that forms them gets the same bits, in every lane.
struct CgSums {
  double v[kCgPartials];
};

__device__ inline CgSums field_sums(const double* partials, int t0, int n) {
  CgSums r;
#pragma unroll
  for (int k = 0; k < kCgPartials; ++k) r.v[k] = 0.0;
  for (int j = lane_id(); j < n; j += 64) {
    const double* p = partials + (size_t)(t0 + j) * kCgPartials;
#pragma unroll
    for (int k = 0; k < kCgPartials; ++k) r.v[k] += p[k];
  }
#pragma unroll
  for (int k = 0; k < kCgPartials; ++k) r.v[k] = wave_sum(r.v[k]);
  return r;
}

// Scalars of the first iteration from the stage-start sums (variational.cpp:
// 524-531): gg0 = |g|^2, d = -g so curv(d) = curv(g) and Re<g, d> = -|g|^2.
__device__ inline CgDyn cg_first(const FieldCG& s, int cur, const CgSums& q) {
  CgDyn n{};
  n.gg = q.v[0];
  n.gg0 = q.v[0];
  n.hh = q.v[1];
  n.gd = -q.v[0];
  n.cur = cur;
  n.done = s.n_inner <= 0 || n.hh <= s.hh_floor ? 1 : 0;
  if (!n.done) n.alpha = (float)(-0.5 * n.gd / n.hh);
  return n;
}

// End of an iteration (variational.cpp:536-548, then :528-531 of the next)
// from the snapshot before it and the sums of its launch: beta, the exits on
// |g|^2 and the iteration budget, the next direction's
//   curv(-g' + beta d) = curv(g') - 2 beta curv(g', d) + beta^2 curv(d)
// (exit when <= hh_floor), Re<g', d'> = -|g'|^2 + beta Re<g', d> and alpha;
// flips the ping-pong slot.
__device__ inline CgDyn cg_advance(const FieldCG& s, const CgDyn& prev, const CgSums& q) {
  CgDyn n = prev;
  if (prev.done) return n;
  const double ggn = q.v[0];
  n.cur ^= 1;
  const float beta = (float)(ggn / (prev.gg > 1e-30 ? prev.gg : 1e-30));
  n.gg = ggn;
  n.it += 1;
  if (ggn < s.tol2 * prev.gg0 || ggn < s.gg_floor * s.npix || n.it >= s.n_inner) {
    n.done = 1;
    return n;
  }
  const double b = (double)beta;
  n.hh = q.v[1] - 2.0 * b * q.v[2] + b * b * prev.hh;
  if (n.hh <= s.hh_floor) {
    n.done = 1;
    return n;
  }
  n.beta = beta;
  n.gd = -ggn + b * q.v[3];
  n.alpha = (float)(-0.5 * n.gd / n.hh);
  return n;
}

// One row of this lane's column as the fused CG iteration forms it: the
// direction d(it) = -g + beta d(it-1) (or -g on the first iteration), the moved
// chroma chi' = chi + alpha d(it), its luma S - Re[chi' c] and the carrier.
// In the stage-start pass (Start) chi' = chi and there is no direction.
struct IterRow {
  Sample p;
  float2 d;
  float dY;     // -Re[d c]: the luma the direction moves
  float2 rhs;   // adjoint solve: the right-hand side at this pixel
};

// Loads of one row of this lane's column, issued a line ahead of their use
// without branches at clamped addresses, so rows and columns outside the
// plane read finite values from its edge (the zero weights there cancel
// them): chi, the composite (none in the adjoint solve, whose S is 0), the
// previous direction (Upd), the weights unmasked, an explicit carrier plane's
// value and the adjoint's right-hand side.
struct IterRaw {
  float2 chi, dp;
  float2 car;
  float2 rhs;
  float4 w;
  float s;
};

template <bool Start, bool Upd, bool Adj, bool FuseW>
__device__ inline IterRaw iter_fetch(const DevField& f, const Carrier& k, int x, const float2* chi,
                                     const float2* dprev, int y) {
  IterRaw r;
  const unsigned i = clamped_index(y, x, f.h, f.w);
  r.chi = gld32(chi, i);
  r.s = Adj ? 0.f : gld32(f.s, i);
  r.dp = Upd ? gld32(dprev, i) : make_float2(0.f, 0.f);
  r.car = k.plane ? gld32(k.plane, i) : make_float2(0.f, 0.f);
  r.rhs = Adj ? gld32(f.rhs, i) : make_float2(0.f, 0.f);
  r.w = FuseW ? make_float4(0.f, 0.f, 0.f, 0.f) : gld32(f.wts, i);
  return r;
}

__device__ inline float4 weights_mask(const DevField& f, float4 v, int y, bool valid) {
  return row_in(f, y, valid) ? v : make_float4(0.f, 0.f, 0.f, 0.f);
}

// A row before this iteration's move: its sample (the input of the gradient
// the previous pass formed), masked weights and what the move needs.
struct OldRow {
  Sample p;
  float4 w;
  float2 dp;
  float2 rhs;
  float s;
};

__device__ inline OldRow old_row(const DevField& f, const Carrier& k, const IterRaw& q, int y, bool valid) {
  OldRow o;
  o.p.car = k.plane ? q.car : k.factored(y);
  o.p.chi = q.chi;
  o.p.yc = luma_res(q.s, q.chi, o.p.car);
  o.w = weights_mask(f, q.w, y, valid);
  o.dp = q.dp;
  o.rhs = q.rhs;
  o.s = q.s;
  return o;
}

// The row of the stage-start pass: chi' = chi, no direction.
__device__ inline IterRow start_row(const Carrier& k, const IterRaw& q, int y) {
  IterRow r;
  r.p.car = k.plane ? q.car : k.factored(y);
  r.p.chi = q.chi;
  r.p.yc = luma_res(q.s, q.chi, r.p.car);
  r.d = make_float2(0.f, 0.f);
  r.dY = 0.f;
  r.rhs = q.rhs;
  return r;
}

// The moved row from the row before the move and its gradient g.
template <bool Upd>
__device__ inline IterRow moved_row(const OldRow& o, float2 g, float alpha, float beta) {
  IterRow r;
  r.p.car = o.p.car;
  r.d = Upd ? dir_update(g, beta, o.dp) : make_float2(-g.x, -g.y);
  r.dY = -re_mul(r.d, r.p.car);
  r.p.chi = axpy(o.p.chi, alpha, r.d);
  r.p.yc = luma_res(o.s, r.p.chi, r.p.car);
  r.rhs = o.rhs;
  return r;
}

// The gradient at a row from its sample a (right neighbour by a lane shift),
// its weights and the fluxes through its upper and lower edges.  The same
// operations as the pass that formed this gradient from the same samples (so
// the same bits).
template <bool Adj>
__device__ inline float2 row_grad_fluxes(const FieldCG& s, const Sample& a, const float4& w, float2 rhs,
                                         const Flux& fup, const Flux& fdn, bool calc) {
  Sample ar;
  ar.yc = shr(a.yc);
  ar.chi = shr(a.chi);
  const Flux fx = flux(w.x, w.z, a, ar);  // through the right edge
  Flux fxl;
  fxl.y = shl(fx.y);
  fxl.c = shl(fx.c);
  float2 g = grad_from_flux(s, a.car, fxl, fx, fup, fdn);
  if (Adj) g = make_float2(g.x - rhs.x, g.y - rhs.y);
  return calc ? g : make_float2(0.f, 0.f);
}
// ... given the row below b: returns the flux through the lower edge in fdn
template <bool Adj>
__device__ inline float2 row_grad(const FieldCG& s, const Sample& a, const float4& w, const Sample& b, float2 rhs,
                                  const Flux& fup, Flux& fdn, bool calc) {
  fdn = flux(w.y, w.w, a, b);  // through the lower edge
  return row_grad_fluxes<Adj>(s, a, w, rhs, fup, fdn, calc);
}

// One pixel's share of curv(a) and of the bilinear form curv(a, b)
// (variational.cpp:352-399) from forward differences to the right and down
// neighbours, with the pixel's own weights (zero on edges that leave the
// plane); aY = -Re[a c].  The pixel's terms are formed in float and summed in
// FP64; `on` false adds nothing.
__device__ inline void curv_terms(const FieldCG& s, float4 w, bool on, float2 a0, float2 ar, float2 ad, float aY0,
                                  float aYr, float aYd, float2 b0, float2 br, float2 bd, float bY0, float bYr,
                                  float bYd, bool with_b, double& A, double& B) {
  const float axY = aYr - aY0, ayY = aYd - aY0;
  const float2 ax = make_float2(ar.x - a0.x, ar.y - a0.y);
  const float2 ay = make_float2(ad.x - a0.x, ad.y - a0.y);
  const float ta = s.lam_y * (w.x * axY * axY + w.y * ayY * ayY) + s.mu_h * w.z * (ax.x * ax.x + ax.y * ax.y) +
                   s.mu_v * w.w * (ay.x * ay.x + ay.y * ay.y);
  A += (double)(on ? ta : 0.f);
  if (with_b) {
    const float bxY = bYr - bY0, byY = bYd - bY0;
    const float2 bx = make_float2(br.x - b0.x, br.y - b0.y);
    const float2 by = make_float2(bd.x - b0.x, bd.y - b0.y);
    const float tb = s.lam_y * (w.x * axY * bxY + w.y * ayY * byY) + s.mu_h * w.z * (ax.x * bx.x + ax.y * bx.y) +
                     s.mu_v * w.w * (ay.x * by.x + ay.y * by.y);
    B += (double)(on ? tb : 0.f);
  }
}

// IRLS lagged-diffusivity weights of one pixel (variational.cpp:406-433) from
// its own sample p0, its right neighbour's luma yr and chroma cr and its lower
// neighbour pd.  An edge that leaves the plane (hr / hd false) gets weight 0
// (its difference is 0 in hvd-core, so its term vanishes either way).
__device__ inline float4 irls_weight(const FieldCG& s, const Sample& p0, float yr, float2 cr, const Sample& pd, bool hr,
                                     bool hd) {
#pragma clang fp contract(off)
  const float e2 = s.eps * s.eps, ec2 = s.eps_c * s.eps_c;
  const float gx = hr ? yr - p0.yc : 0.f;
  const float gy = hd ? pd.yc - p0.yc : 0.f;
  const float2 cx = hr ? make_float2(cr.x - p0.chi.x, cr.y - p0.chi.y) : make_float2(0.f, 0.f);
  const float2 cy = hd ? make_float2(pd.chi.x - p0.chi.x, pd.chi.y - p0.chi.y) : make_float2(0.f, 0.f);
  const float mx = sqrtf(cx.x * cx.x + cx.y * cx.y), my = sqrtf(cy.x * cy.x + cy.y * cy.y);
  // eps / sqrt(.) as eps * rsq(.) (v_rsq_f32, specified to 1 ulp, like the
  // correctly rounded quotient of a correctly rounded root) when every
  // argument is a normal float (>= eps^2); otherwise the IEEE forms
  float4 w;
  if (e2 >= 1.17549435e-38f && ec2 >= 1.17549435e-38f) {
    w.x = hr ? s.eps * __builtin_amdgcn_rsqf(gx * gx + e2) : 0.f;
    w.y = hd ? s.eps * __builtin_amdgcn_rsqf(gy * gy + e2) : 0.f;
    w.z = hr ? s.eps_c * __builtin_amdgcn_rsqf(mx * mx + s.k * gx * gx + ec2) : 0.f;
    w.w = hd ? s.eps_c * __builtin_amdgcn_rsqf(my * my + s.k * gy * gy + ec2) : 0.f;
  } else {
    w.x = hr ? s.eps / sqrtf(gx * gx + e2) : 0.f;
    w.y = hd ? s.eps / sqrtf(gy * gy + e2) : 0.f;
    w.z = hr ? s.eps_c / sqrtf(mx * mx + s.k * gx * gx + ec2) : 0.f;
    w.w = hd ? s.eps_c / sqrtf(my * my + s.k * gy * gy + ec2) : 0.f;
  }
  return w;
}

// The fused CG iteration over one strip (variational.cpp:527-548, one
// iteration per launch).  It forms the gradient g of chi (the same bits the
// previous pass formed and summed: no gradient plane is stored), the direction
// d = d(it) = -g + beta d(it-1) and chi' = chi + alpha d, writes chi' and d,
// and forms g' = grad(chi') (:286-347, less rhs in the adjoint solve) and the
// strip's FP64 partials of
//   |g'|^2,  curv(g'),  curv(g', d),  Re<g', d>,
// from which the next reduction forms beta = |g'|^2 / |g|^2 and, without
// another pass over the planes, the next direction's
//   curv(-g' + beta d) = curv(g') - 2 beta curv(g', d) + beta^2 curv(d)
//   Re<g', -g' + beta d> = -|g'|^2 + beta Re<g', d>.
// The stage-start pass (Start) forms g = grad(chi) and the partials of |g|^2
// and curv(g) (the first direction is -g).
//
// Lanes: g of chi on lanes 1 .. 62 (each needs both neighbours), so d and
// chi' there; g' on lanes kStripHalo .. kStripHalo + kStripCols; the
// kStripCols lanes from kStripHalo own the outputs and the partials (a
// pixel's curvature terms need g' one column right).  Rows: the walk forms g'
// for rows + 1 lines (the curvature terms need g' one row down), chi' from
// the row above the strip and g of chi from two rows above.  Loads are
// clamped and results selected; the walk has no exit between lines.
//
// FuseW (the forward's stage start): the IRLS weights of the stage are formed
// here from chi (variational.cpp:406-433) instead of loaded, and stored for
// the stage's iterations (with chi into f.save + stage when set).
template <bool Start, bool Upd, bool Adj, bool FuseW = false>
__device__ inline void cg_strip(const DevField& f, const FieldCG& s, const CgDyn& q, const StripCtx& c, int rows,
                                double* partials, int stage = 0) {
  static_assert(!FuseW || Start, "weights are formed by the stage-start pass only");
  static_assert(!Start || !Upd, "the stage start has no previous direction");
  const Carrier k = carrier_for(f, c);
  const float2* chi = pick(f.chi, q.cur);
  const float2* dprev = pick(f.d, q.cur);
  float2* chi_out = pick(f.chi, q.cur ^ 1);
  float2* d_out = pick(f.d, q.cur ^ 1);
  const float alpha = q.alpha, beta = q.beta;
  const int lane = lane_id();
  const bool calc = c.valid && lane >= kStripHalo && lane <= kStripHalo + kStripCols;  // g'
  const bool calc0 = c.valid && lane >= 1 && lane <= 62;                               // g of chi
  auto fetch = [&](int y) { return iter_fetch<Start, Upd, Adj, FuseW>(f, k, c.x, chi, dprev, y); };
  auto old = [&](const IterRaw& r, int y) { return old_row(f, k, r, y, c.valid); };
  float2 gp = make_float2(0.f, 0.f);  // g' of the row above
  float gYp = 0.f;
  Flux fyu;  // flux of chi' through the upper edge of the current line
  Flux fdo;  // flux of chi through the lower edge of the row below the line
  double gg = 0.0, A = 0.0, B = 0.0, D = 0.0;
  // One line of the walk (row y = y0 + j, inside the plane).  The rows above
  // / at / below y and their weights live in three slots each that rotate
  // between calls, so no registers are copied from line to line; so do the
  // rows before the move (o1: row y + 1, o2: row y + 2, formed here) and the
  // loads (pin: row y + 2 in the iteration, y + 1 in the stage start, issued
  // by the line before; pout issued here for the next line): a line waits
  // only for loads issued a whole line earlier.
  auto step = [&](int j, const IterRow& up, const IterRow& cur, IterRow& dn, const float4& wu, float4& wc,
                  float4& wn, const IterRaw& pin, IterRaw& pout, const OldRow& o1, OldRow& o2) {
    const int y = c.y0 + j;
    if (Start) {
      pout = fetch(y + 2);
      if (!FuseW) wn = weights_mask(f, pin.w, y + 1, c.valid);
      dn = start_row(k, pin, y + 1);
    } else {
      pout = fetch(y + 3);
      o2 = old(pin, y + 2);
      Flux fd1;
      const float2 g1 = row_grad<Adj>(s, o1.p, o1.w, o2.p, o1.rhs, fdo, fd1, calc0);
      fdo = fd1;
      wn = o1.w;
      dn = moved_row<Upd>(o1, g1, alpha, beta);
    }
    Sample pr;
    pr.yc = shr(cur.p.yc);
    pr.chi = shr(cur.p.chi);
    if (FuseW) {  // this line's weights, from this line and the next
      wc = c.valid ? irls_weight(s, cur.p, pr.yc, pr.chi, dn.p, c.x + 1 < f.w, y + 1 < f.h)
                   : make_float4(0.f, 0.f, 0.f, 0.f);
      if (j < rows && c.out) {
        const unsigned i = (unsigned)(y * f.w + c.x);
        gst(f.wts, (size_t)i, wc);
        if (f.save) gst(f.save + (size_t)stage * f.h * f.w, (size_t)i, cur.p.chi);
      }
    }
    const Flux fx = flux(wc.x, wc.z, cur.p, pr);  // through the right edge
    const Flux fy = flux(wc.y, wc.w, cur.p, dn.p);  // through the lower edge
    Flux fxl;
    fxl.y = shl(fx.y);
    fxl.c = shl(fx.c);
    float2 gv = grad_from_flux(s, cur.p.car, fxl, fx, fyu, fy);
    fyu = fy;
    if (Adj) gv = make_float2(gv.x - cur.rhs.x, gv.y - cur.rhs.y);
    gv = calc ? gv : make_float2(0.f, 0.f);
    const float gY = -re_mul(gv, cur.p.car);
    // curvature terms of the row above, whose down neighbour is this row
    const float2 gpr = shr(gp), dpr = shr(up.d);
    const float gYpr = shr(gYp), dYpr = shr(up.dY);
    curv_terms(s, wu, j >= 1 && c.out, gp, gpr, gv, gYp, gYpr, gY, up.d, dpr, cur.d, up.dY, dYpr, cur.dY, !Start, A,
               B);
    // |g|^2 and Re<g, d> of the pixel, formed in float and summed in FP64
    // (as the curvature terms)
    const bool own = j < rows && c.out;
    gg += (double)(own ? __fmaf_rn(gv.x, gv.x, gv.y * gv.y) : 0.f);
    if (!Start) D += (double)(own ? __fmaf_rn(gv.x, cur.d.x, gv.y * cur.d.y) : 0.f);
    if (!Start && own) {
      const unsigned i = (unsigned)(y * f.w + c.x);
      gst32(chi_out, i, cur.p.chi);
      gst32(d_out, i, cur.d);
    }
    gp = gv;
    gYp = gY;
  };
  // Line j at the plane's last row + 1 (when the plane ends inside the walk):
  // the curvature terms of the row above, whose down neighbour is outside.
  auto plane_end = [&](int j, const IterRow& up, const float4& wu) {
    const float2 gpr = shr(gp), dpr = shr(up.d);
    const float gYpr = shr(gYp), dYpr = shr(up.dY);
    curv_terms(s, wu, j >= 1 && c.out, gp, gpr, gp, gYp, gYpr, gYp, up.d, dpr, up.d, up.dY, dYpr, up.dY, !Start, A, B);
  };
  IterRow r0, r1, r2;
  float4 w0, w1, w2;
  IterRaw p0, p1, p2;
  OldRow oa, ob, oc;
  if (Start) {
    const IterRaw a = fetch(c.y0 - 1), b = fetch(c.y0);
    p0 = fetch(c.y0 + 1);
    r0 = start_row(k, a, c.y0 - 1);
    r1 = start_row(k, b, c.y0);
    if (FuseW) {  // the weights of the line above the strip (rows outside the plane: 0)
      const float yr = shr(r0.p.yc);
      const float2 cr = shr(r0.p.chi);
      w0 = c.valid && c.y0 >= 1 ? irls_weight(s, r0.p, yr, cr, r1.p, c.x + 1 < f.w, c.y0 < f.h)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      w0 = weights_mask(f, a.w, c.y0 - 1, c.valid);
      w1 = weights_mask(f, b.w, c.y0, c.valid);
    }
  } else {
    // rows y0 - 2 .. y0 + 1 before the move; g of chi at y0 - 1 and y0
    const IterRaw am2 = fetch(c.y0 - 2), am1 = fetch(c.y0 - 1), a0 = fetch(c.y0), a1 = fetch(c.y0 + 1);
    p0 = fetch(c.y0 + 2);
    const OldRow om2 = old(am2, c.y0 - 2), om1 = old(am1, c.y0 - 1), o0 = old(a0, c.y0);
    oa = old(a1, c.y0 + 1);
    fdo = flux(om2.w.y, om2.w.w, om2.p, om1.p);
    Flux fd;
    const float2 gm1 = row_grad<Adj>(s, om1.p, om1.w, o0.p, om1.rhs, fdo, fd, calc0);
    const float2 g0 = row_grad<Adj>(s, o0.p, o0.w, oa.p, o0.rhs, fd, fdo, calc0);
    r0 = moved_row<Upd>(om1, gm1, alpha, beta);
    r1 = moved_row<Upd>(o0, g0, alpha, beta);
    w0 = om1.w;
    w1 = o0.w;
  }
  fyu = flux(w0.y, w0.w, r0.p, r1.p);
  // lines 0 .. jl lie inside the plane (jl <= rows: the walk forms g' one
  // line past the strip's last output row)
  const int jl = down_walk_last(c.y0, rows, f.h);
  int j = 0;
  for (; j + 2 <= jl; j += 3) {
    step(j, r0, r1, r2, w0, w1, w2, p0, p1, oa, ob);
    step(j + 1, r1, r2, r0, w1, w2, w0, p1, p2, ob, oc);
    step(j + 2, r2, r0, r1, w2, w0, w1, p2, p0, oc, oa);
  }
  const int rem = jl + 1 - j;  // 0, 1 or 2 lines left
  if (rem >= 1) step(j, r0, r1, r2, w0, w1, w2, p0, p1, oa, ob);
  if (rem >= 2) step(j + 1, r1, r2, r0, w1, w2, w0, p1, p2, ob, oc);
  if (jl < rows) {
    if (rem == 0) plane_end(j, r0, w0);
    else if (rem == 1) plane_end(j + 1, r1, w1);
    else plane_end(j + 2, r2, w2);
  }
  const int slot = f.tile0 + c.lt;
  publish(partials, slot, 0, kCgPartials, 0, gg);
  publish(partials, slot, 0, kCgPartials, 1, A);
  if (!Start) {
    publish(partials, slot, 0, kCgPartials, 2, B);
    publish(partials, slot, 0, kCgPartials, 3, D);
  }
}

// The fused CG iteration of cg_strip walked bottom-up, for strips on odd
// strip rows: two vertically adjacent strips then read the rows around
// their shared edge at the same time (both at the start or both at the end
// of their walks), so the second read is served by the L2.  Every per-pixel
// value is formed by the same expressions as in the top-down walk (a vertical
// flux always from the upper row's weights, upper sample first; a pixel's
// curvature terms from its right and lower neighbours), so the results are
// the same bits; only the order of the FP64 partial sums differs.
//
// The walk starts at T = min(y0 + rows, h) (g' there feeds the curvature of
// the line above) and ends at y0; line y forms chi'(y - 1) from the gradient
// of chi at y - 1, then g'(y), the curvature terms of line y (its lower
// neighbour's g' came one line earlier) and line y's outputs.
template <bool Upd, bool Adj>
__device__ inline void cg_strip_up(const DevField& f, const FieldCG& s, const CgDyn& q, const StripCtx& c, int rows,
                                   double* partials) {
  const Carrier k = carrier_for(f, c);
  const float2* chi = pick(f.chi, q.cur);
  const float2* dprev = pick(f.d, q.cur);
  float2* chi_out = pick(f.chi, q.cur ^ 1);
  float2* d_out = pick(f.d, q.cur ^ 1);
  const float alpha = q.alpha, beta = q.beta;
  const int lane = lane_id();
  const bool calc = c.valid && lane >= kStripHalo && lane <= kStripHalo + kStripCols;  // g'
  const bool calc0 = c.valid && lane >= 1 && lane <= 62;                               // g of chi
  auto fetch = [&](int y) { return iter_fetch<false, Upd, Adj, false>(f, k, c.x, chi, dprev, y); };
  auto old = [&](const IterRaw& r, int y) { return old_row(f, k, r, y, c.valid); };
  float2 gp = make_float2(0.f, 0.f);  // g' of the line below
  float gYp = 0.f;
  Flux fdn;   // flux of chi' through the lower edge of the current line
  Flux fdo;   // flux of chi through the lower edge of the row above the line
  double gg = 0.0, A = 0.0, B = 0.0, D = 0.0;
  const int T = up_walk_top(c.y0, rows, f.h);
  // Line j of the walk (row y = T - j).  Slots as in cg_strip: dn / up / cur
  // rotate over the rows below / at / above the line (`above` is filled
  // here), the weights with them, the rows before the move (o1: row y - 1,
  // o2: row y - 2, formed here) and the loads (pin: row y - 2, pout: y - 3).
  // (the weights of the line below ride along in the rotation unused)
  auto step = [&](int j, const IterRow& dn, const IterRow& cur, IterRow& above, const float4& /*wd*/,
                  const float4& wc, float4& wa, const IterRaw& pin, IterRaw& pout, const OldRow& o1, OldRow& o2) {
    const int y = T - j;
    pout = fetch(y - 3);
    o2 = old(pin, y - 2);
    // gradient of chi at y - 1: its upper edge's flux from row y - 2
    const Flux fup1 = flux(o2.w.y, o2.w.w, o2.p, o1.p);
    const float2 g1 = row_grad_fluxes<Adj>(s, o1.p, o1.w, o1.rhs, fup1, fdo, calc0);
    fdo = fup1;
    wa = o1.w;
    above = moved_row<Upd>(o1, g1, alpha, beta);
    // g' at y: upper edge from chi'(y - 1) and the weights of row y - 1
    Sample pr;
    pr.yc = shr(cur.p.yc);
    pr.chi = shr(cur.p.chi);
    const Flux fx = flux(wc.x, wc.z, cur.p, pr);  // through the right edge
    const Flux fu = flux(wa.y, wa.w, above.p, cur.p);  // through the upper edge
    Flux fxl;
    fxl.y = shl(fx.y);
    fxl.c = shl(fx.c);
    float2 gv = grad_from_flux(s, cur.p.car, fxl, fx, fu, fdn);
    fdn = fu;
    if (Adj) gv = make_float2(gv.x - cur.rhs.x, gv.y - cur.rhs.y);
    gv = calc ? gv : make_float2(0.f, 0.f);
    const float gY = -re_mul(gv, cur.p.car);
    // curvature terms of line y, whose lower neighbour is the previous line
    const bool own = j >= 1 && c.out;
    const float2 gvr = shr(gv), dr = shr(cur.d);
    const float gYr = shr(gY), dYr = shr(cur.dY);
    curv_terms(s, wc, own, gv, gvr, gp, gY, gYr, gYp, cur.d, dr, dn.d, cur.dY, dYr, dn.dY, true, A, B);
    gg += (double)(own ? __fmaf_rn(gv.x, gv.x, gv.y * gv.y) : 0.f);
    D += (double)(own ? __fmaf_rn(gv.x, cur.d.x, gv.y * cur.d.y) : 0.f);
    if (own) {
      const unsigned i = (unsigned)(y * f.w + c.x);
      gst32(chi_out, i, cur.p.chi);
      gst32(d_out, i, cur.d);
    }
    gp = gv;
    gYp = gY;
  };
  // rows T + 2 .. T - 1 before the move; g of chi at T + 1 and T
  IterRow r0, r1, r2;
  float4 w0, w1, w2;
  IterRaw p0, p1, p2;
  OldRow oa, ob, oc;
  {
    const IterRaw b2 = fetch(T + 2), b1 = fetch(T + 1), b0 = fetch(T), bm1 = fetch(T - 1);
    p0 = fetch(T - 2);
    const OldRow o2 = old(b2, T + 2), o1 = old(b1, T + 1), o0 = old(b0, T);
    oa = old(bm1, T - 1);
    const Flux fd1 = flux(o1.w.y, o1.w.w, o1.p, o2.p);  // lower edge of T + 1
    const Flux fu1 = flux(o0.w.y, o0.w.w, o0.p, o1.p);  // upper edge of T + 1
    const float2 g1 = row_grad_fluxes<Adj>(s, o1.p, o1.w, o1.rhs, fu1, fd1, calc0);
    fdo = flux(oa.w.y, oa.w.w, oa.p, o0.p);  // upper edge of T
    const float2 g0 = row_grad_fluxes<Adj>(s, o0.p, o0.w, o0.rhs, fdo, fu1, calc0);
    r0 = moved_row<Upd>(o1, g1, alpha, beta);  // chi'(T + 1)
    r1 = moved_row<Upd>(o0, g0, alpha, beta);  // chi'(T)
    w0 = o1.w;
    w1 = o0.w;
  }
  fdn = flux(w1.y, w1.w, r1.p, r0.p);  // chi' through the lower edge of T
  // lines T .. y0
  const int jl = T - c.y0;
  int j = 0;
  for (; j + 2 <= jl; j += 3) {
    step(j, r0, r1, r2, w0, w1, w2, p0, p1, oa, ob);
    step(j + 1, r1, r2, r0, w1, w2, w0, p1, p2, ob, oc);
    step(j + 2, r2, r0, r1, w2, w0, w1, p2, p0, oc, oa);
  }
  const int rem = jl + 1 - j;  // 0, 1 or 2 lines left
  if (rem >= 1) step(j, r0, r1, r2, w0, w1, w2, p0, p1, oa, ob);
  if (rem >= 2) step(j + 1, r1, r2, r0, w1, w2, w0, p1, p2, ob, oc);
  const int slot = f.tile0 + c.lt;
  publish(partials, slot, 0, kCgPartials, 0, gg);
  publish(partials, slot, 0, kCgPartials, 1, A);
  publish(partials, slot, 0, kCgPartials, 2, B);
  publish(partials, slot, 0, kCgPartials, 3, D);
}

}  // namespace

// Reset per-field solver state before a refine (replaces any memset so the
// sequence stays graph-capturable).
__global__ void k_refine_begin(FieldCG* st, const FieldCG* init, int nf) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) st[f] = init[f];
}

// ResolveChromaAniso samples (variational.cpp:150-178): group-averaged chi0,
// t = 1,3,5,..., x = 1,4,7,...; |pav(t,x) - pav(t,x-1)| and
// |pav(t,x) - pav(t-1,x)|.  One thread per sample over the whole batch.
__global__ void k_aniso_samples(const DevField* F, const FieldCG* st, const int* sample_prefix, int nf,
                                float* gx, float* gy) {
  const int n_total = sample_prefix[nf];
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n_total; idx += gridDim.x * blockDim.x) {
    const int fi = field_of_tile(sample_prefix, nf, idx);
    const DevField& f = F[fi];
    const int group = st[fi].is_pal ? 4 : 2;
    const int nx = (f.w - 1 + 2) / 3;
    const int j = idx - sample_prefix[fi];
    const int t = 1 + 2 * (j / nx);
    const int x = 1 + 3 * (j % nx);
    const float2* c = f.chi[0];
    float2 p = make_float2(0.f, 0.f), pl = p, pu = p;
    for (int k = 0; k < group; ++k) {
      const float2 a = c[(size_t)(group * t + k) * f.w + x];
      const float2 b = c[(size_t)(group * t + k) * f.w + x - 1];
      const float2 u = c[(size_t)(group * (t - 1) + k) * f.w + x];
      p.x += a.x; p.y += a.y;
      pl.x += b.x; pl.y += b.y;
      pu.x += u.x; pu.y += u.y;
    }
    const float inv = 1.f / (float)group;
    const float ex = (p.x - pl.x) * inv, ey = (p.y - pl.y) * inv;
    const float ux = (p.x - pu.x) * inv, uy = (p.y - pu.y) * inv;
    gx[idx] = sqrtf(ex * ex + ey * ey);
    gy[idx] = sqrtf(ux * ux + uy * uy);
  }
}

// Exact k-th smallest of n non-negative floats by 4 x 8-bit radix selection
// (their bit patterns order as unsigned integers; the value
// std::nth_element returns at k).  One block of kSelThreads threads; each
// wave counts into its own 256-bin histogram (hist: 16 x 256 in LDS), the
// bin totals are scanned in parallel.  v may be in LDS or global memory.
__device__ float radix_select(const float* v, int n, int k, unsigned* hist) {
  __shared__ unsigned tot[256];
  __shared__ unsigned s_prefix, s_k;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    s_prefix = 0u;
    s_k = (unsigned)k;
  }
  for (int pass = 3; pass >= 0; --pass) {
    for (int i = threadIdx.x; i < 16 * 256; i += kSelThreads) hist[i] = 0u;
    __syncthreads();
    const unsigned prefix = s_prefix, kk = s_k;  // both read before this pass may update them
    const int hi_shift = (pass + 1) * 8;
    unsigned* h = hist + wv * 256;
    for (int i = threadIdx.x; i < n; i += kSelThreads) {
      const unsigned u = __float_as_uint(v[i]);
      if (hi_shift >= 32 || (u >> hi_shift) == prefix) atomicAdd(&h[(u >> (pass * 8)) & 255u], 1u);
    }
    __syncthreads();
    unsigned cnt = 0u, incl = 0u;
    if (threadIdx.x < 256) {
      for (int w = 0; w < 16; ++w) cnt += hist[w * 256 + threadIdx.x];
      incl = cnt;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
      }
      tot[threadIdx.x] = incl;  // inclusive scan within the wave
    }
    __syncthreads();
    if (threadIdx.x < 256) {
      for (int w = 0; w < wv; ++w) incl += tot[w * 64 + 63];
      const unsigned excl = incl - cnt;
      if (excl <= kk && kk < incl) {  // exactly one bin holds the k-th value
        s_k = kk - excl;
        s_prefix = (prefix << 8) | (unsigned)threadIdx.x;
      }
    }
    __syncthreads();
  }
  return __uint_as_float(s_prefix);
}

// p98 of both sample sets, the ratio and the NTSC / PAL map into aniso
// (variational.cpp:182-195); sets mu_h, mu_v.  One block per field; the
// field's samples are staged in dynamic LDS when they fit (lds_floats).
__global__ void __launch_bounds__(kSelThreads) k_aniso_select(FieldCG* st, const int* sample_prefix, const float* gx,
                                                              const float* gy, float lambda_c, int lds_floats) {
  extern __shared__ float staged[];
  __shared__ unsigned hist[16 * 256];
  const int fi = blockIdx.x;
  FieldCG& s = st[fi];
  float aniso = s.aniso_forced;
  if (!(aniso > 0.f)) {
    const int off = sample_prefix[fi], n = sample_prefix[fi + 1] - off;
    float px = 0.f, py = 0.f;
    if (n > 0) {
      const int k = n - 1 - n / 50;
      const bool fit = n <= lds_floats;
      for (int comp = 0; comp < 2; ++comp) {
        const float* src = (comp ? gy : gx) + off;
        if (fit) {
          for (int i = threadIdx.x; i < n; i += kSelThreads) staged[i] = src[i];
          __syncthreads();
          src = staged;
        }
        const float r = radix_select(src, n, k, hist);
        if (comp) py = r; else px = r;
        __syncthreads();  // the staging buffer is reused
      }
    }
    const float r = py / fmaxf(px, 1e-6f);
    aniso = s.is_pal ? fminf(fmaxf(0.7f + 0.9f * (r - 1.3f), 0.55f), 1.f)
                     : fminf(fmaxf(0.5f + 1.1f * (r - 1.3f), 0.5f), 1.f);
  }
  if (threadIdx.x == 0) {
#pragma clang fp contract(off)
    // the oriented prior's weight, and the renormalisation that keeps the
    // total chroma prior mass of the axis terms (variational.cpp:205-219)
    float mu_h = lambda_c * aniso, mu_v = lambda_c, mu_d = lambda_c * s.diag * 0.5f;
    const float tot0 = lambda_c * (1.f + aniso);
    const float sc = mu_d > 0.f ? tot0 / (tot0 + 2.f * mu_d) : 1.f;
    s.mu_h = mu_h * sc;
    s.mu_v = mu_v * sc;
    s.mu_d = mu_d * sc;
    s.aniso = aniso;
  }
}

// IRLS lagged-diffusivity weights at the current chi (variational.cpp:406-433).
__global__ void __launch_bounds__(64) k_irls_weights(const DevField* __restrict__ F,
    const FieldCG* __restrict__ st, StripGrid sg, int stage) {
  const StripCtx c = strip_ctx(F, sg);
  const int rows = sg.rows;
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  const FieldCG s = st[c.f];
  const Carrier k = carrier_for(f, c);
  const float2* chi = pick(f.chi, s.cur);
  Sample cur = make_sample(f, k, raw_row(f, k, chi, nullptr, c.y0, c.valid), false, 0.f, c.y0, c.valid);
  for (int r = 0; r < rows; ++r) {
    const int y = c.y0 + r;
    if (y >= f.h) break;
    const Sample dn = make_sample(f, k, raw_row(f, k, chi, nullptr, y + 1, c.valid), false, 0.f, y + 1, c.valid);
    const float yr = shr(cur.yc);
    const float2 cr = shr(cur.chi);
    if (c.out) {
      const float4 w = irls_weight(s, cur, yr, cr, dn, c.x + 1 < f.w, y + 1 < f.h);
      gst(f.wts, (size_t)y * f.w + c.x, w);
      if (f.save) gst(f.save + (size_t)stage * f.h * f.w, (size_t)y * f.w + c.x, cur.chi);
    }
    cur = dn;
  }
}

// Stage start (variational.cpp:406-433, 519-525): for a forward solve the
// stage's IRLS weights from chi (stored for the iterations), then g = grad(chi)
// and the partials of |g|^2 and curv(g) (the first direction is d = -g), into
// partial slot 1.
__global__ void __launch_bounds__(64) k_cg_start(const DevField* __restrict__ F, const FieldCG* __restrict__ st,
                                                 StripGrid sg, double* partials, int adjoint, int stage) {
  const StripCtx c = strip_ctx(F, sg);
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  const FieldCG s = st[c.f];
  CgDyn q{};
  q.cur = s.cur;
  double* out = partials + (size_t)sg.nb * kCgPartials;
  // the adjoint's stage weights come from k_irls_weights (they belong to the
  // forward's composite, which the adjoint fields do not carry)
  if (adjoint) cg_strip<true, false, true>(f, s, q, c, sg.rows, out);
  else cg_strip<true, false, false, true>(f, s, q, c, sg.rows, out, stage);
}

namespace {

// CG iteration `it` (variational.cpp:527-548): the wave first sums its field's
// partials of the previous launch (slot (it + 1) & 1) and advances the
// field's scalars; the field's first strip publishes them in snapshot it & 1.
// Then d <- -g + beta d, chi <- chi + alpha d, g <- grad(chi), and this
// launch's partials into slot it & 1.
template <bool Adj>
__device__ inline void cg_iter(const DevField* __restrict__ F, FieldCG* __restrict__ st, const StripGrid& sg,
                               double* partials, int it) {
  const StripCtx c = strip_ctx(F, sg);
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  const FieldCG s = st[c.f];
  const size_t slot = (size_t)sg.nb * kCgPartials;
  const int ns = sg.spf > 0 ? sg.spf : sg.tp[c.f + 1] - sg.tp[c.f];
  const CgSums sums = field_sums(partials + ((it + 1) & 1) * slot, f.tile0, ns);
  // the snapshot is read from memory with a uniform index (a dynamic index
  // into the register copy would put it on the stack)
  CgDyn q = it == 0 ? cg_first(s, s.cur, sums) : cg_advance(s, st[c.f].dyn[(it - 1) & 1], sums);
  if (c.lt == 0 && lane_id() == 0) st[c.f].dyn[it & 1] = q;
  // every lane holds the same values; make that visible to the compiler
  // (scalar registers, a uniform branch)
  q.done = __builtin_amdgcn_readfirstlane(q.done);
  q.cur = __builtin_amdgcn_readfirstlane(q.cur);
  q.alpha = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.alpha)));
  q.beta = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(q.beta)));
  if (q.done) return;
  double* out = partials + (it & 1) * slot;
  // strips on odd strip rows walk bottom-up (uniform per wave)
  if (strip_walks_up(c.lt, f.tiles_x)) {
    if (it > 0) cg_strip_up<true, Adj>(f, s, q, c, sg.rows, out);
    else cg_strip_up<false, Adj>(f, s, q, c, sg.rows, out);
  } else {
    if (it > 0) cg_strip<false, true, Adj>(f, s, q, c, sg.rows, out);
    else cg_strip<false, false, Adj>(f, s, q, c, sg.rows, out);
  }
}

}  // namespace

__global__ void __launch_bounds__(64) k_cg_iter(const DevField* __restrict__ F, FieldCG* __restrict__ st,
                                                StripGrid sg, double* partials, int it) {
  cg_iter<false>(F, st, sg, partials, it);
}

// The same iteration for an adjoint solve (no composite; the right-hand side
// is subtracted from the gradient).
__global__ void __launch_bounds__(64) k_cg_iter_adj(const DevField* __restrict__ F, FieldCG* __restrict__ st,
                                                    StripGrid sg, double* partials, int it) {
  cg_iter<true>(F, st, sg, partials, it);
}

// End of a stage's CG loop of n_inner launches: the last launch's update,
// the final chi slot, and the iteration count of stage `stage` (diagnostics,
// when 0 <= stage < 16).  One 64-thread block per field.
__global__ void __launch_bounds__(64) k_cg_finish(FieldCG* st, StripGrid sg, const double* partials, int n_inner,
                                                  int stage) {
  const int fi = blockIdx.x;
  const FieldCG s = st[fi];
  int cur = s.cur, it = 0;
  if (n_inner > 0) {
    const size_t slot = (size_t)sg.nb * kCgPartials;
    const int t0 = sg.tp[fi], ns = sg.tp[fi + 1] - t0;
    const CgSums sums = field_sums(partials + ((n_inner - 1) & 1) * slot, t0, ns);
    const CgDyn q = cg_advance(s, st[fi].dyn[(n_inner - 1) & 1], sums);
    cur = q.cur;
    it = q.it;
  }
  if (threadIdx.x == 0) {
    st[fi].cur = cur;
    if (stage >= 0 && stage < 16) st[fi].iters[stage] = it;
  }
}

// Output (variational.cpp:552-555): Y = S - Re[chi c]; chi lands in the
// caller's buffer.
__global__ void __launch_bounds__(64) k_refine_finalize(const DevField* __restrict__ F,
    const FieldCG* __restrict__ st,
                                                                       StripGrid sg) {
  const StripCtx c = strip_ctx(F, sg);
  const int rows = sg.rows;
  const DevField f = F[c.f];  // a private copy: stores cannot alias it
  // every lane loads its row factor before any lane leaves: at(y) reads the
  // factor from lane y - ybase, which may be a lane that produces no output
  const Carrier k = carrier_for(f, c);
  if (!c.out) return;
  const int cur = st[c.f].cur;
  for (int r = 0; r < rows; ++r) {
    const int y = c.y0 + r;
    if (y >= f.h) break;
    const size_t i = (size_t)y * f.w + c.x;
    const float2 chi = gld(pick(f.chi, cur), i);
    if (cur) gst(f.chi[0], i, chi);
    if (f.y) gst(f.y, i, gld(f.s, i) - re_mul(chi, k.at(y)));
  }
}


}  // namespace hvdk
