// strip_bounds.cpp -- host check of the solver strip walk's addressing.
//
// Enumerates, for every block and lane of a launch, each global-memory
// access the strip-walk kernels of solve.hip form -- k_irls_weights,
// k_cg_start (forward stage start with fused weights, and the adjoint's),
// k_cg_iter / k_cg_iter_adj (top-down and bottom-up walks) and
// k_refine_finalize -- and checks it against the size of the plane or
// buffer it addresses.  The grid mapping (xcd_remap, field_of_tile,
// strip_col / strip_row0), the clamped load index and the walks' line
// bounds are the kernels' own functions (strip_geom.h); the sequence of rows
// each walk touches follows solve.hip's cg_strip, cg_strip_up, raw_row and
// the finalize loop (cited per access below).  Carriers in both forms: the
// factored one (row phasors broadcast by lane reads, whose lane index must
// stay inside the wave's carrier_span) and an explicit plane.
//
// Usage: strip_bounds [--plant]
//   --plant  form the walk's loads without the clamp (the addressing of an
//            unclamped explicit-carrier read); the check must then fail.
// Prints a summary line (and the first violation) and exits 1 on any
// violation.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "strip_geom.h"

using namespace hvdk;

namespace {

constexpr int kCgPartials = 4;  // solve.hip: FP64 partials per strip and launch
constexpr int kMinRows = 2;     // abi.hip: the shortest strips (sizes the partial buffer)
constexpr int kIrlsMax = 16;    // stage index bound of the saved chi planes

bool g_plant = false;

struct Plane {
  int h, w;
};

struct Report {
  long accesses = 0;
  long violations = 0;
  std::string first;
  void check(bool ok, const char* what, int f, int lt, int lane, int y, long idx, long size) {
    ++accesses;
    if (ok) return;
    if (violations++ == 0) {
      char b[256];
      std::snprintf(b, sizeof b, "%s: field %d strip %d lane %d row %d index %ld outside [0, %ld)", what, f, lt, lane,
                    y, idx, size);
      first = b;
    }
  }
};

// The load index of the walk (iter_fetch, Carrier::load), or the unclamped
// form under --plant.
long load_index(int y, int x, int h, int w) {
  if (g_plant) return (long)y * w + x;
  return (long)clamped_index(y, x, h, w);
}

// One launch over a batch: fields, strip rows, carrier form.
void run_launch(const std::vector<Plane>& F, int rows, bool explicit_carrier, Report& R) {
  const int nf = (int)F.size();
  // strip_setup (abi.hip): per-field strips across and first strip
  std::vector<int> tiles_x(nf), tile0(nf), tp(nf + 1, 0);
  long cap_strips = 0;
  for (int i = 0; i < nf; ++i) {
    tiles_x[i] = (F[i].w + kStripCols - 1) / kStripCols;
    tile0[i] = tp[i];
    tp[i + 1] = tp[i] + solver_strips(F[i].h, F[i].w, rows);
    cap_strips += solver_strips(F[i].h, F[i].w, kMinRows);
  }
  const int nb = tp[nf];
  int spf = tp[1];
  for (int i = 1; i < nf; ++i)
    if (F[i].h != F[0].h || F[i].w != F[0].w) spf = 0;
  const long part_cap = 2L * kCgPartials * cap_strips;  // two launch slots

  // the block -> strip mapping is a permutation of the launch's strips
  std::vector<char> seen(nb, 0);
  for (int b = 0; b < nb; ++b) {
    const int t = xcd_remap(b, nb);
    R.check(t >= 0 && t < nb, "xcd_remap", -1, t, -1, -1, t, nb);
    if (t >= 0 && t < nb) {
      R.check(!seen[t], "xcd_remap (repeated strip)", -1, t, -1, -1, t, nb);
      seen[t] = 1;
    }
  }

  for (int b = 0; b < nb; ++b) {
    const int t = xcd_remap(b, nb);
    if (t < 0 || t >= nb) continue;
    // strip_ctx
    const int f = spf > 0 ? t / spf : field_of_tile(tp.data(), nf, t);
    R.check(f >= 0 && f < nf && tp[f] <= t && t < tp[f + 1], "field_of_tile", f, t, -1, -1, f, nf);
    if (f < 0 || f >= nf) continue;
    const int lt = spf > 0 ? t - f * spf : t - tp[f];
    const int h = F[f].h, w = F[f].w;
    const long npx = (long)h * w;
    const int y0 = strip_row0(lt, tiles_x[f], rows);
    R.check(y0 >= 0 && y0 < h, "strip first row", f, lt, -1, y0, y0, h);
    // FP64 partials: publish (slot 0 or, in the stage start, slot 1) and the
    // field sums the next launch reads
    const long pslot = (long)(tile0[f] + lt) * kCgPartials;
    R.check(pslot >= 0 && (long)nb * kCgPartials + pslot + kCgPartials <= part_cap, "partials", f, lt, -1, -1,
            (long)nb * kCgPartials + pslot + kCgPartials - 1, part_cap);
    for (int j = 0; j < tp[f + 1] - tp[f]; j += 64) {
      const long p = (long)(tp[f] + j) * kCgPartials;
      R.check(p + kCgPartials <= part_cap, "field_sums", f, lt, -1, -1, p, part_cap);
    }

    for (int lane = 0; lane < 64; ++lane) {
      const int x = strip_col(lt, tiles_x[f], lane);
      const bool valid = x >= 0 && x < w;
      const bool out = lane >= kStripHalo && lane < kStripHalo + kStripCols && x < w;
      if (out) R.check(valid, "output lane column", f, lt, lane, -1, x, w);
      const int ybase = y0 - kStripHalo;
      // carrier_for: the column phasor of a valid column (the row phasor of
      // row ybase + lane is loaded only inside the plane)
      if (!explicit_carrier && valid) R.check(x >= 0 && x < w, "colph", f, lt, lane, -1, x, w);
      // A row's carrier: an explicit plane is loaded at the clamped index;
      // the factored one reads lane y - ybase, which must hold a row phasor
      // this strip loaded (carrier_span rows, inside the wave).
      auto carrier = [&](int y, const char* where) {
        if (explicit_carrier) {
          const long i = load_index(y, x, h, w);
          R.check(i >= 0 && i < npx, where, f, lt, lane, y, i, npx);
        } else {
          const int l = y - ybase;
          R.check(l >= 0 && l < carrier_span(rows) && l < 64, where, f, lt, lane, y, l, carrier_span(rows));
        }
      };
      // iter_fetch: chi, S, the previous direction, the explicit carrier,
      // the adjoint's right-hand side and the weights, all at one index
      auto fetch = [&](int y) {
        const long i = load_index(y, x, h, w);
        R.check(i >= 0 && i < npx, "walk load", f, lt, lane, y, i, npx);
        if (explicit_carrier) R.check(i >= 0 && i < npx, "walk carrier load", f, lt, lane, y, i, npx);
        // 32-bit byte offsets of gld32/gst32 on float4 weights
        R.check(i * 16 < (1L << 32), "32-bit offset", f, lt, lane, y, i * 16, 1L << 32);
      };
      auto store = [&](int y, const char* what) {
        const long i = (long)y * w + x;
        R.check(y >= 0 && y < h && x >= 0 && x < w && i < npx, what, f, lt, lane, y, i, npx);
      };

      // ---- k_irls_weights: raw_row / make_sample at y0 .. and the row below
      {
        auto row = [&](int y) {  // raw_row: guarded by row_in
          if (valid && y >= 0 && y < h) {
            const long i = (long)y * w + x;
            R.check(i < npx, "irls load", f, lt, lane, y, i, npx);
            carrier(y, "irls carrier");
          }
        };
        row(y0);
        for (int r = 0; r < rows; ++r) {
          const int y = y0 + r;
          if (y >= h) break;
          row(y + 1);
          if (out) {
            store(y, "irls weights store");
            const long i = (long)(kIrlsMax - 1) * npx + (long)y * w + x;
            R.check(i < (long)kIrlsMax * npx, "saved chi store", f, lt, lane, y, i, (long)kIrlsMax * npx);
          }
        }
      }

      // ---- k_cg_start: cg_strip<Start> (forward with fused weights, adjoint)
      {
        const int jl = down_walk_last(y0, rows, h);
        R.check(jl >= 0 && jl <= rows, "down walk lines", f, lt, lane, y0, jl, rows + 1);
        fetch(y0 - 1);
        fetch(y0);
        fetch(y0 + 1);
        carrier(y0 - 1, "start carrier");  // start_row of the two rows above the walk
        carrier(y0, "start carrier");
        for (int j = 0; j <= jl; ++j) {
          const int y = y0 + j;
          fetch(y + 2);                   // pout
          carrier(y + 1, "start carrier");  // dn = start_row(pin, y + 1)
          if (j < rows && out) {
            store(y, "stage weights store");
            const long i = (long)(kIrlsMax - 1) * npx + (long)y * w + x;
            R.check(i < (long)kIrlsMax * npx, "saved chi store", f, lt, lane, y, i, (long)kIrlsMax * npx);
          }
        }
      }

      // ---- k_cg_iter: cg_strip (top-down) or cg_strip_up (bottom-up)
      if (!strip_walks_up(lt, tiles_x[f])) {
        const int jl = down_walk_last(y0, rows, h);
        for (int y = y0 - 2; y <= y0 + 2; ++y) fetch(y);
        for (int y = y0 - 2; y <= y0 + 1; ++y) carrier(y, "iter carrier");  // old rows
        for (int j = 0; j <= jl; ++j) {
          const int y = y0 + j;
          fetch(y + 3);
          carrier(y + 2, "iter carrier");  // o2 = old(pin, y + 2)
          if (j < rows && out) {
            store(y, "chi store");
            store(y, "direction store");
          }
        }
      } else {
        const int T = up_walk_top(y0, rows, h);
        const int jl = T - y0;
        R.check(jl >= 0 && jl <= rows, "up walk lines", f, lt, lane, T, jl, rows + 1);
        for (int y = T + 2; y >= T - 2; --y) fetch(y);
        for (int y = T + 2; y >= T - 1; --y) carrier(y, "iter carrier (up)");
        for (int j = 0; j <= jl; ++j) {
          const int y = T - j;
          fetch(y - 3);
          carrier(y - 2, "iter carrier (up)");  // o2 = old(pin, y - 2)
          if (j >= 1 && out) {
            store(y, "chi store (up)");
            store(y, "direction store (up)");
          }
        }
      }

      // ---- k_refine_finalize: output lanes, rows y0 .. (inside the plane)
      if (out) {
        for (int r = 0; r < rows; ++r) {
          const int y = y0 + r;
          if (y >= h) break;
          store(y, "finalize chi/S load and chi/Y store");
          carrier(y, "finalize carrier");
        }
      }
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i)
    if (!std::strcmp(argv[i], "--plant")) g_plant = true;
  const std::vector<Plane> shapes = {{1, 1},     {1, 9},     {9, 1},     {67, 3},   {2, 59},  {3, 60},
                                     {242, 760}, {263, 758}, {288, 922}, {313, 1135}, {992, 2048}};
  long total = 0, bad = 0;
  int configs = 0;
  std::string first;
  auto run = [&](const std::vector<Plane>& batch, const char* name, bool all_rows = true) {
    for (int rows = kMinRows; rows <= kMaxStripRows; ++rows)
      for (int ex = 0; ex < 2; ++ex) {
        if (!all_rows && rows != kMinRows && rows != 7 && rows != 16 && rows != kMaxStripRows) continue;
        if (g_plant && bad) return;  // the planted mistake is seen: stop there
        Report R;
        run_launch(batch, rows, ex != 0, R);
        total += R.accesses;
        bad += R.violations;
        ++configs;
        if (R.violations && first.empty()) first = std::string(name) + " rows " + std::to_string(rows) +
                                                   (ex ? " explicit: " : " factored: ") + R.first;
      }
  };
  for (const Plane& p : shapes) {
    if ((long)p.h * p.w > 400000L) continue;  // the largest plane alone below
    char nm[64];
    std::snprintf(nm, sizeof nm, "%dx%d", p.h, p.w);
    run({p}, nm);
    run({p, p, p}, nm);  // equal sizes: the division mapping (spf > 0)
  }
  // a mixed batch (binary-search mapping) and the largest accepted plane
  std::vector<Plane> mixed;
  for (const Plane& p : shapes)
    if ((long)p.h * p.w <= 400000L) mixed.push_back(p);
  run(mixed, "mixed");
  run({{992, 2048}}, "992x2048", false);  // strip heights 2, 7, 16 and 32
  std::printf("%d launch configurations, %ld accesses checked, %ld outside their planes%s\n", configs, total, bad,
              g_plant ? " (loads unclamped)" : "");
  if (bad) std::printf("first: %s\n", first.c_str());
  return bad ? 1 : 0;
}
