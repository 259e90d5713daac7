// strip_geom.h -- the integer address arithmetic of the solver's strip walk
// (solve.hip): which field, columns and rows a strip covers, the order blocks
// take strips in, the clamped plane index every walk load goes through and
// the walk's line bounds.  Plain C++ with a host/device qualifier, so a host
// program can enumerate every address a launch forms and check it against
// the planes (tests/cpp/strip_bounds.cpp).
#pragma once

#if defined(__HIPCC__)
#define HVD_HD __host__ __device__
#else
#define HVD_HD
#endif

namespace hvdk {

// Output columns of a solver strip: 64 lanes less two halo lanes on the left
// and three on the right.  The fused CG iteration recomputes the gradient of
// chi (lanes 1 .. 62, each needs both neighbours), moves chi there, and forms
// the new gradient on lanes 2 .. 61: the outputs' curvature terms need it one
// column past the last output.
constexpr int kStripCols = 59;
constexpr int kStripHalo = 2;  // lane of the strip's first output column

// Rows of carrier factors a solver strip of `rows` lines holds, one per lane
// (two halo rows above and three below); at most 64.
constexpr int carrier_span(int rows) { return rows + 5; }
constexpr int kMaxStripRows = 32;
static_assert(carrier_span(kMaxStripRows) <= 64, "strip carrier rows must fit one wave");

// Solver strips of `rows` lines covering an h x w plane.
HVD_HD inline int solver_strips(int h, int w, int rows) {
  return ((w + kStripCols - 1) / kStripCols) * ((h + rows - 1) / rows);
}

// Block index -> logical tile, grouping consecutive logical tiles on one XCD
// (blocks are dispatched round-robin over the 8 XCDs, each with its own L2),
// so a tile and its vertical neighbours tend to share an L2.
HVD_HD inline int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int xcd = b % 8, idx = b / 8;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Field owning logical tile t: tile_prefix has n+1 entries.
HVD_HD inline int field_of_tile(const int* tile_prefix, int n, int t) {
  int lo = 0, hi = n;  // invariant: tile_prefix[lo] <= t < tile_prefix[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_prefix[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// Column of lane `lane` in strip lt of a field `tiles_x` strips across
// (-2, -1 or >= w on halo lanes), and the strip's first output row.
HVD_HD inline int strip_col(int lt, int tiles_x, int lane) { return (lt % tiles_x) * kStripCols + lane - kStripHalo; }
HVD_HD inline int strip_row0(int lt, int tiles_x, int rows) { return (lt / tiles_x) * rows; }
// Strips on odd strip rows walk their rows bottom-up.
HVD_HD inline bool strip_walks_up(int lt, int tiles_x) { return ((lt / tiles_x) & 1) != 0; }

// Index of (y, x) clamped into an h x w plane: the walk issues its loads a
// line ahead without branches, so rows and columns outside the plane read
// the plane's edge.
HVD_HD inline unsigned clamped_index(int y, int x, int h, int w) {
  const int yy = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
  const int xx = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
  return (unsigned)(yy * w + xx);
}

// The top-down walk's last line: lines 0 .. down_walk_last lie inside the
// plane (at most `rows`: the walk forms the gradient one line past the
// strip's last output row).
HVD_HD inline int down_walk_last(int y0, int rows, int h) {
  const int a = h - 1 - y0;
  return rows < a ? rows : a;
}
// The bottom-up walk's first row T (its gradient feeds the curvature of the
// line above); the walk covers rows T .. y0.
HVD_HD inline int up_walk_top(int y0, int rows, int h) { return y0 + rows < h ? y0 + rows : h; }

}  // namespace hvdk
