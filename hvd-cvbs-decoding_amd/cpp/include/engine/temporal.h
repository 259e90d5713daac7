// engine/temporal.h -- hvd-core's motion-compensated warps and coherence
// gate for the temporal neighbours (temporal.h:1-149), over the GPU library.
// Same names, types and semantics; failures return empty results with the
// reason in GpuLastError().
#ifndef HVD_AMD_ENGINE_TEMPORAL_H_
#define HVD_AMD_ENGINE_TEMPORAL_H_

#include "engine/motion.h"
#include "engine/plane.h"

namespace hvd {

// temporal.h:31-36: per-pixel vectors from a tile motion field (3x3-median
// outlier snap, then bilinear interpolation between tile centres).
struct PerPixelMotion {
  Plane dy;
  Plane dx;
};
PerPixelMotion VectorsPerPixel(const Plane& mdy, const Plane& mdx, int tile, int out_h, int out_w);

// temporal.h:45-50: integer-pel warp by the (interpolated, rounded) vectors.
Plane WarpByTiles(const Plane& a, const Plane& mdy, const Plane& mdx, int tile);
Plane WarpByTiles(const Plane& a, const PerPixelMotion& v);

// temporal.h:57-61: bilinear sub-pixel warp of a baseband plane.
Plane WarpBilinearTiles(const Plane& a, const Plane& dyf, const Plane& dxf, int tile, float row_offset, int out_h,
                        int out_w);
Plane WarpBilinearTiles(const Plane& a, const PerPixelMotion& v, float row_offset, int out_h, int out_w);

// temporal.h:68: local complex coherence of two phasor fields, in [0, 1].
Plane ComplexCoherence(const ComplexPlane& z1, const ComplexPlane& z2, int r);

// temporal.h:80-83: the integer-pel warp of complex data.
ComplexPlane WarpByTilesComplex(const ComplexPlane& a, const Plane& mdy, const Plane& mdx, int tile);
ComplexPlane WarpByTilesComplex(const ComplexPlane& a, const PerPixelMotion& v);

// temporal.h:90-101: tile confidence to per pixel (nearest, squared, blurred)
// and the fast form (squared tile map through VectorsPerPixel).
Plane UpsampleConfidence(const Plane& conf, int tile, int out_h, int out_w);
Plane UpsampleConfidenceFast(const Plane& conf, int tile, int out_h, int out_w);

// temporal.h:110-121: one neighbour frame's raw state.
struct NeighborRawState {
  Plane luma;
  Plane composite;
  ComplexPlane carrier;
  ComplexPlane chroma;
};

// temporal.h:130-134.
struct MotionCompensatedResult {
  Plane composite;
  ComplexPlane carrier;
  Plane confidence;
};

// temporal.h:139-142: warp a neighbour's composite and carrier onto the
// current field (motion estimated from prev.luma and y_cur_init when
// `motion` is null; `vpix` optional precomputed vectors).
MotionCompensatedResult MotionCompensatePrev(const NeighborRawState& prev, const Plane& y_cur_init, int tile,
                                             int search, const MotionField* motion = nullptr, bool fast = false,
                                             const PerPixelMotion* vpix = nullptr);

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_TEMPORAL_H_
