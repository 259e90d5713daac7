// frame_bridge.h -- the frame-buffer bridge of the hvd-core C++ API
// (hvd-core/src/frame_bridge.h) over the MI355X library: 10-bit
// field-sequential int16 frames in, lossless Y/C split frames out.  Same
// types, names and layout conventions as hvd-core's, so the decode-orc stage
// plugin builds against it unchanged.
#ifndef HVD_AMD_FRAME_BRIDGE_H_
#define HVD_AMD_FRAME_BRIDGE_H_

#include <cstdint>
#include <vector>

#include "engine/engine.h"
#include "engine/hvd_config.h"
#include "engine/ntsc_geometry.h"

namespace hvd {

// The frame parameters the decode needs, in the 10-bit domain
// (frame_bridge.h:40-101).
struct FrameParams {
  VideoStandard standard = VideoStandard::kNtsc;
  bool uses_vswitch() const { return standard == VideoStandard::kPal || standard == VideoStandard::kPalM; }

  int frame_width = 0;
  int frame_height = 0;
  int field1_lines = 0;
  int active_video_start = 0;
  int active_video_end = 0;
  int colour_burst_start = 0;
  int colour_burst_end = 0;
  int first_active_frame_line = 0;  // woven-frame convention: field line = frame line / 2
  int last_active_frame_line = 0;   // 0 => from the field height
  float black_level = 0.0F;
  float white_level = 0.0F;
  float blanking_level = 0.0F;
  float chroma_dc = 0.0F;
  double sample_rate = kFs4Fsc;
  double subcarrier_hz = 0.0;

  // Frame-flat lines carrying one extra sample each listing (PAL: 312 and
  // 624, twice each); empty => uniform stride.
  static constexpr int kMaxExtraSampleLines = 8;
  int32_t extra_sample_lines[kMaxExtraSampleLines] = {};
  int extra_sample_line_count = 0;
};

// Offset of frame-flat line `line` in the source buffer (frame_bridge.h:106-112).
inline int64_t FrameLineOffset(const FrameParams& fp, int line) {
  int64_t offset = static_cast<int64_t>(line) * fp.frame_width;
  for (int i = 0; i < fp.extra_sample_line_count; ++i)
    if (fp.extra_sample_lines[i] < line) ++offset;
  return offset;
}

// A decoded Y/C frame, frame_width x frame_height flat (frame_bridge.h:123-131).
struct YcFrameS16 {
  int width = 0;
  int height = 0;
  float chroma_dc = 0.0F;
  std::vector<int16_t> luma;
  std::vector<int16_t> chroma;
  std::vector<double> u_plane;
  std::vector<double> v_plane;
};

FieldGeometry FieldGeometryFromParams(const FrameParams& fp);

// Woven-frame 2-D decode of one frame (frame_bridge.cpp:161-255).
YcFrameS16 DecodeFrameBuffer(const int16_t* frame, const FrameParams& fp, const HvdConfig& cfg, HvdEngine& engine);

// Block 1 / block 2 of a field-sequential frame as IRE fields
// (frame_bridge.cpp:61-99).
void SplitFrameFields(const int16_t* frame, const FrameParams& fp, FieldInput* top, FieldInput* bottom,
                      bool field1_is_bottom = false);

// +1 field 1 top, -1 field 1 bottom, 0 no confident vote (frame_bridge.cpp:102-149).
int DetectFieldParity(const FieldInput& block1, const FieldInput& block2, const FrameParams& fp);

// HvdConfig::field_order for one frame (frame_bridge.cpp:151-159).
bool ResolveField1IsBottom(const FieldInput& block1, const FieldInput& block2, const FrameParams& fp,
                           const HvdConfig& cfg);

// Field-granularity decode of a window of frames; one YcFrameS16 per core
// frame [core_begin, core_end) (frame_bridge.cpp:722-947): the fields of all
// frames decode as one window, 3-D when cfg's temporal strength engages it
// (the stage passes a negative strength when enable_temporal is off).
std::vector<YcFrameS16> DecodeFrameSequenceWindow(const std::vector<const int16_t*>& frames, int core_begin,
                                                  int core_end, const FrameParams& fp, const HvdConfig& cfg,
                                                  HvdEngine& engine, int64_t base_frame_id = 0);

YcFrameS16 DecodeFrameBuffer(const int16_t* frame, const FrameParams& fp, const HvdConfig& cfg);

// Already separated Y/C sources: luma passes through, chroma is demodulated
// (frame_bridge.cpp:263-377).
YcFrameS16 DecodeYcFrameBuffer(const int16_t* luma, const int16_t* chroma, const FrameParams& fp,
                               const HvdConfig& cfg, HvdEngine& engine);
YcFrameS16 DecodeYcFrameBuffer(const int16_t* luma, const int16_t* chroma, const FrameParams& fp,
                               const HvdConfig& cfg);

}  // namespace hvd

#endif  // HVD_AMD_FRAME_BRIDGE_H_
