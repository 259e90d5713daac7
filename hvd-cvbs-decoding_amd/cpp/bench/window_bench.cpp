// window_bench.cpp -- times hvd::DecodeFrameSequenceWindow through
// libhvd_core.so as the decode-orc plugin calls it per chunk
// (orc-plugin_hvd/src/hvd_chroma_decoder_stage.cpp:700-760): host int16
// frames in, YcFrameS16 (luma, chroma, U, V planes) out, one HvdEngine
// reused across calls, HvdConfig defaults with the temporal switch mapped
// the way the plugin maps it (off: strength -1; on: the configured strength,
// 0 = AUTO).
//
//   hvd_window_bench <frames.s16> <nframes> <core_begin> <core_end>
//                    <temporal: off | auto | strength> <reps> [key=value ...]
//
// key=value set FrameParams fields (standard, frame_width, frame_height,
// field1_lines, active_video_start, active_video_end, colour_burst_start,
// colour_burst_end, first_active_frame_line, last_active_frame_line,
// black_level, white_level, blanking_level, chroma_dc, sample_rate,
// subcarrier_hz, extra_sample_lines=a,b,...) and HvdConfig's cg_iterations.
// Prints one JSON object: wall time per call (first call apart), frames.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine/engine.h"
#include "engine/gpu_status.h"
#include "frame_bridge.h"

namespace {

bool SetParam(hvd::FrameParams& fp, hvd::HvdConfig& cfg, const std::string& k, const std::string& v) {
  const double d = std::atof(v.c_str());
  const int i = std::atoi(v.c_str());
  if (k == "standard") fp.standard = static_cast<hvd::VideoStandard>(i);
  else if (k == "frame_width") fp.frame_width = i;
  else if (k == "frame_height") fp.frame_height = i;
  else if (k == "field1_lines") fp.field1_lines = i;
  else if (k == "active_video_start") fp.active_video_start = i;
  else if (k == "active_video_end") fp.active_video_end = i;
  else if (k == "colour_burst_start") fp.colour_burst_start = i;
  else if (k == "colour_burst_end") fp.colour_burst_end = i;
  else if (k == "first_active_frame_line") fp.first_active_frame_line = i;
  else if (k == "last_active_frame_line") fp.last_active_frame_line = i;
  else if (k == "black_level") fp.black_level = static_cast<float>(d);
  else if (k == "white_level") fp.white_level = static_cast<float>(d);
  else if (k == "blanking_level") fp.blanking_level = static_cast<float>(d);
  else if (k == "chroma_dc") fp.chroma_dc = static_cast<float>(d);
  else if (k == "sample_rate") fp.sample_rate = d;
  else if (k == "subcarrier_hz") fp.subcarrier_hz = d;
  else if (k == "cg_iterations") cfg.cg_iterations = i;
  else if (k == "extra_sample_lines") {
    fp.extra_sample_line_count = 0;
    size_t p = 0;
    while (p < v.size() && fp.extra_sample_line_count < hvd::FrameParams::kMaxExtraSampleLines) {
      const size_t q = v.find(',', p);
      const std::string item = v.substr(p, q == std::string::npos ? std::string::npos : q - p);
      if (!item.empty()) fp.extra_sample_lines[fp.extra_sample_line_count++] = std::atoi(item.c_str());
      if (q == std::string::npos) break;
      p = q + 1;
    }
  } else {
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s frames.s16 nframes core_begin core_end off|auto|strength reps [key=value ...]\n",
                 argv[0]);
    return 2;
  }
  const int nframes = std::atoi(argv[2]), core_begin = std::atoi(argv[3]), core_end = std::atoi(argv[4]);
  const std::string temporal = argv[5];
  const int reps = std::max(1, std::atoi(argv[6]));
  hvd::FrameParams fp;
  hvd::HvdConfig cfg;  // the library defaults, as the plugin's config starts
  for (int a = 7; a < argc; ++a) {
    const std::string kv = argv[a];
    const size_t eq = kv.find('=');
    if (eq == std::string::npos || !SetParam(fp, cfg, kv.substr(0, eq), kv.substr(eq + 1))) {
      std::fprintf(stderr, "unknown argument %s\n", argv[a]);
      return 2;
    }
  }
  // the plugin's mapping of its enable_temporal switch onto the strength
  if (temporal == "off") cfg.temporal_strength = -1.0F;
  else if (temporal == "auto") cfg.temporal_strength = 0.0F;
  else cfg.temporal_strength = static_cast<float>(std::atof(temporal.c_str()));

  const size_t frame_n = static_cast<size_t>(hvd::FrameLineOffset(fp, fp.frame_height));
  std::vector<int16_t> in(frame_n * nframes);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f || std::fread(in.data(), sizeof(int16_t), in.size(), f) != in.size()) {
    std::fprintf(stderr, "cannot read %zu samples from %s\n", in.size(), argv[1]);
    return 2;
  }
  std::fclose(f);
  std::vector<const int16_t*> frames;
  for (int t = 0; t < nframes; ++t) frames.push_back(in.data() + frame_n * t);

  hvd::HvdEngine engine;
  using clk = std::chrono::steady_clock;
  auto call = [&]() {
    const auto t0 = clk::now();
    const std::vector<hvd::YcFrameS16> out =
        hvd::DecodeFrameSequenceWindow(frames, core_begin, core_end, fp, cfg, engine);
    const double ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    if (out.size() != static_cast<size_t>(core_end - core_begin)) {
      std::fprintf(stderr, "DecodeFrameSequenceWindow failed: %s\n", hvd::GpuLastError().c_str());
      std::exit(1);
    }
    return ms;
  };
  const double first = call();
  double total = 0.0, best = 1e30;
  for (int r = 0; r < reps; ++r) {
    const double ms = call();
    total += ms;
    best = ms < best ? ms : best;
  }
  const double mean = total / reps;
  std::printf("{\"ms_per_call\": %.3f, \"ms_best_call\": %.3f, \"ms_first_call\": %.3f, \"frames\": %d, "
              "\"core_frames\": %d, \"core_frames_per_s\": %.2f, \"temporal\": \"%s\", \"cg_iterations\": %d}\n",
              mean, best, first, nframes, core_end - core_begin, (core_end - core_begin) * 1e3 / mean,
              temporal.c_str(), cfg.cg_iterations);
  return 0;
}
