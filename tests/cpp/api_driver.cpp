// api_driver.cpp -- drives the hvd-core C++ API (libhvd_core.so) for the
// parity tests in tests/test_gpu_cpp_api.py: reads raw inputs, runs one API
// entry point, writes raw outputs.  Geometry and levels are the synthetic
// NTSC / PAL frames of hvd_amd.synth (FRAME_NTSC / FRAME_PAL).
//
//   api_driver window <ntsc|pal> <nfields> <in.f32> <out>
//       PrepareFieldObs + HvdEngine::DecodeSequenceWindow on nfields full
//       IRE fields; out: per field luma (f32) then chi (interleaved f32).
//   api_driver frames <ntsc|pal> <nframes> <core_begin> <core_end> <in.s16> <out>
//       DecodeFrameSequenceWindow; out: per core frame luma, chroma (s16),
//       u, v (f64).
//   api_driver frames3d <ntsc|pal> <nframes> <core_begin> <core_end> <strength> <in.s16> <out>
//       DecodeFrameSequenceWindow with the 3-D driver (cg_iterations 20,
//       fast mode, the given temporal strength); out as for frames.
//   api_driver window3d <ntsc|pal> <h> <w> <nfields> <strength> <dir>
//       DecodeFieldWindowWithInits on dir/s, dir/car, dir/chi0 (f32,
//       complex interleaved, fields one after another) and dir/par (int32
//       parities), cg_iterations 30; out dir/out: per field luma then chi,
//       then the five SequenceDiagnostics floats.
//   api_driver frame <ntsc|pal> <in.s16> <out>          DecodeFrameBuffer
//   api_driver ycframe <ntsc|pal> <luma.s16> <chroma.s16> <out>  DecodeYcFrameBuffer
//   api_driver refine3d <ntsc|pal> <h> <w> <nn> <nu> <eps_t> <dir>
//       VariationalRefine with nn temporal neighbours (cg_iterations 40,
//       irls_outer 4, slow mode, chroma_aniso 0.7) on dir/s, dir/car,
//       dir/chi0, dir/sw, dir/cw, dir/conf (f32, complex interleaved, the
//       neighbours' planes one after another); out dir/out: luma then chi.
//   api_driver mcprev <ntsc|pal> <h> <w> <th> <tw> <tile> <fast> <dir>
//       MotionCompensatePrev with the given tile motion field (dir/comp,
//       dir/car, dir/mdy, dir/mdx, dir/mconf) and ComplexCoherence(r = 3) of
//       dir/z1, dir/z2; out dir/out: composite, carrier, confidence, gamma.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine/engine.h"
#include "engine/gpu_status.h"
#include "engine/sequence.h"
#include "engine/temporal.h"
#include "engine/variational.h"
#include "frame_bridge.h"

namespace {

hvd::FrameParams Params(const std::string& std) {
  hvd::FrameParams fp;
  if (std == "pal") {
    fp.standard = hvd::VideoStandard::kPal;
    fp.frame_width = 1135;
    fp.frame_height = 625;
    fp.field1_lines = 313;
    fp.active_video_start = 185;
    fp.active_video_end = 1107;
    fp.colour_burst_start = 98;
    fp.colour_burst_end = 138;
    fp.first_active_frame_line = 44;
    fp.last_active_frame_line = 620;
    fp.black_level = 256.0F;
    fp.white_level = 844.0F;
    fp.blanking_level = 256.0F;
    fp.chroma_dc = 512.0F;
    fp.sample_rate = hvd::kFs4FscPal;
    const int extra[4] = {312, 312, 624, 624};
    for (int i = 0; i < 4; ++i) fp.extra_sample_lines[i] = extra[i];
    fp.extra_sample_line_count = 4;
  } else {
    fp.standard = hvd::VideoStandard::kNtsc;
    fp.frame_width = 910;
    fp.frame_height = 525;
    fp.field1_lines = 263;
    fp.active_video_start = 134;
    fp.active_video_end = 894;
    fp.colour_burst_start = 76;
    fp.colour_burst_end = 112;
    fp.first_active_frame_line = 42;
    fp.last_active_frame_line = 0;
    fp.black_level = 282.0F;
    fp.white_level = 800.0F;
    fp.blanking_level = 240.0F;
    fp.chroma_dc = 512.0F;
    fp.sample_rate = hvd::kFs4Fsc;
  }
  return fp;
}

template <typename T>
std::vector<T> ReadAll(const char* path, size_t n) {
  std::vector<T> v(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(T), n, f) != n) {
    std::fprintf(stderr, "cannot read %zu items from %s\n", n, path);
    std::exit(2);
  }
  std::fclose(f);
  return v;
}

struct Out {
  FILE* f;
  explicit Out(const char* path) : f(std::fopen(path, "wb")) {
    if (!f) std::exit(2);
  }
  ~Out() { std::fclose(f); }
  template <typename T>
  void put(const T* p, size_t n) {
    std::fwrite(p, sizeof(T), n, f);
  }
};

void PutFrame(Out& o, const hvd::YcFrameS16& y) {
  o.put(y.luma.data(), y.luma.size());
  o.put(y.chroma.data(), y.chroma.size());
  o.put(y.u_plane.data(), y.u_plane.size());
  o.put(y.v_plane.data(), y.v_plane.size());
}

int Fail() {
  std::fprintf(stderr, "api_driver: %s\n", hvd::GpuLastError().c_str());
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string mode = argv[1];
  const hvd::FrameParams fp = Params(argv[2]);
  const hvd::FieldGeometry g = hvd::FieldGeometryFromParams(fp);
  const size_t frame_n = static_cast<size_t>(hvd::FrameLineOffset(fp, fp.frame_height));
  hvd::HvdEngine engine;
  std::printf("api_driver %s %s\n", mode.c_str(), argv[2]);
  std::fflush(stdout);

  if (mode == "window" && argc == 6) {
    hvd::HvdConfig cfg;
    cfg.cg_iterations = 100;
    cfg.irls_outer = 4;
    cfg.fast = false;
    cfg.cg_tol = 1e-12F;
    cfg.temporal_strength = -1.0F;
    const int nf = std::atoi(argv[3]);
    const size_t fsz = static_cast<size_t>(g.field_height) * g.field_width;
    const std::vector<float> in = ReadAll<float>(argv[4], fsz * nf);
    std::vector<hvd::FieldObs> obs;
    for (int j = 0; j < nf; ++j) {
      hvd::FieldInput fi;
      fi.samples = hvd::Plane(g.field_height, g.field_width);
      std::memcpy(fi.samples.data(), in.data() + fsz * j, fsz * sizeof(float));
      obs.push_back(hvd::PrepareFieldObs(fi, g, cfg, j % 2));
      if (obs.back().s.empty()) return Fail();
    }
    const std::vector<hvd::DecodedField> dec = engine.DecodeSequenceWindow(obs, g, cfg);
    if (static_cast<int>(dec.size()) != nf) return Fail();
    Out o(argv[5]);
    for (const hvd::DecodedField& d : dec) {
      o.put(d.luma.data(), d.luma.size());
      o.put(reinterpret_cast<const float*>(d.chroma.data()), d.chroma.size() * 2);
    }
    return 0;
  }
  if (mode == "frames" && argc == 8) {
    hvd::HvdConfig cfg;
    cfg.cg_iterations = 40;
    cfg.irls_outer = 4;
    cfg.fast = false;
    cfg.cg_tol = 1e-6F;
    cfg.temporal_strength = -1.0F;
    const int nframes = std::atoi(argv[3]);
    const std::vector<int16_t> in = ReadAll<int16_t>(argv[6], frame_n * nframes);
    std::vector<const int16_t*> frames;
    for (int t = 0; t < nframes; ++t) frames.push_back(in.data() + frame_n * t);
    const std::vector<hvd::YcFrameS16> out =
        hvd::DecodeFrameSequenceWindow(frames, std::atoi(argv[4]), std::atoi(argv[5]), fp, cfg, engine);
    if (out.empty()) return Fail();
    Out o(argv[7]);
    for (const hvd::YcFrameS16& y : out) PutFrame(o, y);
    return 0;
  }
  if (mode == "frames3d" && (argc == 9 || argc == 10)) {
    hvd::HvdConfig cfg;
    cfg.cg_iterations = 20;
    cfg.irls_outer = 4;
    cfg.fast = true;
    cfg.temporal_strength = (float)std::atof(argv[6]);
    if (argc == 10) {  // the selective window, with this area cap
      cfg.selective_3d = true;
      cfg.selective_max_area = (float)std::atof(argv[9]);
    }
    const int nframes = std::atoi(argv[3]);
    const std::vector<int16_t> in = ReadAll<int16_t>(argv[7], frame_n * nframes);
    std::vector<const int16_t*> frames;
    for (int t = 0; t < nframes; ++t) frames.push_back(in.data() + frame_n * t);
    const std::vector<hvd::YcFrameS16> out =
        hvd::DecodeFrameSequenceWindow(frames, std::atoi(argv[4]), std::atoi(argv[5]), fp, cfg, engine);
    if (out.empty()) return Fail();
    Out o(argv[8]);
    for (const hvd::YcFrameS16& y : out) PutFrame(o, y);
    return 0;
  }
  if (mode == "window3d" && argc == 8) {
    const int h = std::atoi(argv[3]), w = std::atoi(argv[4]), nf = std::atoi(argv[5]);
    const size_t n = (size_t)h * w;
    const std::string dir = argv[7];
    const std::vector<float> s = ReadAll<float>((dir + "/s").c_str(), nf * n);
    const std::vector<float> car = ReadAll<float>((dir + "/car").c_str(), 2 * nf * n);
    const std::vector<float> chi0 = ReadAll<float>((dir + "/chi0").c_str(), 2 * nf * n);
    const std::vector<int32_t> par = ReadAll<int32_t>((dir + "/par").c_str(), nf);
    std::vector<hvd::FieldObs> obs(nf);
    std::vector<hvd::DecodedField> inits(nf);
    for (int j = 0; j < nf; ++j) {
      obs[j].s = hvd::Plane(h, w);
      std::memcpy(obs[j].s.data(), s.data() + j * n, sizeof(float) * n);
      obs[j].carrier = hvd::ComplexPlane(h, w);
      std::memcpy(reinterpret_cast<float*>(obs[j].carrier.data()), car.data() + 2 * j * n, sizeof(float) * 2 * n);
      obs[j].parity = par[j];
      inits[j].chroma = hvd::ComplexPlane(h, w);
      std::memcpy(reinterpret_cast<float*>(inits[j].chroma.data()), chi0.data() + 2 * j * n, sizeof(float) * 2 * n);
    }
    hvd::HvdConfig cfg;
    cfg.cg_iterations = 30;
    cfg.temporal_strength = (float)std::atof(argv[6]);
    hvd::SequenceDiagnostics d;
    const std::vector<hvd::DecodedField> dec = hvd::DecodeFieldWindowWithInits(obs, inits, g, cfg, &d);
    if (static_cast<int>(dec.size()) != nf) return Fail();
    Out o((dir + "/out").c_str());
    for (const hvd::DecodedField& f : dec) {
      o.put(f.luma.data(), n);
      o.put(reinterpret_cast<const float*>(f.chroma.data()), 2 * n);
    }
    const float dv[5] = {d.resolved_strength, d.ambiguity_ire, d.sigma_ire, d.temporal_eps, d.nr_eps};
    o.put(dv, 5);
    return 0;
  }
  if (mode == "frame" && argc == 5) {
    hvd::HvdConfig cfg;
    cfg.cg_iterations = 40;
    cfg.irls_outer = 4;
    cfg.fast = false;
    cfg.cg_tol = 1e-6F;
    const std::vector<int16_t> in = ReadAll<int16_t>(argv[3], frame_n);
    const hvd::YcFrameS16 y = hvd::DecodeFrameBuffer(in.data(), fp, cfg, engine);
    if (y.luma.empty()) return Fail();
    Out o(argv[4]);
    PutFrame(o, y);
    return 0;
  }
  if (mode == "ycframe" && argc == 6) {
    hvd::HvdConfig cfg;
    const std::vector<int16_t> lum = ReadAll<int16_t>(argv[3], frame_n);
    const std::vector<int16_t> chr = ReadAll<int16_t>(argv[4], frame_n);
    const hvd::YcFrameS16 y = hvd::DecodeYcFrameBuffer(lum.data(), chr.data(), fp, cfg, engine);
    if (y.luma.empty()) return Fail();
    Out o(argv[5]);
    PutFrame(o, y);
    return 0;
  }
  if (mode == "refine3d" && argc == 9) {
    const int h = std::atoi(argv[3]), w = std::atoi(argv[4]), nn = std::atoi(argv[5]);
    const size_t n = (size_t)h * w;
    const std::string dir = argv[8];
    auto plane = [&](const std::vector<float>& v, size_t off) {
      hvd::Plane p(h, w);
      std::memcpy(p.data(), v.data() + off, sizeof(float) * n);
      return p;
    };
    auto cplane = [&](const std::vector<float>& v, size_t off) {
      hvd::ComplexPlane p(h, w);
      std::memcpy(reinterpret_cast<float*>(p.data()), v.data() + off, sizeof(float) * 2 * n);
      return p;
    };
    const std::vector<float> s = ReadAll<float>((dir + "/s").c_str(), n);
    const std::vector<float> car = ReadAll<float>((dir + "/car").c_str(), 2 * n);
    const std::vector<float> chi0 = ReadAll<float>((dir + "/chi0").c_str(), 2 * n);
    const std::vector<float> sw = ReadAll<float>((dir + "/sw").c_str(), nn * n);
    const std::vector<float> cw = ReadAll<float>((dir + "/cw").c_str(), 2 * nn * n);
    const std::vector<float> conf = ReadAll<float>((dir + "/conf").c_str(), nn * n);
    std::vector<hvd::NeighborTerm> nbrs(nn);
    for (int j = 0; j < nn; ++j) {
      nbrs[j].composite = plane(sw, j * n);
      nbrs[j].carrier = cplane(cw, 2 * j * n);
      nbrs[j].confidence = plane(conf, j * n);
    }
    hvd::HvdConfig cfg;
    cfg.cg_iterations = 40;
    cfg.irls_outer = 4;
    cfg.fast = false;
    cfg.chroma_aniso = 0.7F;
    cfg.temporal_strength = (float)std::atof(argv[6]);
    cfg.temporal_eps = (float)std::atof(argv[7]);
    const hvd::RefineResult r = hvd::VariationalRefine(plane(s, 0), cplane(car, 0), cplane(chi0, 0), cfg, nbrs);
    if (r.luma.empty()) return Fail();
    Out o((dir + "/out").c_str());
    o.put(r.luma.data(), n);
    o.put(reinterpret_cast<const float*>(r.chroma.data()), 2 * n);
    return 0;
  }
  if (mode == "mcprev" && argc == 10) {
    const int h = std::atoi(argv[3]), w = std::atoi(argv[4]), th = std::atoi(argv[5]), tw = std::atoi(argv[6]);
    const int tile = std::atoi(argv[7]);
    const bool fast = std::atoi(argv[8]) != 0;
    const std::string dir = argv[9];
    const size_t n = (size_t)h * w, nt = (size_t)th * tw;
    auto load = [&](const char* name, int ph, int pw, int ch) {
      const std::vector<float> v = ReadAll<float>((dir + "/" + name).c_str(), (size_t)ph * pw * ch);
      return v;
    };
    auto plane = [](const std::vector<float>& v, int ph, int pw) {
      hvd::Plane p(ph, pw);
      std::memcpy(p.data(), v.data(), sizeof(float) * p.size());
      return p;
    };
    auto cplane = [](const std::vector<float>& v, int ph, int pw) {
      hvd::ComplexPlane p(ph, pw);
      std::memcpy(reinterpret_cast<float*>(p.data()), v.data(), sizeof(float) * 2 * p.size());
      return p;
    };
    hvd::NeighborRawState prev;
    prev.luma = hvd::Plane(h, w);
    prev.composite = plane(load("comp", h, w, 1), h, w);
    prev.carrier = cplane(load("car", h, w, 2), h, w);
    hvd::MotionField m;
    m.tile = tile;
    m.dy = plane(load("mdy", th, tw, 1), th, tw);
    m.dx = plane(load("mdx", th, tw, 1), th, tw);
    m.confidence = plane(load("mconf", th, tw, 1), th, tw);
    const hvd::MotionCompensatedResult r = hvd::MotionCompensatePrev(prev, hvd::Plane(h, w), tile, 8, &m, fast);
    const hvd::Plane g =
        hvd::ComplexCoherence(cplane(load("z1", h, w, 2), h, w), cplane(load("z2", h, w, 2), h, w), 3);
    if (r.composite.empty() || g.empty()) return Fail();
    (void)nt;
    Out o((dir + "/out").c_str());
    o.put(r.composite.data(), n);
    o.put(reinterpret_cast<const float*>(r.carrier.data()), 2 * n);
    o.put(r.confidence.data(), n);
    o.put(g.data(), n);
    return 0;
  }
  std::fprintf(stderr, "usage: see the header of api_driver.cpp\n");
  return 2;
}
