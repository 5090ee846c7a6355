// cli.cpp - fluid-simulator, the headless counterpart of the reference's
// main loop (src/main.cu:18-107): load OpenSayal.conf.json, create the
// solver, call Fluid::update once per frame with the frame's d_t, and in
// place of the SDL window write what GraphicsHandler would have drawn and
// the fields themselves to .npy files.  The reference's --save-report /
// --compare-performance (inc/logger.hpp:58-95, a frame time in report.log)
// become a JSON performance record.
//
// The program talks to the solver only through the reference-shaped C++
// wrapper (include/opensayal.hpp) over the C ABI of libopensayal_amd.so.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include <memory>

#include "../../include/opensayal.hpp"
#include "json_value.hpp"

namespace {

const char* kUsage =
    "usage: fluid-simulator [--config PATH] [--steps N] [--dump DIR] [--dump-every K]\n"
    "                       [--frame] [--traces] [--log-every K]\n"
    "                       [--save-report | --compare-performance] [--report PATH]\n";

struct Options {
  std::string config = "OpenSayal.conf.json";  // ConfigParser's default file
  long steps = 100;
  std::string dump_dir;
  long dump_every = 0;
  bool frame = false;
  bool traces = false;
  long log_every = 0;
  bool save_report = false;
  bool compare = false;
  std::string report = "report.json";
};

int usage_error(const std::string& msg) {
  std::fprintf(stderr, "Error: %s\n%s", msg.c_str(), kUsage);
  return 1;
}

bool parse_long(const char* s, long lo, long* out) {
  char* end = nullptr;
  errno = 0;
  long v = std::strtol(s, &end, 10);
  if (errno || !end || *end || end == s || v < lo) return false;
  *out = v;
  return true;
}

int parse_args(int argc, char** argv, Options* o) {
  for (int a = 1; a < argc; a++) {
    const std::string k = argv[a];
    auto value = [&](const char** v) {
      if (a + 1 >= argc) return false;
      *v = argv[++a];
      return true;
    };
    const char* v = nullptr;
    if (k == "--config") {
      if (!value(&v)) return usage_error("--config needs a path");
      o->config = v;
    } else if (k == "--steps") {
      if (!value(&v) || !parse_long(v, 0, &o->steps)) return usage_error("--steps needs N >= 0");
    } else if (k == "--dump") {
      if (!value(&v)) return usage_error("--dump needs a directory");
      o->dump_dir = v;
    } else if (k == "--dump-every") {
      if (!value(&v) || !parse_long(v, 1, &o->dump_every))
        return usage_error("--dump-every needs K >= 1");
    } else if (k == "--frame") {
      o->frame = true;
    } else if (k == "--traces") {
      o->traces = true;
    } else if (k == "--log-every") {
      if (!value(&v) || !parse_long(v, 1, &o->log_every))
        return usage_error("--log-every needs K >= 1");
    } else if (k == "--save-report") {
      o->save_report = true;
    } else if (k == "--compare-performance") {
      o->compare = true;
    } else if (k == "--report") {
      if (!value(&v)) return usage_error("--report needs a path");
      o->report = v;
    } else if (k == "--help" || k == "-h") {
      std::fputs(kUsage, stdout);
      std::exit(0);
    } else {
      return usage_error("unknown argument " + k);
    }
  }
  // the reference takes at most one of the two report flags (src/main.cu:19-25)
  if (o->save_report && o->compare)
    return usage_error("--save-report and --compare-performance are exclusive");
  if (o->dump_every && o->dump_dir.empty()) return usage_error("--dump-every needs --dump");
  if ((o->frame || o->traces) && o->dump_dir.empty())
    return usage_error("--frame and --traces write into the --dump directory");
  return 0;
}

// visual.path_line of the config (src/config_parser.cpp:149-181).
struct PathLine {
  bool enable = false;
  int length = 20;
  int distance = 20;
};

int read_path_line(const std::string& path, PathLine* pl) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return 1;
  std::ostringstream ss;
  ss << f.rdbuf();
  try {
    const osy::Json j = osy::Json::parse(ss.str());
    auto as_int = [](const osy::Json& v) { return (int)v.to_int64(); };
    auto as_bool = [](const osy::Json& v) { return v.to_bool(); };
    pl->enable = osy::json_get_or<bool>(j, "visual.path_line.enable", false, as_bool);
    pl->length = osy::json_get_or<int>(j, "visual.path_line.length", 20, as_int);
    pl->distance = osy::json_get_or<int>(j, "visual.path_line.distance", 20, as_int);
  } catch (const osy::JsonError& e) {
    std::fprintf(stderr, "Error: config: %s\n", e.what());
    return 1;
  }
  return 0;
}

// NumPy .npy, format version 1.0: magic, version, little-endian header
// length, a Python dict literal padded with spaces to a multiple of 64 bytes
// ending in '\n', then the raw C-order data.
bool write_npy(const std::string& path, const char* descr, const std::vector<long>& shape,
               const void* data, size_t bytes) {
  std::string dims;
  for (size_t k = 0; k < shape.size(); k++) dims += std::to_string(shape[k]) + ", ";
  if (shape.size() > 1) dims.resize(dims.size() - 2);
  else if (shape.size() == 1) dims.resize(dims.size() - 1);
  std::string header = std::string("{'descr': '") + descr +
                       "', 'fortran_order': False, 'shape': (" + dims + "), }";
  const size_t pre = 10;  // magic (6) + version (2) + header length (2)
  size_t total = pre + header.size() + 1;
  header.append((64 - total % 64) % 64, ' ');
  header.push_back('\n');
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const uint16_t hl = (uint16_t)header.size();
  const unsigned char hlb[2] = {(unsigned char)(hl & 255), (unsigned char)(hl >> 8)};
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hlb, 1, 2, f) == 2 &&
            std::fwrite(header.data(), 1, header.size(), f) == header.size() &&
            (bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes);
  return std::fclose(f) == 0 && ok;
}

// Binary PPM (P6) of an RGBA8888 frame, alpha dropped: a picture any viewer
// opens, for looking at a run without a window.
bool write_ppm(const std::string& path, const std::vector<uint32_t>& px, int W, int H) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::fprintf(f, "P6\n%d %d\n255\n", W, H);
  std::vector<unsigned char> rgb((size_t)W * H * 3);
  for (size_t k = 0; k < px.size(); k++) {
    rgb[3 * k] = (unsigned char)(px[k] >> 24);
    rgb[3 * k + 1] = (unsigned char)(px[k] >> 16);
    rgb[3 * k + 2] = (unsigned char)(px[k] >> 8);
  }
  bool ok = std::fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
  return std::fclose(f) == 0 && ok;
}

int fail(const char* what, const std::exception& e) {
  std::fprintf(stderr, "Error: %s: %s\n", what, e.what());
  return 1;
}

struct Dumper {
  const Options& o;
  const opensayal::Config& cfg;
  const PathLine& pl;
  opensayal::Fluid& fl;

  std::string name(long step, const char* field, const char* ext) const {
    char b[64];
    std::snprintf(b, sizeof b, "step_%06ld_%s.%s", step, field, ext);
    return o.dump_dir + "/" + b;
  }

  int dump(long step, float d_t) const {
    const int W = cfg.width, H = cfg.height;
    const size_t bytes = (size_t)W * H * 4;
    struct F { osy_field id; const char* name; bool on; };
    const F fields[] = {{OSY_FIELD_VEL_X, "vel_x", true},
                        {OSY_FIELD_VEL_Y, "vel_y", true},
                        {OSY_FIELD_SMOKE, "smoke", true},
                        {OSY_FIELD_PRESSURE, "pressure", cfg.enable_pressure != 0}};
    for (const F& f : fields) {
      if (!f.on) continue;
      const std::vector<float> buf = fl.field(f.id);
      if (!write_npy(name(step, f.name, "npy"), "<f4", {H, W}, buf.data(), bytes)) {
        std::fprintf(stderr, "Error: cannot write %s of step %ld\n", f.name, step);
        return 1;
      }
    }
    if (o.frame) {
      const std::vector<uint32_t> px = fl.render_rgba();
      if (!write_npy(name(step, "frame", "npy"), "<u4", {H, W}, px.data(), bytes) ||
          !write_ppm(name(step, "frame", "ppm"), px, W, H)) {
        std::fprintf(stderr, "Error: cannot write the frame of step %ld\n", step);
        return 1;
      }
    }
    if (o.traces || pl.enable) {
      const long th = H / pl.distance, tw = W / pl.distance, L = pl.length;
      std::vector<int32_t> tx, ty;
      fl.traces(pl.distance, pl.length, d_t, tx, ty);
      const size_t each = tx.size() * 4;
      if (!write_npy(name(step, "traces_x", "npy"), "<i4", {th, tw, L}, tx.data(), each) ||
          !write_npy(name(step, "traces_y", "npy"), "<i4", {th, tw, L}, ty.data(), each)) {
        std::fprintf(stderr, "Error: cannot write the path lines of step %ld\n", step);
        return 1;
      }
    }
    return 0;
  }
};

std::string json_escape(const std::string& s) {
  std::string out;
  for (char c : s) {
    if (c == '"' || c == '\\') {
      out.push_back('\\');
      out.push_back(c);
    } else if ((unsigned char)c < 0x20) {
      char b[8];
      std::snprintf(b, sizeof b, "\\u%04x", (unsigned)(unsigned char)c);
      out += b;
    } else {
      out.push_back(c);
    }
  }
  return out;
}

// The previous record's frame time, or -1 when there is none.
double previous_frame_us(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return -1.0;
  std::ostringstream ss;
  ss << f.rdbuf();
  try {
    const osy::Json j = osy::Json::parse(ss.str());
    return osy::json_get_or<double>(j, "avg_frame_us", -1.0,
                                    [](const osy::Json& v) { return v.to_double(); });
  } catch (const osy::JsonError&) {
    return -1.0;
  }
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  if (int rc = parse_args(argc, argv, &o)) return rc;

  opensayal::Config cfg;
  try {
    cfg = opensayal::ConfigParser(o.config).parse();
  } catch (const opensayal::Error& e) {
    return fail("config", e);
  }
  PathLine pl;
  if (read_path_line(o.config, &pl)) return 1;
  if ((o.traces || pl.enable) && (pl.distance < 1 || pl.length < 1 || pl.length > 65536)) {
    std::fprintf(stderr, "Error: visual.path_line needs distance >= 1 and length in [1, 65536]\n");
    return 1;
  }
  if (!o.dump_dir.empty()) {
    struct stat st;
    if (stat(o.dump_dir.c_str(), &st) != 0 && mkdir(o.dump_dir.c_str(), 0755) != 0) {
      std::fprintf(stderr, "Error: cannot create %s\n", o.dump_dir.c_str());
      return 1;
    }
  }

  std::unique_ptr<opensayal::Fluid> fluid;
  try {
    fluid = std::make_unique<opensayal::Fluid>(cfg);
  } catch (const opensayal::Error& e) {
    return fail("create", e);
  }
  opensayal::Fluid& fl = *fluid;
  Dumper dumper{o, cfg, pl, fl};
  try {
    // The frame loop of src/main.cu:55-103 without events: the source stays
    // inactive (no mouse), d_t is sim.time.d_t or, with enable_real_time, the
    // previous frame's wall time times real_time_multiplier (0 on the first).
    using clk = std::chrono::steady_clock;
    const double cells = (double)cfg.width * cfg.height;
    std::vector<double> frame_us;  // the last 10 frame times (FPS_AVG_SIZE)
    frame_us.reserve(o.steps);
    float d_t = cfg.d_t;
    const auto t0 = clk::now();
    auto prev = t0;
    for (long s = 1; s <= o.steps; s++) {
      const auto now = clk::now();
      if (s == 1) prev = now;  // the first frame has passed_time 0
      const double passed_s = std::chrono::duration<double>(now - prev).count();
      d_t = cfg.enable_real_time ? (float)passed_s * cfg.real_time_multiplier : cfg.d_t;
      fl.update(opensayal::Source{}, d_t);
      const auto end = clk::now();
      frame_us.push_back(std::chrono::duration<double, std::micro>(end - now).count());
      prev = now;
      if (o.log_every && s % o.log_every == 0) {
        const size_t n = frame_us.size() < 10 ? frame_us.size() : 10;
        double avg = 0;
        for (size_t k = frame_us.size() - n; k < frame_us.size(); k++) avg += frame_us[k];
        avg /= (double)n;
        std::printf("step %ld: FPS: %d, DT: %.0fus\n", s, avg > 0 ? (int)(1e6 / avg) : 0, avg);
        std::fflush(stdout);
      }
      if (o.dump_every && s % o.dump_every == 0 && s != o.steps)
        if (int rc = dumper.dump(s, d_t)) return rc;
    }
    const double wall_s = std::chrono::duration<double>(clk::now() - t0).count();
    if (!o.dump_dir.empty())
      if (int rc = dumper.dump(o.steps, d_t)) return rc;

    const size_t n = frame_us.size() < 10 ? frame_us.size() : 10;
    double avg_frame_us = 0;
    for (size_t k = frame_us.size() - n; k < frame_us.size(); k++) avg_frame_us += frame_us[k];
    if (n) avg_frame_us /= (double)n;
    const double ms_per_step = o.steps ? wall_s * 1e3 / (double)o.steps : 0.0;
    const double gcell = wall_s > 0 ? cells * (double)o.steps / wall_s / 1e9 : 0.0;
    osy_fluid_info info;
    opensayal::check(osy_fluid_get_info(fl.handle(), &info));
    static const char* kStorage[] = {"f32", "f16", "bf16"};
    char rec[1024];
    std::snprintf(rec, sizeof rec,
                  "{\"config\": \"%s\", \"width\": %d, \"height\": %d, \"projection_n\": %d, "
                  "\"storage\": \"%s\", \"steps\": %ld, \"real_time\": %s, \"last_d_t\": %.9g, "
                  "\"wall_s\": %.6f, \"ms_per_step\": %.6f, \"gcell_steps_per_s\": %.6f, "
                  "\"avg_frame_us\": %.3f, \"fps_avg\": %d, \"device_bytes\": %zu}",
                  json_escape(o.config).c_str(), cfg.width, cfg.height, cfg.projection_n,
                  kStorage[info.storage < 3 ? info.storage : 0], o.steps,
                  cfg.enable_real_time ? "true" : "false", (double)d_t, wall_s, ms_per_step, gcell,
                  avg_frame_us, avg_frame_us > 0 ? (int)(1e6 / avg_frame_us) : 0,
                  info.device_bytes);
    std::printf("%s\n", rec);
    if (o.save_report) {
      std::ofstream f(o.report);
      if (!f) {
        std::fprintf(stderr, "Error: couldn't open report file %s\n", o.report.c_str());
        return 1;
      }
      f << rec << "\n";
    }
    if (o.compare) {  // Logger::report (inc/logger.hpp:68-86)
      const double before = previous_frame_us(o.report);
      if (before <= 0) {
        std::fprintf(stderr, "Warning: no previous record in %s to compare with\n",
                     o.report.c_str());
      } else {
        std::printf("performance gain: %g%%\n", (before - avg_frame_us) / before * 100.0);
      }
    }
    return 0;
  } catch (const opensayal::Error& e) {
    return fail("solver", e);
  }
}
