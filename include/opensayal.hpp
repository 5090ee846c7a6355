/* opensayal.hpp - the reference's C++ surface for the solver, over the C ABI
 * of opensayal.h.
 *
 * The names and argument meanings are those of the reference's classes, so
 * a caller written against them changes its includes and little else:
 *   ConfigParser(file).parse()            src/config_parser.cpp:11-184
 *   Fluid(Config), update(Source, d_t)    src/fluid.cu:41-71, 770-795
 *   Fluid::width / height / min_pressure / max_pressure  inc/fluid.cuh:15-116
 *   Source { active, smoke, velocity, position }         inc/fluid.cuh:8-13
 * Config is the flat osy_config (the reference nests the same values as
 * config.sim.*, config.fluid.*).  Where the reference exits or crashes on an
 * error, these throw opensayal::Error with the library's message.  Header
 * only; link libopensayal_amd.so.
 */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "opensayal.h"

namespace opensayal {

using Config = osy_config;

class Error : public std::runtime_error {
 public:
  Error(int status, const char* what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status) {
  if (status != OSY_OK) throw Error(status, osy_last_error());
}

/* Vector2d<T> (inc/helper.cuh:7-76), the part a Source needs. */
template <typename T>
struct Vector2d {
  T x{}, y{};
  Vector2d() = default;
  Vector2d(T x_, T y_) : x(x_), y(y_) {}
  T get_x() const { return x; }
  T get_y() const { return y; }
  void set_x(T v) { x = v; }
  void set_y(T v) { y = v; }
};

/* Source (inc/fluid.cuh:8-13): the mouse source of one frame, in cells. */
struct Source {
  bool active = false;
  float smoke = 0.0f;
  float velocity = 0.0f;
  Vector2d<int> position;
};

/* ConfigParser (src/config_parser.cpp:11-13): reads OpenSayal.conf.json with
 * the reference's keys, defaults and dotted lookup. */
class ConfigParser {
 public:
  ConfigParser() : file_("OpenSayal.conf.json") {}
  explicit ConfigParser(std::string config_file_name) : file_(std::move(config_file_name)) {}
  Config parse() const {
    Config c;
    check(osy_config_parse_file(file_.c_str(), &c));
    return c;
  }

 private:
  std::string file_;
};

/* Fluid (inc/fluid.cuh:15-116): the solver on the current HIP device. */
class Fluid {
 public:
  explicit Fluid(const Config& config) {  // src/fluid.cu:41-71
    check(osy_fluid_create(&config, &h_));
    osy_fluid_info info;
    check(osy_fluid_get_info(h_, &info));
    width = info.width;
    height = info.height;
    enable_pressure = config.enable_pressure != 0;
  }
  ~Fluid() {  // src/fluid.cu:73-84
    if (h_) osy_fluid_destroy(h_);
  }
  Fluid(const Fluid&) = delete;
  Fluid& operator=(const Fluid&) = delete;
  Fluid(Fluid&& o) noexcept { *this = std::move(o); }
  Fluid& operator=(Fluid&& o) noexcept {
    if (this != &o) {
      if (h_) osy_fluid_destroy(h_);
      h_ = std::exchange(o.h_, nullptr);
      width = o.width;
      height = o.height;
      enable_pressure = o.enable_pressure;
      min_pressure = o.min_pressure;
      max_pressure = o.max_pressure;
    }
    return *this;
  }

  /* Fluid::update (src/fluid.cu:770-795): one synchronous step, after which
   * min_pressure / max_pressure hold the step's pressure range. */
  void update(const Source& source, float d_t) {
    const osy_source s{source.active ? 1 : 0, source.smoke, source.velocity,
                       source.position.get_x(), source.position.get_y()};
    check(osy_fluid_update(h_, &s, d_t));
    refresh_pressure_range();
  }

  /* `steps` updates with an inactive source, from a captured HIP graph. */
  void step(int32_t steps, float d_t) {
    check(osy_fluid_step(h_, steps, d_t));
    refresh_pressure_range();
  }

  /* d_vel_x, d_vel_y, d_smoke, d_pressure (float) or d_is_solid, total_s
   * (int32, as float-sized elements), W*H values in the reference's indx
   * order. */
  template <typename T = float>
  std::vector<T> field(osy_field f) const {
    static_assert(sizeof(T) == 4, "fields have 4-byte elements");
    std::vector<T> out((size_t)width * height);
    check(osy_fluid_get_field(h_, f, out.data(), out.size() * 4));
    return out;
  }
  void set_field(osy_field f, const std::vector<float>& values) {
    if (values.size() != (size_t)width * height)
      throw Error(OSY_E_INVALID, "set_field: expected width*height values");
    check(osy_fluid_set_field(h_, f, values.data(), values.size() * 4));
  }

  /* GraphicsHandler::update_fluid_pixels (src/graphics_handler.cu:258-300):
   * the RGBA8888 frame, W*H pixels, row 0 the top of the screen. */
  std::vector<uint32_t> render_rgba() const {
    std::vector<uint32_t> px((size_t)width * height);
    check(osy_fluid_render_rgba(h_, px.data(), px.size() * 4));
    return px;
  }

  /* GraphicsHandler::update_traces (src/graphics_handler.cu:398-412):
   * traces_x / traces_y in indx_traces order. */
  void traces(int32_t distance, int32_t length, float d_t, std::vector<int32_t>& traces_x,
              std::vector<int32_t>& traces_y) const {
    if (distance < 1) throw Error(OSY_E_INVALID, "traces: distance must be at least 1");
    const size_t n = (size_t)(width / distance) * (height / distance) * (size_t)(length > 0 ? length : 0);
    traces_x.assign(n, 0);
    traces_y.assign(n, 0);
    check(osy_fluid_traces(h_, distance, length, d_t, traces_x.data(), traces_y.data(), n * 4));
  }

  osy_fluid* handle() const { return h_; }

  int width = 0, height = 0;
  float min_pressure = 0.0f, max_pressure = 0.0f;

 private:
  void refresh_pressure_range() {
    if (enable_pressure) check(osy_fluid_pressure_range(h_, &min_pressure, &max_pressure));
  }

  osy_fluid* h_ = nullptr;
  bool enable_pressure = false;
};

}  // namespace opensayal
