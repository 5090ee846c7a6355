// config.cpp - OpenSayal.conf.json loader: the key surface, defaults and
// lookup rules of the reference's ConfigParser (src/config_parser.cpp:11-184,
// inc/config_parser.hpp:118-150), rebuilt on json_value.hpp.
#include <algorithm>
#include <climits>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "json_value.hpp"
#include "osy_internal.hpp"

namespace osy {
namespace {

int32_t as_int(const Json& v) {
  int64_t x = v.to_int64();
  if (x < INT32_MIN || x > INT32_MAX) throw JsonError("integer out of int32 range");
  return (int32_t)x;
}
float as_float(const Json& v) { return (float)v.to_double(); }
int32_t as_bool(const Json& v) { return v.to_bool() ? 1 : 0; }

int32_t get_int(const Json& j, const char* key, int32_t def) {
  return json_get_or<int32_t>(j, key, def, as_int);
}
float get_float(const Json& j, const char* key, float def) {
  return json_get_or<float>(j, key, def, as_float);
}
int32_t get_bool(const Json& j, const char* key, int32_t def) {
  return json_get_or<int32_t>(j, key, def, as_bool);
}

}  // namespace

void config_defaults(osy_config* c, int32_t width, int32_t height) {
  std::memset(c, 0, sizeof(*c));
  c->width = width;
  c->height = height;
  c->cell_size = 1.0f;
  c->cell_pixel_size = 1;
  c->enable_drain = 1;
  c->enable_pressure = 0;
  c->enable_smoke = 1;
  c->enable_interactive = 0;
  c->projection_n = 50;
  c->projection_o = 1.9f;
  c->pipe_height = height / 4;
  c->pipe_length = 0;
  c->smoke_length = 1;
  c->smoke_height = height / 4;
  c->smoke_count = 1;
  c->wind_speed = 0.0f;
  c->wind_smoke = 1.0f;
  c->g = 0.0f;
  c->d_t = 0.05f;
  c->enable_real_time = 0;
  c->real_time_multiplier = 1.0f;
  c->enable_decay = 0;
  c->decay_rate = 0.05f;
  c->obstacle_enable = 1;
  c->obstacle_center_x = width / 2;
  c->obstacle_center_y = height / 2;
  c->obstacle_radius = std::min(height, width) / 30.0f;
  c->density = 1.0f;
  c->drag_coeff = 0.0f;
  c->viscosity = 0.0f;
  c->block_size_x = 64;
  c->block_size_y = 1;
  c->omp_thread_count = 0;
  c->storage = OSY_STORAGE_F32;
}

// ConfigParser::parse (src/config_parser.cpp:15-184).  Render-only keys
// (visual.*) are accepted and not stored.  fluid.viscosity keeps the
// reference's default of 0.001 when absent, which runs the diffusion pass.
int config_parse_text(const std::string& text, osy_config* out) {
  Json j;
  try {
    j = Json::parse(text);
  } catch (const JsonError& e) {
    return set_error(OSY_E_CONFIG, e.what());
  }
  try {
    osy_config c;
    const int32_t H = get_int(j, "sim.height", 1080);
    const int32_t W = get_int(j, "sim.width", 1920);
    config_defaults(&c, W, H);
    c.omp_thread_count = get_int(j, "thread.openMP.thread_count", 0);
    c.block_size_x = get_int(j, "thread.cuda.block_size_x", 64);
    c.block_size_y = get_int(j, "thread.cuda.block_size_y", 1);
    c.cell_pixel_size = get_int(j, "sim.cell_pixel_size", 1);
    c.cell_size = get_float(j, "sim.cell_size", 1.0f);
    c.enable_drain = get_bool(j, "sim.enable_drain", 1);
    c.enable_pressure = get_bool(j, "sim.enable_pressure", 0);
    c.enable_smoke = get_bool(j, "sim.enable_smoke", 1);
    c.enable_interactive = get_bool(j, "sim.enable_interactive", 0);
    c.projection_n = get_int(j, "sim.projection.n", 50);
    c.projection_o = get_float(j, "sim.projection.o", 1.9f);
    c.pipe_height = get_int(j, "sim.wind_tunnel.pipe_height", H / 4);
    c.pipe_length = 0;  // hard-coded by the reference (src/config_parser.cpp:64)
    c.smoke_length = get_int(j, "sim.wind_tunnel.smoke_length", 1);
    c.smoke_height = get_int(j, "sim.wind_tunnel.smoke_height", H / 4);
    c.smoke_count = get_int(j, "sim.wind_tunnel.smoke_count", 1);
    c.wind_speed = get_float(j, "sim.wind_tunnel.speed", 0.0f);
    c.wind_smoke = get_float(j, "sim.wind_tunnel.smoke", 1.0f);
    c.g = get_float(j, "sim.physics.g", 0.0f);
    c.d_t = get_float(j, "sim.time.d_t", 0.05f);
    c.enable_real_time = get_bool(j, "sim.time.enable_real_time", 0);
    c.real_time_multiplier = get_float(j, "sim.time.real_time_multiplier", 1.0f);
    c.enable_decay = get_bool(j, "sim.smoke.enable_decay", 0);
    c.decay_rate = get_float(j, "sim.smoke.decay_rate", 0.05f);
    c.obstacle_enable = get_bool(j, "sim.obstacle.enable", 1);
    c.obstacle_center_x = get_int(j, "sim.obstacle.center_x", W / 2);
    c.obstacle_center_y = get_int(j, "sim.obstacle.center_y", H / 2);
    c.obstacle_radius =
        get_float(j, "sim.obstacle.radius", std::min(H, W) / 30.0f);
    c.density = get_float(j, "fluid.density", 1.0f);
    c.drag_coeff = get_float(j, "fluid.drag_coeff", 0.0f);
    c.viscosity = get_float(j, "fluid.viscosity", 0.001f);
    // this implementation's storage key (string), default f32
    c.storage = OSY_STORAGE_F32;
    std::string storage = json_get_or<std::string>(
        j, "sim.storage", std::string("f32"),
        [](const Json& v) { return v.str(); });
    if (storage == "f32") c.storage = OSY_STORAGE_F32;
    else if (storage == "f16") c.storage = OSY_STORAGE_F16;
    else if (storage == "bf16") c.storage = OSY_STORAGE_BF16;
    else return set_error(OSY_E_CONFIG, "sim.storage must be \"f32\", \"f16\" or \"bf16\"");
    *out = c;
    return OSY_OK;
  } catch (const JsonError& e) {
    return set_error(OSY_E_CONFIG, std::string("config: ") + e.what());
  }
}

int config_parse_file(const char* path, osy_config* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return set_error(OSY_E_CONFIG, std::string("cannot open config file ") + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  return config_parse_text(ss.str(), out);
}

}  // namespace osy
