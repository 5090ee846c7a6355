// fluid_host.cpp - the solver object behind the C ABI: replaces the
// reference's class Fluid (inc/fluid.cuh:15-116, src/fluid.cu:41-161, 770-795)
// with device state laid out for gfx950, one HIP stream per handle and the
// step sequence replayed from captured HIP graphs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "fluid_internal.hpp"
#include "projection.hpp"

namespace osy {

namespace {
thread_local std::string g_last_error;
}

int set_error(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}
void clear_error() { g_last_error.clear(); }

int config_validate(const osy_config& c) {
  if (c.width < 3 || c.height < 3 || c.width > 32768 || c.height > 32768)
    return set_error(OSY_E_INVALID, "sim.width and sim.height must be in [3, 32768]");
  const int cs = (int)c.cell_size;  // Fluid::cell_size is int (inc/fluid.cuh:46)
  if (!(c.cell_size >= 1.0f) || cs > 64)
    return set_error(OSY_E_INVALID, "sim.cell_size must be in [1, 64] (it is truncated to int)");
  if (c.projection_n < 0 || c.projection_n > 100000)
    return set_error(OSY_E_INVALID, "sim.projection.n must be in [0, 100000]");
  if (!(c.projection_o >= 0.0f && c.projection_o <= 2.0f))
    return set_error(OSY_E_INVALID, "sim.projection.o must be within [0, 2]");
  if (!std::isfinite(c.viscosity))
    return set_error(OSY_E_INVALID, "fluid.viscosity must be finite");
  if (c.smoke_count != 1) {
    const int spacing = (c.smoke_count - 1) != 0
                            ? (c.pipe_height - c.smoke_count * c.smoke_height) / (c.smoke_count - 1)
                            : 0;
    if (spacing + c.smoke_height == 0)
      return set_error(OSY_E_INVALID,
                       "sim.wind_tunnel: smoke stripe period is 0 (the reference divides by zero)");
  }
  if (std::abs((long long)c.obstacle_center_x) > (1LL << 24) ||
      std::abs((long long)c.obstacle_center_y) > (1LL << 24) ||
      std::fabs(c.obstacle_radius) >= 16777216.0f)
    return set_error(OSY_E_INVALID, "sim.obstacle centre and radius must be below 2^24");
  if (!std::isfinite(c.d_t) || !std::isfinite(c.g) || !std::isfinite(c.wind_speed) ||
      !std::isfinite(c.density))
    return set_error(OSY_E_INVALID, "d_t, g, wind_tunnel.speed and density must be finite");
  if (c.storage < OSY_STORAGE_F32 || c.storage > OSY_STORAGE_BF16)
    return set_error(OSY_E_INVALID, "storage must be f32, f16 or bf16");
  return OSY_OK;
}

}  // namespace osy

using namespace osy;

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Sizes of the parts of a handle's device allocation: six fields of the
// storage type (u, v, smoke, double-buffered), pressure, a reference-order
// staging buffer, the cell flags and gate bytes, and reduction scratch.
struct BlockLayout {
  size_t fld, f32, ref, flg, red, dred, rsd;
  size_t total() const { return 6 * fld + f32 + ref + 2 * flg + red + dred + rsd; }
};

BlockLayout block_layout(int W, int H, int storage) {
  const size_t P = align_up((size_t)W, 64);
  const size_t cells = P * (size_t)H;
  const size_t elem = storage == OSY_STORAGE_F32 ? 4 : 2;
  BlockLayout L;
  L.fld = align_up(cells * elem, 256);
  L.f32 = align_up(cells * 4, 256);
  L.ref = align_up((size_t)W * H * 4, 256);
  L.flg = align_up(cells, 256);
  L.red = align_up(sizeof(float) * (2 * kReduceBlocks + 2), 256);
  L.dred = align_up(sizeof(double) * (2 * kReduceBlocks + 2), 256);
  // resident projection: flag words, finished count, error word
  L.rsd = align_up(sizeof(unsigned) * (kResidentMaxTiles + 2), 256);
  return L;
}

// Starting launch shape before any autotuning: the register-resident
// kernel with 8 rows per wavefront and 5 sweeps per launch (a 128 x 64
// region owning 108 x 44 cells), the fastest measured at 1920 x 1080.
ProjTuning default_tuning(int W, bool pressure) {
  (void)W;
  ProjTuning t;
  t.kernel = PROJ_REGS;
  t.k = 5;
  t.tw = 128 - 4 * t.k;
  t.th = 8 * kRegWaves - 4 * t.k;
  if (!t.valid(pressure)) t.kernel = PROJ_COLOUR;
  return t;
}

int create_impl(const osy_config* cfg, osy_fluid** out) {
  if (!cfg || !out) return set_error(OSY_E_INVALID, "null argument");
  *out = nullptr;
  int st = config_validate(*cfg);
  if (st != OSY_OK) return st;
  int ndev = 0;
  OSY_HIP_CHECK(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return set_error(OSY_E_HIP, "no HIP device");

  osy_fluid* fl = new (std::nothrow) osy_fluid();
  if (!fl) return set_error(OSY_E_NOMEM, "host allocation failed");
  fl->cfg = *cfg;
  fl->W = cfg->width;
  fl->H = cfg->height;
  fl->P = (int)align_up((size_t)cfg->width, 64);
  fl->cs = (int)cfg->cell_size;
  fl->storage = cfg->storage;
  fl->elem = cfg->storage == OSY_STORAGE_F32 ? 4 : 2;
  fl->tuning = default_tuning(fl->W, cfg->enable_pressure != 0);
  hipError_t e = hipGetDevice(&fl->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&fl->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete fl;
    return set_error(OSY_E_HIP, std::string("stream: ") + hipGetErrorString(e));
  }

  const BlockLayout L = block_layout(fl->W, fl->H, fl->storage);
  const size_t fld = L.fld, f32 = L.f32, ref = L.ref, flg = L.flg, red = L.red, dred = L.dred;
  int khz = 0;
  e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, fl->device);
  if (e != hipSuccess) {
    delete fl;
    return set_error(OSY_E_HIP, std::string("device attribute: ") + hipGetErrorString(e));
  }
  fl->wall_khz = khz > 0 ? khz : 100000;
  fl->spin_ticks = (long long)fl->wall_khz * 2000;  // 2 s
  fl->block_bytes = L.total();
  e = hipMalloc(&fl->block, fl->block_bytes);
  if (e != hipSuccess) {
    delete fl;
    return set_error(OSY_E_NOMEM, std::string("device allocation of ") +
                                      std::to_string(fl->block_bytes) +
                                      " bytes failed: " + hipGetErrorString(e));
  }
  char* b = (char*)fl->block;
  for (int k = 0; k < 2; k++) {
    fl->f.u[k] = b; b += fld;
    fl->f.v[k] = b; b += fld;
    fl->f.smoke[k] = b; b += fld;
  }
  fl->f.p = (float*)b; b += f32;
  fl->staging = (float*)b; b += ref;
  fl->f.flags = (uint8_t*)b; b += flg;
  fl->f.gates = (uint8_t*)b; b += flg;
  fl->f.scratch = (float*)b; b += red;
  fl->f.dscratch = (double*)b; b += dred;
  fl->f.prange = fl->f.scratch + 2 * kReduceBlocks;
  fl->f.rflags = (unsigned*)b;
  fl->f.error = (int*)(fl->f.rflags + kResidentMaxTiles + 1);
  b += L.rsd;

  e = init_projection_kernels();
  if (e == hipSuccess) e = hipMemsetAsync(fl->block, 0, fl->block_bytes, fl->stream);
  if (e == hipSuccess) {
    MaskParams m{};
    m.W = fl->W; m.H = fl->H; m.P = fl->P;
    m.enable_drain = cfg->enable_drain;
    m.obstacle_enable = cfg->obstacle_enable;
    m.cx = (double)cfg->obstacle_center_x;
    m.cy = (double)cfg->obstacle_center_y;
    m.radius = (double)cfg->obstacle_radius;
    m.pipe_length = cfg->pipe_length;
    m.pipe_height = cfg->pipe_height;
    launch_build_flags(m, fl->f.flags, fl->f.gates, fl->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(fl->stream);
  if (e != hipSuccess) {
    delete fl;
    return set_error(OSY_E_HIP, std::string("init: ") + hipGetErrorString(e));
  }
  *out = fl;
  return OSY_OK;
}

int enqueue_eager(osy_fluid* fl, const osy_source* src, float d_t, hipStream_t s) {
  if (fl->tuning.kernel == PROJ_RESIDENT) fl->resident_pending = true;
  StepParams sp = fl->params(d_t, src);
  enqueue_step(fl->storage, sp, fl->f, &fl->uv, &fl->sm, fl->tuning, s);
  OSY_HIP_CHECK(hipGetLastError());
  return OSY_OK;
}

int enqueue_graph_steps(osy_fluid* fl, int32_t steps, float d_t, hipStream_t s) {
  if (steps > 0 && fl->tuning.kernel == PROJ_RESIDENT) fl->resident_pending = true;
  for (int32_t k = 0; k < steps; k++) {
    osy_fluid::Graph& g = fl->graphs[fl->uv][fl->sm];
    if (!g.exec || g.d_t != d_t) {
      if (g.exec) {
        OSY_HIP_CHECK(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
      }
      // Capture one step starting from the current buffer parity.  Capture
      // runs on the handle's stream; the recorded graph launches anywhere.
      int uv = fl->uv, sm = fl->sm;
      hipGraph_t graph = nullptr;
      OSY_HIP_CHECK(hipStreamBeginCapture(fl->stream, hipStreamCaptureModeThreadLocal));
      StepParams sp = fl->params(d_t, nullptr);
      enqueue_step(fl->storage, sp, fl->f, &uv, &sm, fl->tuning, fl->stream);
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(fl->stream, &graph);
      if (le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_error(OSY_E_HIP, std::string("graph capture: ") +
                                        hipGetErrorString(le != hipSuccess ? le : ce));
      }
      hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ie != hipSuccess) {
        g.exec = nullptr;
        return set_error(OSY_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
      }
      g.d_t = d_t;
      g.uv_out = uv;
      g.sm_out = sm;
    }
    OSY_HIP_CHECK(hipGraphLaunch(g.exec, s));
    fl->uv = g.uv_out;
    fl->sm = g.sm_out;
  }
  return OSY_OK;
}

int field_elem_ok(int field) { return field >= OSY_FIELD_VEL_X && field <= OSY_FIELD_TOTAL_S; }

}  // namespace

int osy::fluid_enqueue_graph_steps(osy_fluid* fl, int32_t steps, hipStream_t s) {
  return enqueue_graph_steps(fl, steps, fl->cfg.d_t, s);
}

size_t osy::fluid_block_bytes(int W, int H, int storage) {
  return block_layout(W, H, storage).total();
}

namespace osy {

bool tuning_fits(const osy_fluid* fl, const ProjTuning& t) {
  if (!t.valid(fl->cfg.enable_pressure != 0)) return false;
  if (t.kernel != PROJ_RESIDENT) return true;
  const int tiles = t.tiles(fl->W, fl->H);
  int slots = 0;
  if (resident_capacity(t.reg_rows(), &slots) != hipSuccess) return false;
  return tiles <= kResidentMaxTiles && tiles <= slots;
}

int fluid_check_error(osy_fluid* fl) {
  if (!fl->resident_pending) return OSY_OK;
  fl->resident_pending = false;
  int err = 0;
  OSY_HIP_CHECK(hipMemcpy(&err, fl->f.error, sizeof(int), hipMemcpyDeviceToHost));
  if (err) {
    const int zero = 0;
    OSY_HIP_CHECK(hipMemcpy(fl->f.error, &zero, sizeof(int), hipMemcpyHostToDevice));
    return set_error(OSY_E_STATE,
                     "projection: a neighbour wait of the resident kernel timed out; the "
                     "fields are invalid");
  }
  return OSY_OK;
}

}  // namespace osy

extern "C" {

int osy_abi_version(void) { return OSY_ABI_VERSION; }

const char* osy_last_error(void) { return g_last_error.c_str(); }

int osy_config_defaults(osy_config* cfg) {
  if (!cfg) return set_error(OSY_E_INVALID, "null config");
  config_defaults(cfg, 1920, 1080);
  return OSY_OK;
}

int osy_config_parse_file(const char* path, osy_config* cfg) {
  if (!path || !cfg) return set_error(OSY_E_INVALID, "null argument");
  return config_parse_file(path, cfg);
}

int osy_config_parse_string(const char* text, osy_config* cfg) {
  if (!text || !cfg) return set_error(OSY_E_INVALID, "null argument");
  return config_parse_text(std::string(text), cfg);
}

int osy_fluid_create(const osy_config* cfg, osy_fluid** out) {
  try {
    return create_impl(cfg, out);
  } catch (const std::exception& ex) {
    return set_error(OSY_E_NOMEM, ex.what());
  }
}

int osy_fluid_create_from_json(const char* path, osy_fluid** out) {
  osy_config c;
  int st = osy_config_parse_file(path, &c);
  if (st != OSY_OK) return st;
  return osy_fluid_create(&c, out);
}

int osy_fluid_device_bytes(const osy_config* cfg, uint64_t* bytes) {
  if (!cfg || !bytes) return set_error(OSY_E_INVALID, "null argument");
  int st = config_validate(*cfg);
  if (st != OSY_OK) return st;
  *bytes = fluid_block_bytes(cfg->width, cfg->height, cfg->storage);
  return OSY_OK;
}

int osy_fluid_destroy(osy_fluid* fluid) {
  if (!fluid) return OSY_OK;
  if (fluid->batch) batch_forget(fluid);
  (void)hipStreamSynchronize(fluid->stream);
  delete fluid;
  return OSY_OK;
}


int osy_fluid_update(osy_fluid* fl, const osy_source* src, float d_t) {
  if (!fl) return set_error(OSY_E_INVALID, "null handle");
  if (fl->batch) return set_error(OSY_E_STATE, "the handle is stepped by a batch");
  if (!std::isfinite(d_t)) return set_error(OSY_E_INVALID, "d_t must be finite");
  int st = (src && src->active) ? enqueue_eager(fl, src, d_t, fl->stream)
                                : enqueue_graph_steps(fl, 1, d_t, fl->stream);
  if (st != OSY_OK) return st;
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));  // src/fluid.cu:794
  return fluid_check_error(fl);
}

int osy_fluid_step(osy_fluid* fl, int32_t steps, float d_t) {
  int st = osy_fluid_step_async(fl, steps, d_t, nullptr);
  if (st != OSY_OK) return st;
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  return fluid_check_error(fl);
}

int osy_fluid_step_async(osy_fluid* fl, int32_t steps, float d_t, void* hip_stream) {
  if (!fl) return set_error(OSY_E_INVALID, "null handle");
  if (fl->batch) return set_error(OSY_E_STATE, "the handle is stepped by a batch");
  if (steps < 0) return set_error(OSY_E_INVALID, "steps must be >= 0");
  if (!std::isfinite(d_t)) return set_error(OSY_E_INVALID, "d_t must be finite");
  hipStream_t s = hip_stream ? (hipStream_t)hip_stream : fl->stream;
  return enqueue_graph_steps(fl, steps, d_t, s);
}

int osy_fluid_synchronize(osy_fluid* fl) {
  if (!fl) return set_error(OSY_E_INVALID, "null handle");
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  return fluid_check_error(fl);
}

int osy_fluid_project(osy_fluid* fl, int32_t sweeps, void* hip_stream) {
  if (!fl) return set_error(OSY_E_INVALID, "null handle");
  if (fl->batch) return set_error(OSY_E_STATE, "the handle is stepped by a batch");
  if (sweeps < 0) return set_error(OSY_E_INVALID, "sweeps must be >= 0");
  hipStream_t s = hip_stream ? (hipStream_t)hip_stream : fl->stream;
  StepParams sp = fl->params(fl->cfg.d_t, nullptr);
  sp.n = sweeps;
  if (fl->tuning.kernel == PROJ_RESIDENT) fl->resident_pending = true;
  enqueue_projection_only(fl->storage, sp, fl->f, &fl->uv, fl->tuning, s);
  OSY_HIP_CHECK(hipGetLastError());
  return OSY_OK;
}

int osy_fluid_get_field(osy_fluid* fl, int32_t field, void* dst, size_t bytes) {
  if (!fl || !dst) return set_error(OSY_E_INVALID, "null argument");
  if (!field_elem_ok(field)) return set_error(OSY_E_INVALID, "unknown field");
  if (bytes != (size_t)fl->W * fl->H * 4)
    return set_error(OSY_E_INVALID, "buffer must hold width*height 4-byte elements");
  if (int st = batch_wait(fl)) return st;
  StepParams sp = fl->params(fl->cfg.d_t, nullptr);
  enqueue_export(field, fl->storage, sp, fl->f, fl->uv, fl->sm, fl->staging, fl->stream);
  OSY_HIP_CHECK(hipGetLastError());
  OSY_HIP_CHECK(hipMemcpyAsync(dst, fl->staging, bytes, hipMemcpyDeviceToHost, fl->stream));
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  return fluid_check_error(fl);
}

int osy_fluid_set_field(osy_fluid* fl, int32_t field, const void* src, size_t bytes) {
  if (!fl || !src) return set_error(OSY_E_INVALID, "null argument");
  if (field == OSY_FIELD_IS_SOLID || field == OSY_FIELD_TOTAL_S)
    return set_error(OSY_E_INVALID, "is_solid and total_s are derived from the config; read-only");
  if (!field_elem_ok(field)) return set_error(OSY_E_INVALID, "unknown field");
  if (bytes != (size_t)fl->W * fl->H * 4)
    return set_error(OSY_E_INVALID, "buffer must hold width*height 4-byte elements");
  if (int st = batch_wait(fl)) return st;
  OSY_HIP_CHECK(hipMemcpyAsync(fl->staging, src, bytes, hipMemcpyHostToDevice, fl->stream));
  StepParams sp = fl->params(fl->cfg.d_t, nullptr);
  enqueue_import(field, fl->storage, sp, fl->f, fl->uv, fl->sm, fl->staging, fl->stream);
  OSY_HIP_CHECK(hipGetLastError());
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  return OSY_OK;
}

int osy_fluid_pressure_range(osy_fluid* fl, float* min_p, float* max_p) {
  if (!fl || !min_p || !max_p) return set_error(OSY_E_INVALID, "null argument");
  float r[2] = {0.0f, 0.0f};
  if (fl->cfg.enable_pressure) {
    if (fl->batch) {  // a batch step does not reduce; do it here
      if (int st = batch_wait(fl)) return st;
      enqueue_pressure_range(fl->params(fl->cfg.d_t, nullptr), fl->f, fl->stream);
      OSY_HIP_CHECK(hipGetLastError());
    }
    OSY_HIP_CHECK(hipMemcpyAsync(r, fl->f.prange, sizeof(r), hipMemcpyDeviceToHost, fl->stream));
    OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  }
  *min_p = r[0];
  *max_p = r[1];
  return OSY_OK;
}

int osy_fluid_render_rgba(osy_fluid* fl, uint32_t* dst, size_t bytes) {
  if (!fl || !dst) return set_error(OSY_E_INVALID, "null argument");
  if (bytes != (size_t)fl->W * fl->H * 4)
    return set_error(OSY_E_INVALID, "buffer must hold width*height 4-byte pixels");
  if (int st = batch_wait(fl)) return st;
  const StepParams sp = fl->params(fl->cfg.d_t, nullptr);
  if (fl->cfg.enable_pressure && fl->batch) {  // a batch step does not reduce
    enqueue_pressure_range(sp, fl->f, fl->stream);
    OSY_HIP_CHECK(hipGetLastError());
  }
  enqueue_render(fl->storage, sp, fl->f, fl->sm, fl->cfg.enable_pressure != 0,
                 fl->cfg.enable_smoke != 0, (uint32_t*)fl->staging, fl->stream);
  OSY_HIP_CHECK(hipGetLastError());
  OSY_HIP_CHECK(hipMemcpyAsync(dst, fl->staging, bytes, hipMemcpyDeviceToHost, fl->stream));
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  return fluid_check_error(fl);
}

int osy_fluid_traces(osy_fluid* fl, int32_t distance, int32_t length, float d_t,
                     int32_t* traces_x, int32_t* traces_y, size_t bytes) {
  if (!fl || !traces_x || !traces_y) return set_error(OSY_E_INVALID, "null argument");
  if (distance < 1) return set_error(OSY_E_INVALID, "distance must be at least 1");
  if (length < 1 || length > 65536)
    return set_error(OSY_E_INVALID, "length must be in [1, 65536]");
  if (!std::isfinite(d_t)) return set_error(OSY_E_INVALID, "d_t must be finite");
  const size_t seeds = (size_t)(fl->W / distance) * (size_t)(fl->H / distance);
  const size_t each = seeds * (size_t)length * 4;
  if (bytes != each)
    return set_error(OSY_E_INVALID,
                     "each buffer must hold (W/distance)*(H/distance)*length int32 entries");
  if (each == 0) return OSY_OK;
  if (int st = batch_wait(fl)) return st;
  if (fl->trace_cap < 2 * each) {
    OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
    if (fl->trace_buf) OSY_HIP_CHECK(hipFree(fl->trace_buf));
    fl->trace_buf = nullptr;
    fl->trace_cap = 0;
    OSY_HIP_CHECK(hipMalloc(&fl->trace_buf, 2 * each));
    fl->trace_cap = 2 * each;
  }
  int32_t* dx = (int32_t*)fl->trace_buf;
  int32_t* dy = dx + each / 4;
  const StepParams sp = fl->params(d_t, nullptr);
  enqueue_traces(fl->storage, sp, fl->f, fl->uv, distance, length, d_t, dx, dy, fl->stream);
  OSY_HIP_CHECK(hipGetLastError());
  OSY_HIP_CHECK(hipMemcpyAsync(traces_x, dx, each, hipMemcpyDeviceToHost, fl->stream));
  OSY_HIP_CHECK(hipMemcpyAsync(traces_y, dy, each, hipMemcpyDeviceToHost, fl->stream));
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  return OSY_OK;
}

int osy_fluid_divergence_norm(osy_fluid* fl, double* l2, double* linf) {
  if (!fl || !l2 || !linf) return set_error(OSY_E_INVALID, "null argument");
  if (int st = batch_wait(fl)) return st;
  StepParams sp = fl->params(fl->cfg.d_t, nullptr);
  enqueue_divergence(fl->storage, sp, fl->f, fl->uv, fl->stream);
  OSY_HIP_CHECK(hipGetLastError());
  double r[2];
  OSY_HIP_CHECK(hipMemcpyAsync(r, fl->f.dscratch + 2 * kReduceBlocks, sizeof(r),
                               hipMemcpyDeviceToHost, fl->stream));
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  *l2 = r[0];
  *linf = r[1];
  return OSY_OK;
}

int osy_fluid_set_tuning(osy_fluid* fl, const osy_tuning* t) {
  if (!fl || !t) return set_error(OSY_E_INVALID, "null argument");
  if (fl->batch) return set_error(OSY_E_STATE, "the handle is stepped by a batch");
  ProjTuning pt;
  pt.kernel = t->projection_kernel;
  pt.tw = t->tile_width;
  pt.th = t->tile_height;
  pt.k = t->sweeps_per_launch;
  if (pt.kernel == PROJ_COLOUR) {
    pt = ProjTuning();
    pt.kernel = PROJ_COLOUR;
  } else if (pt.kernel == PROJ_RESIDENT && pt.valid(fl->cfg.enable_pressure != 0) &&
             !tuning_fits(fl, pt)) {
    int slots = 0;
    const hipError_t e = resident_capacity(pt.reg_rows(), &slots);
    return set_error(OSY_E_INVALID,
                     "resident projection: " + std::to_string(pt.tiles(fl->W, fl->H)) +
                         " tiles do not all fit on the device at once (" +
                         std::to_string(slots) + " workgroup slots at " +
                         std::to_string(pt.reg_rows()) + " rows per wavefront; " +
                         hipGetErrorString(e) + ")");
  } else if (!pt.valid(fl->cfg.enable_pressure != 0)) {
    return set_error(OSY_E_INVALID,
                     "tiled projection needs 1 <= k <= 64, an even tile width >= 2, a tile "
                     "height >= 1, a region of at most " +
                         std::to_string(2 * 64 * kTileWaves) +
                         " cells per row and at most 160 KiB of LDS; register and resident "
                         "tiles a width of 128 - 4k and a region height of 8 x 4, 8, 10, "
                         "12 or 16 rows (resident: both tile sides >= 2k)");
  }
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  for (auto& row : fl->graphs)
    for (auto& g : row)
      if (g.exec) {
        OSY_HIP_CHECK(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
      }
  fl->tuning = pt;
  return OSY_OK;
}

int osy_fluid_set_wait_limit(osy_fluid* fl, int64_t microseconds) {
  if (!fl) return set_error(OSY_E_INVALID, "null handle");
  if (microseconds < 0) return set_error(OSY_E_INVALID, "the wait limit must be >= 0");
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  for (auto& row : fl->graphs)
    for (auto& g : row)
      if (g.exec) {
        OSY_HIP_CHECK(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
      }
  fl->spin_ticks = (long long)(microseconds * (int64_t)fl->wall_khz / 1000);
  return OSY_OK;
}

int osy_fluid_get_tuning(const osy_fluid* fl, osy_tuning* t) {
  if (!fl || !t) return set_error(OSY_E_INVALID, "null argument");
  t->projection_kernel = fl->tuning.kernel;
  t->tile_width = fl->tuning.tw;
  t->tile_height = fl->tuning.th;
  t->sweeps_per_launch = fl->tuning.k;
  return OSY_OK;
}

}  // extern "C"

namespace {

using TuneKey = std::tuple<int, int, int, int, int>;  // W, H, n, pressure, storage
std::mutex g_tune_mu;
std::map<TuneKey, ProjTuning> g_tune_cache;

TuneKey tune_key(const osy_fluid* fl) {
  return TuneKey(fl->W, fl->H, fl->cfg.projection_n, fl->cfg.enable_pressure != 0, fl->storage);
}

// Every launch shape the autotuner times: the colour-phase kernel, the
// register-resident kernels (several launches, and one resident launch when
// the tiles fit) at 4, 8, 10, 12 and 16 rows per wavefront and 1..12 sweeps
// per launch or stage, and a few LDS-tiled shapes.
std::vector<ProjTuning> tuning_candidates(bool pressure) {
  std::vector<ProjTuning> out;
  ProjTuning c;
  c.kernel = PROJ_COLOUR;
  out.push_back(c);
  for (int rows : {4, 8, 10, 12, 16})
    for (int k : {1, 2, 3, 4, 5, 6, 7, 8, 10, 12}) {
      for (int kern : {PROJ_REGS, PROJ_RESIDENT}) {
        ProjTuning r;
        r.kernel = kern;
        r.k = k;
        r.tw = 128 - 4 * k;
        r.th = kRegWaves * rows - 4 * k;
        if (r.valid(pressure)) out.push_back(r);
      }
    }
  const int lds_shapes[][3] = {{112, 16, 3}, {112, 16, 4}, {176, 24, 4}, {240, 24, 4}};
  for (const auto& sh : lds_shapes) {
    ProjTuning l;
    l.kernel = PROJ_TILED;
    l.tw = sh[0];
    l.th = sh[1];
    l.k = sh[2];
    if (l.valid(pressure)) out.push_back(l);
  }
  return out;
}

int drop_graphs(osy_fluid* fl) {
  for (auto& row : fl->graphs)
    for (auto& g : row)
      if (g.exec) {
        OSY_HIP_CHECK(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
      }
  return OSY_OK;
}

osy_tuning to_abi(const ProjTuning& t) {
  osy_tuning o;
  o.projection_kernel = t.kernel;
  o.tile_width = t.tw;
  o.tile_height = t.th;
  o.sweeps_per_launch = t.k;
  return o;
}

}  // namespace

extern "C" {

int osy_fluid_autotune(osy_fluid* fl, osy_tuning* chosen, double* ms) {
  if (!fl) return set_error(OSY_E_INVALID, "null handle");
  if (fl->batch) return set_error(OSY_E_STATE, "the handle is stepped by a batch");
  return fluid_autotune(fl, chosen, ms);
}

}  // extern "C"

int osy::fluid_autotune(osy_fluid* fl, osy_tuning* chosen, double* ms) {
  const bool pressure = fl->cfg.enable_pressure != 0;
  StepParams sp = fl->params(fl->cfg.d_t, nullptr);
  sp.forces_in_projection = 0;
  double best_ms = 0.0;
  ProjTuning best = fl->tuning;
  if (sp.n > 0) {
    OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
    // a timed-out step before the trials is reported, not cleared by them
    if (int st = fluid_check_error(fl)) return st;
    // snapshot of the current u, v and pressure; every trial starts from it
    const size_t fld = (size_t)fl->P * fl->H * fl->elem;
    const size_t pb = (size_t)fl->P * fl->H * 4;
    void* snap = nullptr;
    hipError_t e = hipMalloc(&snap, 2 * fld + pb);
    if (e != hipSuccess)
      return set_error(OSY_E_NOMEM, std::string("autotune snapshot: ") + hipGetErrorString(e));
    const int uv0 = fl->uv;
    char* sb = (char*)snap;
    auto restore = [&](hipStream_t s) -> hipError_t {
      hipError_t r = hipMemcpyAsync(fl->f.u[uv0], sb, fld, hipMemcpyDeviceToDevice, s);
      if (r == hipSuccess)
        r = hipMemcpyAsync(fl->f.v[uv0], sb + fld, fld, hipMemcpyDeviceToDevice, s);
      if (r == hipSuccess)
        r = hipMemcpyAsync(fl->f.p, sb + 2 * fld, pb, hipMemcpyDeviceToDevice, s);
      fl->uv = uv0;
      return r;
    };
    e = hipMemcpyAsync(sb, fl->f.u[uv0], fld, hipMemcpyDeviceToDevice, fl->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sb + fld, fl->f.v[uv0], fld, hipMemcpyDeviceToDevice, fl->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(sb + 2 * fld, fl->f.p, pb, hipMemcpyDeviceToDevice, fl->stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    best_ms = 1e30;
    for (const ProjTuning& cand : tuning_candidates(pressure)) {
      if (e != hipSuccess) break;
      if (!tuning_fits(fl, cand)) continue;
      float t_min = 1e30f;
      for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {  // rep 0 warms up
        e = restore(fl->stream);
        if (e == hipSuccess) e = hipEventRecord(e0, fl->stream);
        int cur = fl->uv;
        if (e == hipSuccess) {
          enqueue_projection_only(fl->storage, sp, fl->f, &cur, cand, fl->stream);
          e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, fl->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (rep > 0 && t < t_min) t_min = t;
      }
      if (e == hipSuccess && t_min < best_ms) {
        best_ms = t_min;
        best = cand;
      }
    }
    hipError_t r = restore(fl->stream);
    if (e == hipSuccess) e = r;
    r = hipStreamSynchronize(fl->stream);
    if (e == hipSuccess) e = r;
    // a resident trial whose wait timed out only mistimed itself: the state
    // was restored from the snapshot, so its error word is cleared
    const int zero = 0;
    r = hipMemcpy(fl->f.error, &zero, sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = r;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(snap);
    if (e != hipSuccess) return set_error(OSY_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
    int st = drop_graphs(fl);
    if (st != OSY_OK) return st;
    fl->tuning = best;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune_cache[tune_key(fl)] = best;
  }
  if (chosen) *chosen = to_abi(fl->tuning);
  if (ms) *ms = best_ms;
  return OSY_OK;
}

extern "C" {

int osy_fluid_autotune_cached(osy_fluid* fl) {
  if (!fl) return set_error(OSY_E_INVALID, "null handle");
  if (fl->batch) return set_error(OSY_E_STATE, "the handle is stepped by a batch");
  ProjTuning t;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune_cache.find(tune_key(fl));
    if (it == g_tune_cache.end()) return 0;
    t = it->second;
  }
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  int st = drop_graphs(fl);
  if (st != OSY_OK) return st;
  fl->tuning = t;
  return 1;
}

int osy_fluid_get_info(const osy_fluid* fl, osy_fluid_info* info) {
  if (!fl || !info) return set_error(OSY_E_INVALID, "null argument");
  info->width = fl->W;
  info->height = fl->H;
  info->cell_size = fl->cs;
  info->projection_n = fl->cfg.projection_n;
  info->pitch = fl->P;
  info->storage = fl->storage;
  info->projection_kernel = fl->tuning.kernel;
  info->sweeps_per_launch = fl->tuning.kernel == PROJ_COLOUR ? 1 : fl->tuning.k;
  info->device_bytes = fl->block_bytes;
  return OSY_OK;
}

}  // extern "C"

extern "C" {

int osy_stream_create(void** stream) {
  if (!stream) return set_error(OSY_E_INVALID, "null argument");
  hipStream_t s = nullptr;
  OSY_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return OSY_OK;
}

int osy_stream_destroy(void* stream) {
  if (stream) OSY_HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
  return OSY_OK;
}

int osy_stream_synchronize(void* stream) {
  OSY_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return OSY_OK;
}

int osy_event_create(void** event) {
  if (!event) return set_error(OSY_E_INVALID, "null argument");
  hipEvent_t e = nullptr;
  OSY_HIP_CHECK(hipEventCreate(&e));
  *event = e;
  return OSY_OK;
}

int osy_event_destroy(void* event) {
  if (event) OSY_HIP_CHECK(hipEventDestroy((hipEvent_t)event));
  return OSY_OK;
}

int osy_event_record(void* event, void* stream) {
  if (!event) return set_error(OSY_E_INVALID, "null event");
  OSY_HIP_CHECK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return OSY_OK;
}

int osy_event_synchronize(void* event) {
  if (!event) return set_error(OSY_E_INVALID, "null event");
  OSY_HIP_CHECK(hipEventSynchronize((hipEvent_t)event));
  return OSY_OK;
}

int osy_event_elapsed_ms(void* start, void* end, float* ms) {
  if (!start || !end || !ms) return set_error(OSY_E_INVALID, "null argument");
  OSY_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
  return OSY_OK;
}

int osy_calibration_copy(uint64_t bytes, int32_t reps) {
  if (bytes == 0 || bytes % 16 != 0 || reps < 0)
    return set_error(OSY_E_INVALID, "bytes must be a positive multiple of 16, reps >= 0");
  void* a = nullptr;
  void* b = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipMemset(a, 0, bytes);
  for (int32_t r = 0; r < reps && e == hipSuccess; r++) {
    enqueue_copy(b, a, bytes, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess) return set_error(OSY_E_HIP, std::string("copy: ") + hipGetErrorString(e));
  return OSY_OK;
}

}  // extern "C"
