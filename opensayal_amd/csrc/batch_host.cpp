// batch_host.cpp - batches of solver handles stepped by one persistent kernel
// launch per step and storage group (kernels_batch.hip).  The reference
// steps one Fluid per process (src/main.cu:49,97); a batch runs many of them,
// each bit-identical to stepping it alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "batch.hpp"
#include "fluid_internal.hpp"

using namespace osy;

struct osy_batch {
  uint64_t budget = 0, used = 0;
  hipStream_t stream = nullptr;
  hipStream_t last = nullptr;  // stream of the last enqueued steps
  std::vector<osy_fluid*> members;
  // the shape the batch autotuner picks for the wind-tunnel mixes (256 grids
  // of 128..1024 and of 64..384 cells wide): 12 rows, 10 sweeps, two per CU
  int rows = 12, k = 10, occupancy = 2;
  // members of at least this many cells step through their own per-grid
  // graphs (osy_batch_set_per_grid_cells); 0 = none
  int64_t per_grid_cells = int64_t(1) << 20;
  std::vector<osy_fluid*> solo;  // those members, in membership order
  bool dirty = true;

  struct Group {
    int storage = 0;
    bool pressure = false;
    BatchArgs args{};
    std::vector<int> member;  // indices into members
    std::vector<int> flip_uv, flip_sm;
    std::vector<BatchStage> stages;  // host copies of the launch's tables
    std::vector<uint4> items;
  };
  std::vector<Group> groups;
  void* dev = nullptr;  // every group's tables and counters, and the error word
  int* error = nullptr;
  hipGraphExec_t exec = nullptr;
  long long spin_ticks = 0;
  double tick_ns = 10.0;  // wall-clock tick
  bool timing = false;    // stamp every work item (osy_batch_set_timing)
  bool stamped = false;   // a step ran since the tables were built with stamps

  ~osy_batch() { release(); if (stream) (void)hipStreamDestroy(stream); }
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    if (dev) (void)hipFree(dev);
    dev = nullptr;
    error = nullptr;
    groups.clear();
    solo.clear();
    dirty = true;
    stamped = false;
  }
};

namespace {

struct StagePlan {
  int kind, q, sweeps, forces, smoke, ntiles, ntx;
};

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// The chain of stages of one simulation's step with projection shape (tw,
// th, k): the order the per-grid path runs its phases in (kernels_step.hip).
std::vector<StagePlan> plan_stages(const osy_fluid* fl, const StepParams& sp, int tw, int th, int k) {
  std::vector<StagePlan> v;
  const int n = sp.n;
  const int ns = n > 0 ? ceil_div(n, k) : 0;
  const int atx = ceil_div(sp.W, kBatchAdvTX), aty = ceil_div(sp.H, kBatchAdvTY);
  (void)fl;
  if (n == 0 || !sp.smoke_on)
    v.push_back({BK_FORCES, 0, 0, n == 0 ? 1 : 0, sp.smoke_on ? 0 : 1, atx * aty, atx});
  const int ptx = ceil_div(sp.W, tw), pty = ceil_div(sp.H, th);
  for (int q = 0; q < ns; q++) {
    const int sweeps = std::min(k, n - q * k);
    v.push_back({BK_PROJ, q, sweeps, q == 0 ? 1 : 0, 0, ptx * pty, ptx});
  }
  v.push_back({BK_ADV_UV, ns, 0, 0, 0, atx * aty, atx});
  if (sp.smoke_on) v.push_back({BK_ADV_SMOKE, ns, 0, 0, 0, atx * aty, atx});
  return v;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

int sync_batch(osy_batch* b) {
  if (b->last) OSY_HIP_CHECK(hipStreamSynchronize(b->last));
  OSY_HIP_CHECK(hipStreamSynchronize(b->stream));
  b->last = nullptr;
  if (b->error) {
    int err = 0;
    OSY_HIP_CHECK(hipMemcpy(&err, b->error, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
      const int zero = 0;
      OSY_HIP_CHECK(hipMemcpy(b->error, &zero, sizeof(int), hipMemcpyHostToDevice));
      return set_error(OSY_E_STATE,
                       "batch step: a stage dependency wait timed out; member results are "
                       "invalid");
    }
  }
  return OSY_OK;
}

int rebuild(osy_batch* b) {
  b->release();
  if (b->members.empty()) {
    b->dirty = false;
    return OSY_OK;
  }
  const int rows = b->rows, k = b->k;
  const int tw = 128 - 4 * k, th = 8 * rows - 4 * k;

  // groups by (storage, pressure), in order of first appearance; members
  // that fill the GPU alone step on their own graphs instead
  for (size_t m = 0; m < b->members.size(); m++) {
    osy_fluid* fl = b->members[m];
    if (b->per_grid_cells > 0 && (int64_t)fl->W * fl->H >= b->per_grid_cells) {
      b->solo.push_back(fl);
      continue;
    }
    const bool pressure = fl->cfg.enable_pressure != 0;
    auto it = std::find_if(b->groups.begin(), b->groups.end(), [&](const osy_batch::Group& g) {
      return g.storage == fl->storage && g.pressure == pressure;
    });
    if (it == b->groups.end()) {
      b->groups.emplace_back();
      b->groups.back().storage = fl->storage;
      b->groups.back().pressure = pressure;
      it = b->groups.end() - 1;
    }
    it->member.push_back((int)m);
  }

  struct Tables {
    std::vector<BatchSim> sims;
    std::vector<BatchStage> stages;
    std::vector<uint4> items;
    std::vector<unsigned> deps;
  };
  std::vector<Tables> tabs(b->groups.size());
  size_t bytes = 256;  // error word
  for (size_t gi = 0; gi < b->groups.size(); gi++) {
    osy_batch::Group& g = b->groups[gi];
    Tables& t = tabs[gi];
    std::vector<std::vector<StagePlan>> chains;
    size_t levels = 0;
    for (int m : g.member) {
      osy_fluid* fl = b->members[m];
      StepParams sp = fl->params(fl->cfg.d_t, nullptr);
      sp.forces_in_projection = 0;
      chains.push_back(plan_stages(fl, sp, tw, th, k));
      levels = std::max(levels, chains.back().size());
      BatchSim s{};
      s.sp = sp;
      for (int c = 0; c < 2; c++) {
        s.u[c] = fl->f.u[c];
        s.v[c] = fl->f.v[c];
        s.smoke[c] = fl->f.smoke[c];
      }
      s.p = fl->f.p;
      s.flags = fl->f.flags;
      s.gates = fl->f.gates;
      s.uv0 = fl->uv;
      s.sm0 = fl->sm;
      const int ns = sp.n > 0 ? ceil_div(sp.n, k) : 0;
      s.flip_uv = (ns + 1) & 1;
      s.flip_sm = sp.smoke_on ? 1 : 0;
      g.flip_uv.push_back(s.flip_uv);
      g.flip_sm.push_back(s.flip_sm);
      t.sims.push_back(s);
    }
    // stages and items level by level: a stage's dependency (the previous
    // stage of its simulation) is on the level before, so it precedes it
    std::vector<int> prev(g.member.size(), -1), prev_tiles(g.member.size(), 0);
    std::vector<int> prev_first(g.member.size(), 0), prev_kind(g.member.size(), -1);
    for (size_t L = 0; L < levels; L++) {
      for (size_t si = 0; si < chains.size(); si++) {
        if (L >= chains[si].size()) continue;
        const StagePlan& p = chains[si][L];
        const StepParams& sp = t.sims[si].sp;
        BatchStage st{};
        st.sim = (int)si;
        st.kind = p.kind;
        st.q = p.q;
        st.sweeps = p.sweeps;
        st.forces = p.forces;
        st.smoke = p.smoke;
        st.dep = prev[si];
        st.dep_tiles = prev_tiles[si];
        st.ntx = p.ntx;
        // a projection launch after a projection or forces stage waits only
        // for the predecessor tiles within 2k of its owned tile: those wrote
        // what its region reads, or read what its tile writes (the pressure
        // sums included, which stay on the owned tile)
        st.tile_dep = p.kind == BK_PROJ && (prev_kind[si] == BK_PROJ || prev_kind[si] == BK_FORCES);
        const int first = (int)t.items.size();
        const int id = (int)t.stages.size();
        const int ptw = prev_kind[si] == BK_PROJ ? tw : kBatchAdvTX;
        const int pth = prev_kind[si] == BK_PROJ ? th : kBatchAdvTY;
        const int pntx = ceil_div(sp.W, ptw), pnty = ceil_div(sp.H, pth);
        // one lane of a wavefront polls each dependency: lists longer than
        // 64 (very flat tiles) fall back to waiting for the whole stage
        const int span_x = ceil_div(tw + 4 * k, ptw) + 1, span_y = ceil_div(th + 4 * k, pth) + 1;
        if (span_x * span_y > 64) st.tile_dep = 0;
        t.stages.push_back(st);
        for (int tile = 0; tile < p.ntiles; tile++) {
          const unsigned d0 = (unsigned)t.deps.size();
          if (st.tile_dep) {
            const int tx = tile % p.ntx, ty = tile / p.ntx;
            const int xl = std::max(0, tx * tw - 2 * k), xh = std::min(sp.W, tx * tw + tw + 2 * k);
            const int yl = std::max(0, ty * th - 2 * k), yh = std::min(sp.H, ty * th + th + 2 * k);
            for (int by = yl / pth; by <= std::min(pnty - 1, (yh - 1) / pth); by++)
              for (int bx = xl / ptw; bx <= std::min(pntx - 1, (xh - 1) / ptw); bx++)
                t.deps.push_back((unsigned)(prev_first[si] + by * pntx + bx));
          }
          t.items.push_back(make_uint4((unsigned)id, (unsigned)tile, d0,
                                       (unsigned)t.deps.size() - d0));
        }
        prev[si] = id;
        prev_tiles[si] = p.ntiles;
        prev_first[si] = first;
        prev_kind[si] = p.kind;
      }
    }
    bytes += align256(t.sims.size() * sizeof(BatchSim)) +
             align256(t.stages.size() * sizeof(BatchStage)) +
             align256(t.items.size() * sizeof(uint4)) +
             align256(std::max<size_t>(1, t.deps.size()) * sizeof(unsigned)) + align256(8) +
             align256(t.stages.size() * 8) + align256(t.items.size() * 8) +
             (b->timing ? align256(t.items.size() * 32) : 0);
  }

  hipError_t e = hipMalloc(&b->dev, bytes);
  if (e != hipSuccess) {
    b->dev = nullptr;
    return set_error(OSY_E_NOMEM, std::string("batch tables: ") + hipGetErrorString(e));
  }
  OSY_HIP_CHECK(hipMemsetAsync(b->dev, 0, bytes, b->stream));
  char* p = (char*)b->dev;
  b->error = (int*)p;
  p += 256;
  for (size_t gi = 0; gi < b->groups.size(); gi++) {
    osy_batch::Group& g = b->groups[gi];
    Tables& t = tabs[gi];
    BatchArgs& a = g.args;
    a.sims = (const BatchSim*)p;
    OSY_HIP_CHECK(hipMemcpyAsync(p, t.sims.data(), t.sims.size() * sizeof(BatchSim),
                                 hipMemcpyHostToDevice, b->stream));
    p += align256(t.sims.size() * sizeof(BatchSim));
    a.stages = (const BatchStage*)p;
    OSY_HIP_CHECK(hipMemcpyAsync(p, t.stages.data(), t.stages.size() * sizeof(BatchStage),
                                 hipMemcpyHostToDevice, b->stream));
    p += align256(t.stages.size() * sizeof(BatchStage));
    a.items = (const uint4*)p;
    OSY_HIP_CHECK(hipMemcpyAsync(p, t.items.data(), t.items.size() * sizeof(uint4),
                                 hipMemcpyHostToDevice, b->stream));
    p += align256(t.items.size() * sizeof(uint4));
    a.deps = (const unsigned*)p;
    if (!t.deps.empty())
      OSY_HIP_CHECK(hipMemcpyAsync(p, t.deps.data(), t.deps.size() * sizeof(unsigned),
                                   hipMemcpyHostToDevice, b->stream));
    p += align256(std::max<size_t>(1, t.deps.size()) * sizeof(unsigned));
    a.head = (unsigned long long*)p;
    p += align256(8);
    a.done = (unsigned long long*)p;
    p += align256(t.stages.size() * 8);
    a.tdone = (unsigned long long*)p;
    p += align256(t.items.size() * 8);
    a.stamps = nullptr;
    if (b->timing) {
      a.stamps = (unsigned long long*)p;
      p += align256(t.items.size() * 32);
    }
    g.stages = t.stages;
    g.items = t.items;
    a.nitems = (int)t.items.size();
    a.error = b->error;
    a.spin_ticks = b->spin_ticks;
    a.tw = tw;
    a.th = th;
    a.k = k;
    OSY_HIP_CHECK(batch_grid_size(g.storage, g.pressure, rows, b->occupancy, &a.grid));
  }
  OSY_HIP_CHECK(hipStreamSynchronize(b->stream));

  if (b->groups.empty()) {  // every member steps on its own graphs
    b->dirty = false;
    return OSY_OK;
  }
  // one step = one persistent launch per group, captured once
  hipGraph_t graph = nullptr;
  OSY_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (const osy_batch::Group& g : b->groups)
    enqueue_batch_step(g.storage, g.pressure, rows, b->occupancy, g.args, b->stream);
  hipError_t le = hipGetLastError();
  hipError_t ce = hipStreamEndCapture(b->stream, &graph);
  if (le != hipSuccess || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return set_error(OSY_E_HIP, std::string("batch graph capture: ") +
                                    hipGetErrorString(le != hipSuccess ? le : ce));
  }
  hipError_t ie = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) {
    b->exec = nullptr;
    return set_error(OSY_E_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(ie));
  }
  b->dirty = false;
  return OSY_OK;
}

int enqueue_steps(osy_batch* b, int32_t steps, hipStream_t s) {
  if (b->dirty) {
    int st = sync_batch(b);
    if (st == OSY_OK) st = rebuild(b);
    if (st != OSY_OK) return st;
  }
  if ((!b->exec && b->solo.empty()) || steps == 0) return OSY_OK;
  for (int32_t i = 0; i < steps; i++) {
    for (osy_fluid* fl : b->solo)
      if (int st = fluid_enqueue_graph_steps(fl, 1, s)) return st;
    if (b->exec) OSY_HIP_CHECK(hipGraphLaunch(b->exec, s));
  }
  if (s != b->stream) b->last = s;
  if (b->timing) b->stamped = true;
  if (steps & 1) {
    for (const osy_batch::Group& g : b->groups)
      for (size_t i = 0; i < g.member.size(); i++) {
        osy_fluid* fl = b->members[g.member[i]];
        fl->uv ^= g.flip_uv[i];
        fl->sm ^= g.flip_sm[i];
      }
  }
  return OSY_OK;
}

bool tuning_ok(int rows, int k, int occupancy) {
  return batch_rows_supported(rows) && batch_occupancy_supported(rows, occupancy) && k >= 1 &&
         8 * rows - 4 * k >= 1 && 128 - 4 * k >= 2;
}

}  // namespace

namespace osy {

int batch_wait(osy_fluid* fl) {
  if (!fl->batch) return OSY_OK;
  return sync_batch(fl->batch);
}

void batch_forget(osy_fluid* fl) {
  osy_batch* b = fl->batch;
  if (!b) return;
  (void)sync_batch(b);
  b->members.erase(std::remove(b->members.begin(), b->members.end(), fl), b->members.end());
  b->used -= fl->block_bytes;
  fl->batch = nullptr;
  b->release();
}

}  // namespace osy

extern "C" {

int osy_batch_create(uint64_t budget_bytes, osy_batch** out) {
  if (!out) return set_error(OSY_E_INVALID, "null argument");
  *out = nullptr;
  osy_batch* b = new (std::nothrow) osy_batch();
  if (!b) return set_error(OSY_E_NOMEM, "host allocation failed");
  b->budget = budget_bytes;
  int dev = 0, khz = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return set_error(OSY_E_HIP, std::string("batch create: ") + hipGetErrorString(e));
  }
  // one dependency wait is bounded by 2 s of wall clock
  b->spin_ticks = (long long)(khz > 0 ? khz : 100000) * 2000;
  b->tick_ns = 1e6 / (khz > 0 ? khz : 100000);
  *out = b;
  return OSY_OK;
}

int osy_batch_destroy(osy_batch* b) {
  if (!b) return OSY_OK;
  (void)sync_batch(b);
  for (osy_fluid* fl : b->members) fl->batch = nullptr;
  delete b;
  return OSY_OK;
}

int osy_batch_add(osy_batch* b, osy_fluid* fl) {
  if (!b || !fl) return set_error(OSY_E_INVALID, "null argument");
  if (fl->batch) return set_error(OSY_E_STATE, "the handle is already in a batch");
  if (fl->cfg.viscosity != 0.0f)
    return set_error(OSY_E_UNSUPPORTED,
                     "a batch does not run the diffusion pass: members need fluid.viscosity 0");
  if (b->budget && b->used + fl->block_bytes > b->budget)
    return set_error(OSY_E_BUDGET, "the handle's " + std::to_string(fl->block_bytes) +
                                       " device bytes exceed the batch budget (" +
                                       std::to_string(b->budget - b->used) + " bytes left)");
  int st = sync_batch(b);
  if (st != OSY_OK) return st;
  OSY_HIP_CHECK(hipStreamSynchronize(fl->stream));
  b->members.push_back(fl);
  b->used += fl->block_bytes;
  fl->batch = b;
  b->release();
  return OSY_OK;
}

int osy_batch_remove(osy_batch* b, osy_fluid* fl) {
  if (!b || !fl) return set_error(OSY_E_INVALID, "null argument");
  if (fl->batch != b) return set_error(OSY_E_STATE, "the handle is not in this batch");
  int st = sync_batch(b);
  batch_forget(fl);
  return st;
}

int osy_batch_step(osy_batch* b, int32_t steps) {
  int st = osy_batch_step_async(b, steps, nullptr);
  if (st != OSY_OK) return st;
  return sync_batch(b);
}

int osy_batch_step_async(osy_batch* b, int32_t steps, void* hip_stream) {
  if (!b) return set_error(OSY_E_INVALID, "null handle");
  if (steps < 0) return set_error(OSY_E_INVALID, "steps must be >= 0");
  if (b->last && hip_stream && (hipStream_t)hip_stream != b->last) {
    int st = sync_batch(b);
    if (st != OSY_OK) return st;
  }
  return enqueue_steps(b, steps, hip_stream ? (hipStream_t)hip_stream : b->stream);
}

int osy_batch_synchronize(osy_batch* b) {
  if (!b) return set_error(OSY_E_INVALID, "null handle");
  return sync_batch(b);
}

int osy_batch_set_tuning(osy_batch* b, const osy_batch_tuning* t) {
  if (!b || !t) return set_error(OSY_E_INVALID, "null argument");
  const int occ = t->occupancy == 0 ? 1 : t->occupancy;
  if (!tuning_ok(t->rows, t->sweeps_per_tile, occ))
    return set_error(OSY_E_INVALID,
                     "batch tuning needs rows 8 or 12, 1 <= sweeps_per_tile < 2 * rows, and "
                     "occupancy 1 (rows 8) or 2");
  int st = sync_batch(b);
  if (st != OSY_OK) return st;
  b->rows = t->rows;
  b->k = t->sweeps_per_tile;
  b->occupancy = occ;
  b->release();
  return OSY_OK;
}

int osy_batch_get_tuning(const osy_batch* b, osy_batch_tuning* t) {
  if (!b || !t) return set_error(OSY_E_INVALID, "null argument");
  t->rows = b->rows;
  t->sweeps_per_tile = b->k;
  t->occupancy = b->occupancy;
  return OSY_OK;
}

int osy_batch_autotune(osy_batch* b, osy_batch_tuning* chosen, double* ms) {
  if (!b) return set_error(OSY_E_INVALID, "null handle");
  int st = sync_batch(b);
  if (st != OSY_OK) return st;
  if (b->members.empty()) return set_error(OSY_E_STATE, "the batch has no members");
  // members that step on their own graphs: the per-grid autotuner picks
  // their projection shape (their graphs run one after another on the
  // batch's stream, so a resident projection is safe there)
  for (osy_fluid* fl : b->members)
    if (b->per_grid_cells > 0 && (int64_t)fl->W * fl->H >= b->per_grid_cells) {
      st = fluid_autotune(fl, nullptr, nullptr);
      if (st != OSY_OK) return st;
    }
  // save every member's device state and buffer parity
  size_t total = 0;
  for (osy_fluid* fl : b->members) total += fl->block_bytes;
  void* save = nullptr;
  hipError_t e = hipMalloc(&save, total);
  if (e != hipSuccess)
    return set_error(OSY_E_NOMEM, std::string("autotune snapshot: ") + hipGetErrorString(e));
  std::vector<std::pair<int, int>> parity;
  size_t off = 0;
  for (osy_fluid* fl : b->members) {
    e = hipMemcpyAsync((char*)save + off, fl->block, fl->block_bytes, hipMemcpyDeviceToDevice,
                       b->stream);
    if (e != hipSuccess) break;
    off += fl->block_bytes;
    parity.emplace_back(fl->uv, fl->sm);
  }
  const int rows0 = b->rows, k0 = b->k, occ0 = b->occupancy;
  int best_rows = rows0, best_k = k0, best_occ = occ0;
  double best = 1e30;
  hipEvent_t a = nullptr, z = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&z);
  st = e == hipSuccess ? OSY_OK : set_error(OSY_E_HIP, hipGetErrorString(e));
  const int variants[3][2] = {{8, 1}, {8, 2}, {12, 2}};
  for (const auto& var : variants) {
    const int rows = var[0], occ = var[1];
    for (int k = 1; k <= 12 && st == OSY_OK; k++) {
      if (!tuning_ok(rows, k, occ)) continue;
      b->rows = rows;
      b->k = k;
      b->occupancy = occ;
      b->release();
      st = enqueue_steps(b, 1, b->stream);  // warm-up (and rebuild)
      if (st != OSY_OK) break;
      const int reps = 3;
      if ((e = hipEventRecord(a, b->stream)) != hipSuccess) break;
      st = enqueue_steps(b, reps, b->stream);
      if (st != OSY_OK) break;
      if ((e = hipEventRecord(z, b->stream)) != hipSuccess) break;
      st = sync_batch(b);
      if (st != OSY_OK) break;
      float t = 0.0f;
      if ((e = hipEventElapsedTime(&t, a, z)) != hipSuccess) break;
      if (t / reps < best) {
        best = t / reps;
        best_rows = rows;
        best_k = k;
        best_occ = occ;
      }
    }
  }
  if (st == OSY_OK && e != hipSuccess) st = set_error(OSY_E_HIP, hipGetErrorString(e));
  // restore
  (void)hipStreamSynchronize(b->stream);
  off = 0;
  for (size_t i = 0; i < b->members.size(); i++) {
    osy_fluid* fl = b->members[i];
    if (i < parity.size()) {
      (void)hipMemcpyAsync(fl->block, (char*)save + off, fl->block_bytes, hipMemcpyDeviceToDevice,
                           b->stream);
      fl->uv = parity[i].first;
      fl->sm = parity[i].second;
    }
    off += fl->block_bytes;
  }
  (void)hipStreamSynchronize(b->stream);
  (void)hipFree(save);
  if (a) (void)hipEventDestroy(a);
  if (z) (void)hipEventDestroy(z);
  b->rows = st == OSY_OK ? best_rows : rows0;
  b->k = st == OSY_OK ? best_k : k0;
  b->occupancy = st == OSY_OK ? best_occ : occ0;
  b->release();
  if (st != OSY_OK) return st;
  if (chosen) {
    chosen->rows = b->rows;
    chosen->sweeps_per_tile = b->k;
    chosen->occupancy = b->occupancy;
  }
  if (ms) *ms = best;
  return OSY_OK;
}

int osy_batch_set_per_grid_cells(osy_batch* b, int64_t cells) {
  if (!b) return set_error(OSY_E_INVALID, "null handle");
  if (cells < 0) return set_error(OSY_E_INVALID, "cells must be >= 0");
  int st = sync_batch(b);
  if (st != OSY_OK) return st;
  if (b->per_grid_cells != cells) {
    b->per_grid_cells = cells;
    b->release();
  }
  return OSY_OK;
}

int osy_batch_set_timing(osy_batch* b, int32_t enable) {
  if (!b) return set_error(OSY_E_INVALID, "null handle");
  int st = sync_batch(b);
  if (st != OSY_OK) return st;
  if (b->timing != (enable != 0)) {
    b->timing = enable != 0;
    b->release();  // the tables are rebuilt with (or without) the stamps
  }
  return OSY_OK;
}

int osy_batch_item_times(osy_batch* b, osy_batch_item_time* out, int64_t capacity,
                         int64_t* count, double* tick_ns) {
  if (!b || !count) return set_error(OSY_E_INVALID, "null argument");
  int st = sync_batch(b);
  if (st != OSY_OK) return st;
  if (!b->timing || !b->stamped)
    return set_error(OSY_E_STATE, "no timed batch step: enable timing, then step");
  int64_t n = 0;
  for (const osy_batch::Group& g : b->groups) n += g.args.nitems;
  *count = n;
  if (tick_ns) *tick_ns = b->tick_ns;
  if (capacity < n || (n > 0 && !out))
    return set_error(OSY_E_INVALID, "the record array holds " + std::to_string(capacity) +
                                        " items; the last step has " + std::to_string(n));
  std::vector<unsigned long long> raw;
  int64_t o = 0;
  for (const osy_batch::Group& g : b->groups) {
    raw.resize(4 * (size_t)g.args.nitems);
    OSY_HIP_CHECK(hipMemcpy(raw.data(), g.args.stamps, raw.size() * 8, hipMemcpyDeviceToHost));
    // stage index within its member's chain: count the member's earlier stages
    std::vector<int> chain_pos(g.stages.size(), 0), seen(g.member.size(), 0);
    for (size_t si = 0; si < g.stages.size(); si++) chain_pos[si] = seen[g.stages[si].sim]++;
    for (int i = 0; i < g.args.nitems; i++, o++) {
      const BatchStage& stg = g.stages[g.items[i].x];
      osy_batch_item_time& r = out[o];
      r.member = g.member[stg.sim];
      r.kind = stg.kind;
      r.stage = chain_pos[g.items[i].x];
      r.tile = (int32_t)g.items[i].y;
      r.claim = raw[4 * i];
      r.ready = raw[4 * i + 1];
      r.done = raw[4 * i + 2];
      r.hw_id = raw[4 * i + 3];
    }
  }
  return OSY_OK;
}

int osy_batch_get_info(const osy_batch* b, osy_batch_info* info) {
  if (!b || !info) return set_error(OSY_E_INVALID, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->count = (int32_t)b->members.size();
  info->budget_bytes = b->budget;
  info->used_bytes = b->used;
  for (const osy_fluid* fl : b->members) info->cells += (int64_t)fl->W * fl->H;
  info->groups = (int32_t)b->groups.size();
  for (const osy_batch::Group& g : b->groups) info->work_items += g.args.nitems;
  return OSY_OK;
}

}  // extern "C"
