// kernels_batch.hip - one persistent launch steps a whole batch of
// simulations (the per-simulation step is Fluid::update, src/fluid.cu:770-795).
//
// Every workgroup loops: claim the next work item from the batch's atomic
// counter, wait for its dependencies (the overlapping tiles of the previous
// projection stage, per-item counters; or the whole previous stage, a
// per-stage counter of finished tiles; read with agent-scope acquire), run
// the stage's body on the item's tile with the same device code as the
// per-grid kernels, publish completion with an agent-scope release.  Items
// are ordered so that dependencies come first; a workgroup only ever waits
// for items already claimed by running workgroups, so the launch drains
// whatever the occupancy.  A wait is bounded by `spin_ticks`; when the bound
// is hit the item is skipped, the error word set, and completion still
// published, so the launch ends in every case.
#include "advect_device.hpp"
#include "batch.hpp"
#include "projection_device.hpp"

#pragma clang fp contract(off)

namespace osy {
namespace {

static_assert(kBatchAdvTX == kAdvTX, "advection tile width");

union BatchShared {
  RegShared reg;
  AdvShared<kBatchAdvTY> adv;
};

// A value every lane holds the same copy of, as a scalar (so the claim, the
// loop exit and the epoch arithmetic are wavefront-uniform).
__device__ inline unsigned long long uniform_u64(unsigned long long x) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)x);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(x >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Wavefront 0 (every lane reads the same counter): wait until *counter >=
// target, polling with relaxed loads (an
// acquire per poll would invalidate the L2 each time), then one agent-scope
// acquire.  Returns false on timeout, or at once when an earlier wait of the
// launch timed out (so a failed launch drains in about one bound rather than
// one per dependency level).
__device__ inline bool wait_count(const unsigned long long* counter, unsigned long long target,
                                  const int* error, long long spin_ticks) {
  if (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    const long long t0 = wall_clock64();
    for (;;) {
      __builtin_amdgcn_s_sleep(2);
      if (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target)
        break;
      if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
      if (wall_clock64() - t0 > spin_ticks) return false;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// Wavefront 0: wait until every counter named by the item's dependency list
// (at most 64, one per lane) has reached `target`.  All lanes poll at once
// with relaxed loads, so a poll costs one memory round trip whatever the
// list length; then one agent-scope acquire.  One bound for the whole wait.
__device__ inline bool wait_tiles(const unsigned long long* tdone, const unsigned* deps, int n,
                                  unsigned long long target, const int* error,
                                  long long spin_ticks) {
  const int lane = (int)(threadIdx.x & 63);
  const unsigned long long* c = tdone + (lane < n ? deps[lane] : 0u);
  const long long t0 = wall_clock64();
  for (;;) {
    const bool ready =
        lane >= n || __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target;
    if (__all(ready)) break;
    if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (wall_clock64() - t0 > spin_ticks) return false;
    __builtin_amdgcn_s_sleep(2);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

template <typename T, bool PRESSURE, int R>
__device__ __attribute__((always_inline)) inline void run_item(const BatchStage& st, const BatchSim& sim, int tile,
                                unsigned epoch_bit, int tw, int th, int k, BatchShared& sh) {
  const StepParams& sp = sim.sp;
  const int uv = sim.uv0 ^ (int)(epoch_bit & (unsigned)sim.flip_uv);
  const int sm = sim.sm0 ^ (int)(epoch_bit & (unsigned)sim.flip_sm);
  const GridView g{sp.W, sp.H, sp.P, sp.cs, const_cast<uint8_t*>(sim.flags)};
  const int kind = st.kind;
  if (kind == BK_PROJ) {
    RegArgs a;
    a.tw = tw;
    a.th = th;
    a.k = k;
    a.ntx = st.ntx;
    a.sweeps = st.sweeps;
    a.forces = st.forces;
    const int in = uv ^ (st.q & 1);
    project_regs_tile<T, PRESSURE, R, true>(sp, sim.gates, (const T*)sim.u[in], (const T*)sim.v[in],
                                      (T*)sim.u[in ^ 1], (T*)sim.v[in ^ 1], sim.p, a, tile,
                                      sh.reg);
    return;
  }
  const int bx = tile % st.ntx, by = tile / st.ntx;
  if (kind == BK_FORCES) {
    forces_tile<T, kBatchAdvTY, kRegWaves>(sp, st.forces, st.smoke, (T*)sim.u[uv], (T*)sim.v[uv],
                                           (T*)sim.smoke[sm], sim.p, bx, by);
    return;
  }
  if (kind == BK_ADV_UV) {
    const int in = uv ^ (st.q & 1);
    if (sp.cs == 1)
      advect_velocity_tile<T, true, kBatchAdvTY, kRegWaves>(
          sp, g, (const T*)sim.u[in], (const T*)sim.v[in], (T*)sim.u[in ^ 1], (T*)sim.v[in ^ 1],
          bx, by, sh.adv);
    else
      advect_velocity_tile<T, false, kBatchAdvTY, kRegWaves>(
          sp, g, (const T*)sim.u[in], (const T*)sim.v[in], (T*)sim.u[in ^ 1], (T*)sim.v[in ^ 1],
          bx, by, sh.adv);
    return;
  }
  // BK_ADV_SMOKE: the velocities after this step's advection
  const int cur = uv ^ ((st.q + 1) & 1);
  if (sp.cs == 1)
    advect_smoke_tile<T, true, kBatchAdvTY, kRegWaves>(
        sp, g, (const T*)sim.u[cur], (const T*)sim.v[cur], (const T*)sim.smoke[sm],
        (T*)sim.smoke[sm ^ 1], bx, by, sh.adv);
  else
    advect_smoke_tile<T, false, kBatchAdvTY, kRegWaves>(
        sp, g, (const T*)sim.u[cur], (const T*)sim.v[cur], (const T*)sim.smoke[sm],
        (T*)sim.smoke[sm ^ 1], bx, by, sh.adv);
}

// The workgroup's loop keeps every thread-0 region between two barriers of
// the same iteration: the claim of the next item shares one region with the
// completion count of the current one.  (Two thread-0 regions adjacent across
// the loop's back-edge let the compiler merge them and move a barrier into a
// loop that thread 0 leaves, which deadlocks the other threads.)
// OCC2: registers capped for two workgroups per CU (at the price of some
// spilling for R = 8); the batch autotuner decides whether that pays.
template <typename T, bool PRESSURE, int R, bool OCC2>
__global__ void __launch_bounds__(kBatchThreads, OCC2 ? 4 : 1) k_batch_step(BatchArgs A) {
  __shared__ BatchShared sh;
  __shared__ unsigned long long s_claim;
  __shared__ int s_ok;
  const unsigned long long per = (unsigned long long)A.nitems + (unsigned long long)A.grid;
  if (threadIdx.x == 0)
    s_claim = __hip_atomic_fetch_add(A.head, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  unsigned long long raw = uniform_u64(s_claim);
  for (;;) {
    const unsigned long long epoch = raw / per;
    const int idx = (int)(raw - epoch * per);
    if (idx >= A.nitems) break;
    const uint4 it = A.items[idx];
    const int sid = __builtin_amdgcn_readfirstlane((int)it.x);
    const int tile = __builtin_amdgcn_readfirstlane((int)it.y);
    const BatchStage st = A.stages[sid];
    unsigned long long t_claim = 0, t_ready = 0;
    if (threadIdx.x < 64) {  // wavefront 0 waits; thread 0 reports
      if (A.stamps) t_claim = wall_clock64();
      bool ok = true;
      if (st.tile_dep)
        ok = wait_tiles(A.tdone, A.deps + it.z, (int)it.w, epoch + 1, A.error, A.spin_ticks);
      else if (st.dep >= 0)
        ok = wait_count(A.done + st.dep, (epoch + 1) * (unsigned long long)st.dep_tiles,
                        A.error, A.spin_ticks);
      if (threadIdx.x == 0) {
        if (!ok) __hip_atomic_store(A.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = ok ? 1 : 0;
        if (A.stamps) t_ready = wall_clock64();
      }
    }
    // thread 0's acquire invalidated the caches this workgroup reads
    // through (its CU's L1 and its XCD's L2); the barrier orders every
    // wavefront's loads after it
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(s_ok)) {
      const int simi = __builtin_amdgcn_readfirstlane(st.sim);
      run_item<T, PRESSURE, R>(st, A.sims[simi], tile, (unsigned)(epoch & 1ULL), A.tw, A.th, A.k,
                               sh);
    }
    // every wavefront's stores have reached the L2 before the barrier; thread
    // 0's release then writes the L2 back once for the whole workgroup
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
      if (A.stamps) {
        unsigned long long* r = A.stamps + 4 * (size_t)idx;
        r[0] = t_claim;
        r[1] = t_ready;
        r[2] = wall_clock64();
        r[3] = __smid();
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_fetch_add(A.done + sid, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(A.tdone + idx, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_claim = __hip_atomic_fetch_add(A.head, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    raw = uniform_u64(s_claim);
  }
}

template <typename T, bool PRESSURE>
const void* batch_kernel(int rows, int occupancy) {
  if (rows == 12) return (const void*)k_batch_step<T, PRESSURE, 12, true>;
  return occupancy == 2 ? (const void*)k_batch_step<T, PRESSURE, 8, true>
                        : (const void*)k_batch_step<T, PRESSURE, 8, false>;
}

const void* batch_kernel_for(int storage, bool pressure, int rows, int occupancy) {
  switch (storage) {
    case OSY_STORAGE_F16:
      return pressure ? batch_kernel<__half, true>(rows, occupancy)
                      : batch_kernel<__half, false>(rows, occupancy);
    case OSY_STORAGE_BF16:
      return pressure ? batch_kernel<__hip_bfloat16, true>(rows, occupancy)
                      : batch_kernel<__hip_bfloat16, false>(rows, occupancy);
    default:
      return pressure ? batch_kernel<float, true>(rows, occupancy)
                      : batch_kernel<float, false>(rows, occupancy);
  }
}

}  // namespace

hipError_t batch_grid_size(int storage, bool pressure, int rows, int occupancy, int* grid) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, batch_kernel_for(storage, pressure, rows, occupancy), kBatchThreads, 0);
  if (e != hipSuccess) return e;
  *grid = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
  return hipSuccess;
}

void enqueue_batch_step(int storage, bool pressure, int rows, int occupancy, const BatchArgs& a,
                        hipStream_t s) {
  const void* fn = batch_kernel_for(storage, pressure, rows, occupancy);
  void* args[] = {(void*)&a};
  (void)hipLaunchKernel(fn, dim3(a.grid), dim3(kBatchThreads), args, 0, s);
}

}  // namespace osy
