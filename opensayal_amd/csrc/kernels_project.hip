// kernels_project.hip - over-relaxed red-black projection (src/fluid.cu:229-295).
//
// A sweep is the even colour phase ((i+j) even) then the odd one, over
// interior fluid cells.  Same-colour cells share no face, so every colour
// phase is a set of independent cell updates and any schedule that finishes
// one phase before the next starts reproduces the reference bit for bit.
//
// Two kernels:
//  * k_project_colour: one launch per colour phase, in place (2n launches);
//    the plain restatement, kept for tiny grids.
//  * k_project_tiled: temporal blocking.  A workgroup loads its tile plus a
//    halo of u, v and cell flags into LDS, runs up to k sweeps there, and
//    writes the tile to the other half of a ping-pong pair.  A colour phase
//    changes only faces of the cells it updates, so a value can be wrong
//    after phase s only within s cells of the region edge; a halo of 2k
//    cells keeps every face of the tile exact through 2k phases.  Each phase
//    updates only the cells whose results still reach the tile (a box that
//    shrinks by one cell per phase), so the redundant halo work falls off as
//    the sweeps proceed.
#include "projection.hpp"
#include "projection_device.hpp"

#pragma clang fp contract(off)

namespace osy {
namespace {

// apply_projection_at + update_pressure_at (src/fluid.cu:221-262) for the
// cell at k = j*P + i, given its flag byte.  Face gates: the reference checks
// !is_solid of the neighbour across each face; for a fluid cell that is the
// face's fluid-fluid flag (CF_FU of (i,j) and (i+1,j), CF_FV of (i,j) and
// (i,j+1)).
template <typename T>
__device__ inline void project_cell(const StepParams& sp, const uint8_t* flags, T* u, T* v,
                                    float* p, long k, uint8_t f) {
  const int P = sp.P;
  const float uu = ldf(u, k), ur = ldf(u, k + 1);
  const float vv = ldf(v, k), vt = ldf(v, k + P);
  const float div = ((ur - uu) + vt) - vv;
  const float vd = sp.o * (div / (float)cf_total_s(f));
  if (sp.pressure_on) p[k] = p[k] + (((vd * sp.density) * (float)sp.cs) / sp.d_t);
  if (f & CF_FU) stf(u, k, uu + vd);
  if (flags[k + 1] & CF_FU) stf(u, k + 1, ur - vd);
  if (f & CF_FV) stf(v, k, vv + vd);
  if (flags[k + P] & CF_FV) stf(v, k + P, vt - vd);
}

// One colour phase; thread (t, j) owns cell i = 2t + ((j + colour) & 1),
// the mapping of apply_projection_even/odd_kernel (src/fluid.cu:264-280).
template <typename T>
__global__ void __launch_bounds__(256) k_project_colour(StepParams sp, const uint8_t* flags, T* u,
                                                        T* v, float* p, int colour) {
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int i = 2 * (int)(blockIdx.x * 64 + threadIdx.x) + ((j + colour) & 1);
  if (i < 1 || i > sp.W - 2 || j < 1 || j > sp.H - 2) return;
  const long k = (long)j * sp.P + i;
  const uint8_t f = flags[k];
  if (!(f & CF_UPD)) return;
  project_cell(sp, flags, u, v, p, k, f);
}

// ---- tiled kernel ----------------------------------------------------------

struct TiledArgs {
  int tw, th, k;  // owned tile, sweeps per launch (halo 2k)
  int pw, rh;     // region: pw cell pairs (2*pw cells) x rh rows
  int wpr;        // wavefronts per region row
  int ntx;        // tiles per grid row
  int sweeps;     // sweeps this launch (<= k)
  int forces;     // apply the velocity forces on load, pressure sums from 0
};

// LDS layout: U, V (and the pressure sums S) hold rh rows of two half-rows,
// the even columns of the region row then the odd ones, so the lanes that
// update one colour of a row touch consecutive words.  Lane l of column
// block cb handles cell pair q = 64*cb + l of every row its wavefront visits.
template <typename T, bool PRESSURE>
__global__ void __launch_bounds__(kTileThreads)
    k_project_tiled(StepParams sp, const uint8_t* __restrict__ flags, const T* __restrict__ u_in,
                    const T* __restrict__ v_in, T* __restrict__ u_out, T* __restrict__ v_out,
                    float* __restrict__ p, TiledArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int pw = a.pw, rh = a.rh, K = a.k;
  const int pwp = pw + 1;
  const int rs = 2 * pwp;  // LDS row stride (elements)
  const int rw = 2 * pw;   // region width (cells)
  const int fs = rw + 1;   // flag row stride (bytes)
  const int ncell = rh * rs;
  float* U = (float*)smem;
  float* V = U + ncell;
  float* S = V + ncell;  // pressure sums (PRESSURE only)
  uint8_t* G = (uint8_t*)(PRESSURE ? S + ncell : V + ncell);
  uint8_t* F = G + ncell;

  const int lane = threadIdx.x & 63;
  // wavefront-uniform values live in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cb = wave % a.wpr;             // column block of this wavefront
  const int r0 = wave / a.wpr;             // its first row
  const int rstep = kTileWaves / a.wpr;    // its row stride
  const int q = cb * 64 + lane;            // this lane's cell pair
  const bool lane_on = q < pw;

  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = tile % a.ntx, ty = tile / a.ntx;
  const int x0 = tx * a.tw, y0 = ty * a.th;
  const int gx0 = x0 - 2 * K, gy0 = y0 - 2 * K;  // region origin; gx0 is even
  const int W = sp.W, H = sp.H, P = sp.P;

  // flags of the region plus one column right and one row above
  for (int r = r0; r <= rh; r += rstep) {
    const int gj = gy0 + r;
    const bool row_in = (unsigned)gj < (unsigned)H;
    if (lane_on) {
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int gi = gx0 + 2 * q + b;
        F[r * fs + 2 * q + b] =
            (row_in && (unsigned)gi < (unsigned)W) ? flags[(long)gj * P + gi] : (uint8_t)0;
      }
    }
  }
  for (int r = threadIdx.x; r <= rh; r += kTileThreads) {
    const int gi = gx0 + rw, gj = gy0 + r;
    F[r * fs + rw] = ((unsigned)gi < (unsigned)W && (unsigned)gj < (unsigned)H)
                         ? flags[(long)gj * P + gi]
                         : (uint8_t)0;
  }
  // u, v (and pressure) of the region
  for (int r = r0; r < rh; r += rstep) {
    if (!lane_on) break;
    const int gj = gy0 + r;
    const int gi = gx0 + 2 * q;
    const bool row_in = (unsigned)gj < (unsigned)H;
    const bool in0 = row_in && (unsigned)gi < (unsigned)W;
    const bool in1 = row_in && (unsigned)(gi + 1) < (unsigned)W;
    const long gk = (long)gj * P + gi;
    float a0, a1, b0, b1;
    const int l0 = r * rs + q, l1 = l0 + pwp;
    load_pair(u_in, gk, in0, in1, a0, a1);
    load_pair(v_in, gk, in0, in1, b0, b1);
    if (a.forces) {  // the forces' results as stored (rounded to T)
      if (in0) force_uv(sp, gi, gj, a0, b0);
      if (in1) force_uv(sp, gi + 1, gj, a1, b1);
      a0 = rnd<T>(a0);
      a1 = rnd<T>(a1);
      b0 = rnd<T>(b0);
      b1 = rnd<T>(b1);
    }
    U[l0] = a0;
    U[l1] = a1;
    V[l0] = b0;
    V[l1] = b1;
    if constexpr (PRESSURE) {
      S[l0] = (in0 && !a.forces) ? p[gk] : 0.0f;
      S[l1] = (in1 && !a.forces) ? p[gk + 1] : 0.0f;
    }
  }
  __syncthreads();
  // gate bytes
  for (int r = r0; r < rh; r += rstep) {
    if (!lane_on) break;
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int lc = 2 * q + b;
      const uint8_t f = F[r * fs + lc];
      const uint8_t fr = F[r * fs + lc + 1];
      const uint8_t ft = F[(r + 1) * fs + lc];
      uint32_t g = 0;
      if (f & CF_UPD) {
        g = G_UPD;
        if (f & CF_FU) g |= G_L;
        if (fr & CF_FU) g |= G_R;
        if (f & CF_FV) g |= G_B;
        if (ft & CF_FV) g |= G_T;
        const int s = cf_total_s(f);
        if (s == 0 || s == 3)
          g |= G_SLOW;
        else
          g |= (uint32_t)(s >> 1) << G_L2_SHIFT;
      }
      G[r * rs + b * pwp + q] = (uint8_t)g;
    }
  }
  __syncthreads();

  const float o = sp.o;
  const int nphase = 2 * a.sweeps;
  for (int s = 0; s < nphase; s++) {
    const int c = s & 1;           // even cells first, as the reference
    const int e = nphase - 1 - s;  // phases still to come after this one
    const int lo = 2 * K - 1 - e;  // box of cells that still reach the tile
    const int chi = 2 * K + a.tw + e;
    const int rhi = 2 * K + a.th + e;
    int r = r0;
    if (lo > r0) r += ((lo - r0 + rstep - 1) / rstep) * rstep;
    // the pair's cell of colour c alternates with the row; b and the offset
    // of u(i+1) are the same for every lane of a row
    int b = (c + gy0 + r) & 1;
    const int bstep = rstep & 1;
    int own = r * rs + q + b * pwp;
    for (; r < rhi; r += rstep) {
      const int lc = 2 * q + b;
      if (lane_on && lc >= lo && lc < chi) {
        // every face a cell of this colour touches is touched by no other
        // cell of the phase, so the four writes are unconditional
        const int right = b ? own - pwp + 1 : own + pwp;  // u(i+1): next pair's even cell
        const int top = own + rs;
        const uint32_t g = G[own];
        const float uu = U[own], ur = U[right], vv = V[own], vt = V[top];
        const float div = ((ur - uu) + vt) - vv;
        const float d = relaxed_quotient(div, g, o);
        if constexpr (PRESSURE) {
          if (g & G_UPD) S[own] = S[own] + (((d * sp.density) * (float)sp.cs) / sp.d_t);
        }
        U[own] = rnd<T>(gated_add(uu, d, g, 0));
        U[right] = rnd<T>(gated_sub(ur, d, g, 1));
        V[own] = rnd<T>(gated_add(vv, d, g, 2));
        V[top] = rnd<T>(gated_sub(vt, d, g, 3));
      }
      own += bstep ? (b ? rstep * rs - pwp : rstep * rs + pwp) : rstep * rs;
      b ^= bstep;
    }
    __syncthreads();
  }

  // write back the owned tile
  for (int r = r0 + ((2 * K - r0 + rstep - 1) / rstep) * rstep; r < 2 * K + a.th; r += rstep) {
    if (!lane_on) break;
    const int gj = gy0 + r;
    if (gj >= H) break;
    const int lc = 2 * q;
    const int gi = gx0 + lc;
    const bool in0 = lc >= 2 * K && lc < 2 * K + a.tw && gi < W;
    const bool in1 = lc + 1 >= 2 * K && lc + 1 < 2 * K + a.tw && gi + 1 < W;
    if (!in0 && !in1) continue;
    const long gk = (long)gj * P + gi;
    const int l0 = r * rs + q, l1 = l0 + pwp;
    store_pair(u_out, gk, in0, in1, U[l0], U[l1]);
    store_pair(v_out, gk, in0, in1, V[l0], V[l1]);
    if constexpr (PRESSURE) {
      if (in0) p[gk] = S[l0];
      if (in1) p[gk + 1] = S[l1];
    }
  }
}

// 8 and 10 rows per wavefront are held within 128 VGPRs, so two workgroups
// share a CU (8 rows with pressure sums spill a few values outside the
// sweep loop, which beats one workgroup per CU); 4 rows within 80, so three
// share one (the others take what the compiler allocates).
template <typename T, bool PRESSURE, int R>
__global__ void __launch_bounds__(64 * kRegWaves, (R == 8 || R == 10) ? 4 : (R == 4 ? 6 : 1))
    k_project_regs(StepParams sp, const uint8_t* __restrict__ gates, const T* __restrict__ u_in,
                   const T* __restrict__ v_in, T* __restrict__ u_out, T* __restrict__ v_out,
                   float* __restrict__ p, RegArgs a) {
  __shared__ RegShared sh;
  project_regs_tile<T, PRESSURE, R, false>(sp, gates, u_in, v_in, u_out, v_out, p, a,
                                    xcd_tile(blockIdx.x, gridDim.x), sh);
}

template <typename T, bool PRESSURE>
void launch_tiled(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                  const ProjTuning& t, hipStream_t s) {
  TiledArgs a;
  a.tw = t.tw;
  a.th = t.th;
  a.k = t.k;
  a.pw = t.pw();
  a.rh = t.region_h();
  a.wpr = t.waves_per_row();
  a.ntx = (sp.W + a.tw - 1) / a.tw;
  const int nty = (sp.H + a.th - 1) / a.th;
  const size_t lds = t.lds_bytes(PRESSURE);
  for (int done = 0; done < sp.n; done += t.k) {
    a.sweeps = (sp.n - done < t.k) ? sp.n - done : t.k;
    a.forces = (done == 0 && sp.forces_in_projection) ? 1 : 0;
    const int nxt = 1 - *cur;
    hipLaunchKernelGGL((k_project_tiled<T, PRESSURE>), dim3(a.ntx * nty), dim3(kTileThreads), lds,
                       s, sp, g.flags, (const T*)f.u[*cur], (const T*)f.v[*cur], (T*)f.u[nxt],
                       (T*)f.v[nxt], f.p, a);
    *cur = nxt;
  }
}

// ---- resident kernel ---------------------------------------------------------
//
// All n sweeps in one launch: one workgroup per tile keeps its region in
// registers from the first sweep to the last.  The sweeps run in stages of
// k (the last stage takes the rest); after a stage the owned cells are
// exact and the halo is stale, so the workgroup stores the ring of owned
// cells its neighbours read as halo (those within 2k of the tile's edge),
// publishes the stage on its flag, waits for the adjacent tiles' flags and
// reloads only its halo from their rings.  Stage s reads buffer s & 1 and
// writes buffer (s + 1) & 1, so a ring is overwritten (two stages later)
// only after every neighbour has finished the stage that read it, which it
// did before publishing the stage this tile waited for.  The last stage
// stores the whole tile (and the pressure sums) into buffer nstages & 1.
// The last workgroup to finish clears the flags for the next launch.
//
// Every tile's workgroup must be resident at once (the host checks the
// tile count against resident_capacity); a wait is bounded by spin_ticks,
// after which the error word is set and the launch drains without waiting.
struct ResidentArgs {
  int nstages;      // ceil(n / k)
  int last_sweeps;  // sweeps of the last stage
  int nty;          // tile rows
  unsigned* flags;  // [tiles]: stages published; [tiles]: finished workgroups
  int* error;
  long long spin_ticks;
};

// Wavefront 0: wait until the adjacent tiles (up to 8, one per lane) have
// published `stage`.  False on timeout, or at once when another wait of the
// launch has timed out.  The rings are read with agent-coherent loads
// issued after this wait (behind the workgroup barrier that follows it), so
// no cache needs invalidating.
__device__ inline bool wait_neighbours(const ResidentArgs& ra, int ntx, int tile, unsigned stage) {
  const int lane = (int)(threadIdx.x & 63);
  const int tx = tile % ntx, ty = tile / ntx;
  const int nx = tx + lane % 3 - 1, ny = ty + lane / 3 - 1;
  const bool poll = lane < 9 && lane != 4 && (unsigned)nx < (unsigned)ntx &&
                    (unsigned)ny < (unsigned)ra.nty;
  const unsigned* f = ra.flags + (poll ? ny * ntx + nx : 0);
  const long long t0 = wall_clock64();
  for (;;) {
    const bool ready =
        !poll || __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= stage;
    if (__all(ready)) break;
    if (__hip_atomic_load(ra.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (wall_clock64() - t0 > ra.spin_ticks) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

template <typename T, bool PRESSURE, int R>
__global__ void __launch_bounds__(64 * kRegWaves, (R == 8 || R == 10) ? 4 : (R == 4 ? 6 : 1))
    k_project_resident(StepParams sp, const uint8_t* __restrict__ gates, T* u0, T* v0, T* u1,
                       T* v1, float* __restrict__ p, RegArgs a, ResidentArgs ra) {
  __shared__ RegShared sh;
  __shared__ int s_last;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const RegPlace pl = reg_place<R>(a, tile);
  RegTile<R> t;
  reg_load<T, PRESSURE, R>(t, sp, gates, u0, v0, p, a, pl);
  for (int s = 0; s < ra.nstages; s++) {
    const bool odd = (s & 1) != 0;
    if (s > 0) {
      if (threadIdx.x < 64) {
        const bool ok = wait_neighbours(ra, a.ntx, tile, (unsigned)s);
        if (!ok && threadIdx.x == 0)
          __hip_atomic_store(ra.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      // the barrier orders every wavefront's halo loads after the wait
      __syncthreads();
      reg_reload_halo<T, R>(t, odd ? u1 : u0, odd ? v1 : v0, sp, a, pl);
    }
    const bool last = s + 1 == ra.nstages;
    // one barrier per colour phase (measured 0.4% faster than two here)
    reg_sweeps<T, PRESSURE, R, true>(t, sh, sp, pl, a.th, last ? ra.last_sweeps : a.k);
    T* uo = odd ? u0 : u1;
    T* vo = odd ? v0 : v1;
    if (!last) {
      reg_store<T, PRESSURE, R, true>(t, uo, vo, p, sp, a, pl);
      // every wavefront's ring stores (agent-coherent) have completed before
      // the barrier, so they are visible GPU-wide before the flag is
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_store(ra.flags + tile, (unsigned)(s + 1), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
      reg_store<T, PRESSURE, R, false>(t, uo, vo, p, sp, a, pl);
    }
  }
  // every other workgroup has finished waiting when the last one counts
  // itself in, so that one may clear the flags
  const int ntiles = a.ntx * ra.nty;
  if (threadIdx.x == 0)
    s_last = __hip_atomic_fetch_add(ra.flags + ntiles, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) == (unsigned)(ntiles - 1);
  __syncthreads();
  if (s_last) {
    for (int i = (int)threadIdx.x; i <= ntiles; i += (int)blockDim.x)
      __hip_atomic_store(ra.flags + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename T, bool PRESSURE, int R>
const void* resident_kernel_r() {
  return (const void*)k_project_resident<T, PRESSURE, R>;
}

template <typename T, bool PRESSURE>
const void* resident_kernel(int rows) {
  switch (rows) {
    case 4: return resident_kernel_r<T, PRESSURE, 4>();
    case 8: return resident_kernel_r<T, PRESSURE, 8>();
    case 10: return resident_kernel_r<T, PRESSURE, 10>();
    case 12: return resident_kernel_r<T, PRESSURE, 12>();
    default: return resident_kernel_r<T, PRESSURE, 16>();
  }
}

template <typename T, bool PRESSURE>
void launch_resident(const StepParams& sp, const FieldPtrs& f, int* cur, const ProjTuning& t,
                     long long spin_ticks, hipStream_t s) {
  if (sp.n <= 0) return;
  RegArgs a;
  a.tw = t.tw;
  a.th = t.th;
  a.k = t.k;
  a.ntx = (sp.W + a.tw - 1) / a.tw;
  a.sweeps = t.k;
  a.forces = sp.forces_in_projection ? 1 : 0;
  ResidentArgs ra;
  ra.nstages = (sp.n + t.k - 1) / t.k;
  ra.last_sweeps = sp.n - (ra.nstages - 1) * t.k;
  ra.nty = (sp.H + a.th - 1) / a.th;
  ra.flags = f.rflags;
  ra.error = f.error;
  ra.spin_ticks = spin_ticks;
  const int in = *cur, other = 1 - in;
  T* u0 = (T*)f.u[in];
  T* v0 = (T*)f.v[in];
  T* u1 = (T*)f.u[other];
  T* v1 = (T*)f.v[other];
  void* args[] = {(void*)&sp, (void*)&f.gates, (void*)&u0, (void*)&v0, (void*)&u1,
                  (void*)&v1, (void*)&f.p, (void*)&a, (void*)&ra};
  (void)hipLaunchKernel(resident_kernel<T, PRESSURE>(t.reg_rows()), dim3(a.ntx * ra.nty),
                        dim3(64 * kRegWaves), args, 0, s);
  if (ra.nstages & 1) *cur = other;
}

template <typename T, bool PRESSURE, int R>
void launch_regs_r(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                   const ProjTuning& t, hipStream_t s) {
  RegArgs a;
  a.tw = t.tw;
  a.th = t.th;
  a.k = t.k;
  a.ntx = (sp.W + a.tw - 1) / a.tw;
  const int nty = (sp.H + a.th - 1) / a.th;
  for (int done = 0; done < sp.n; done += t.k) {
    a.sweeps = (sp.n - done < t.k) ? sp.n - done : t.k;
    a.forces = (done == 0 && sp.forces_in_projection) ? 1 : 0;
    const int nxt = 1 - *cur;
    hipLaunchKernelGGL((k_project_regs<T, PRESSURE, R>), dim3(a.ntx * nty), dim3(64 * kRegWaves),
                       0, s, sp, f.gates, (const T*)f.u[*cur], (const T*)f.v[*cur], (T*)f.u[nxt],
                       (T*)f.v[nxt], f.p, a);
    *cur = nxt;
  }
}

template <typename T, bool PRESSURE>
void launch_regs(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                 const ProjTuning& t, hipStream_t s) {
  switch (t.reg_rows()) {
    case 4: launch_regs_r<T, PRESSURE, 4>(sp, g, f, cur, t, s); break;
    case 8: launch_regs_r<T, PRESSURE, 8>(sp, g, f, cur, t, s); break;
    case 10: launch_regs_r<T, PRESSURE, 10>(sp, g, f, cur, t, s); break;
    case 12: launch_regs_r<T, PRESSURE, 12>(sp, g, f, cur, t, s); break;
    default: launch_regs_r<T, PRESSURE, 16>(sp, g, f, cur, t, s); break;
  }
}

template <typename T>
hipError_t allow_large_lds() {
  const int bytes = (int)kLdsLimit;
  hipError_t e = hipFuncSetAttribute((const void*)k_project_tiled<T, false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_project_tiled<T, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  return e;
}


}  // namespace

hipError_t init_projection_kernels() {
  hipError_t e = allow_large_lds<float>();
  if (e == hipSuccess) e = allow_large_lds<__half>();
  if (e == hipSuccess) e = allow_large_lds<__hip_bfloat16>();
  return e;
}

// Workgroups of kRegWaves wavefronts a CU holds at once, from the register
// kernels' launch bounds (which the compiler meets): at most 80 VGPRs at 4
// rows (6 wavefronts per SIMD: 3 workgroups), 128 at 8 and 10 rows (4: 2
// workgroups), up to 256 otherwise (1).  LDS (12 KiB per workgroup) and the
// wavefront slots allow more.  (The runtime's occupancy query was not used:
// it has returned 0 for a kernel that fits, depending on what ran before.)
hipError_t resident_capacity(int rows, int* slots) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int per_cu = rows == 4 ? 3 : (rows == 8 || rows == 10) ? 2 : 1;
  *slots = e == hipSuccess ? per_cu * cus : 0;
  return e;
}

template <typename T>
void enqueue_projection(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                        const ProjTuning& t, hipStream_t s) {
  if (t.kernel == PROJ_RESIDENT && t.valid(sp.pressure_on)) {
    if (sp.pressure_on)
      launch_resident<T, true>(sp, f, cur, t, sp.spin_ticks, s);
    else
      launch_resident<T, false>(sp, f, cur, t, sp.spin_ticks, s);
    return;
  }
  if (t.kernel == PROJ_REGS && t.valid(sp.pressure_on)) {
    if (sp.pressure_on)
      launch_regs<T, true>(sp, g, f, cur, t, s);
    else
      launch_regs<T, false>(sp, g, f, cur, t, s);
    return;
  }
  if (t.kernel == PROJ_TILED && t.valid(sp.pressure_on)) {
    if (sp.pressure_on)
      launch_tiled<T, true>(sp, g, f, cur, t, s);
    else
      launch_tiled<T, false>(sp, g, f, cur, t, s);
    return;
  }
  T* u = (T*)f.u[*cur];
  T* v = (T*)f.v[*cur];
  const dim3 grid((sp.W / 2 + 63) / 64 + 1, (sp.H + 3) / 4), block(64, 4);
  for (int sweep = 0; sweep < sp.n; sweep++) {
    hipLaunchKernelGGL(k_project_colour<T>, grid, block, 0, s, sp, g.flags, u, v, f.p, 0);
    hipLaunchKernelGGL(k_project_colour<T>, grid, block, 0, s, sp, g.flags, u, v, f.p, 1);
  }
}

template void enqueue_projection<float>(const StepParams&, const GridView&, const FieldPtrs&, int*,
                                        const ProjTuning&, hipStream_t);
template void enqueue_projection<__half>(const StepParams&, const GridView&, const FieldPtrs&,
                                         int*, const ProjTuning&, hipStream_t);
template void enqueue_projection<__hip_bfloat16>(const StepParams&, const GridView&,
                                                 const FieldPtrs&, int*, const ProjTuning&,
                                                 hipStream_t);

}  // namespace osy
