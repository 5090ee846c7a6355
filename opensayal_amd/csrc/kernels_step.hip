// kernels_step.hip - the per-step phases of Fluid::update (src/fluid.cu:770-795)
// as HIP kernels for gfx950, plus mask build, reductions and layout copies.
//
// One step is: the projection launches (the first of which applies the
// external forces to every cell it loads), the velocity advection (which
// reads the projected field through the edge extrapolation, so the
// extrapolated field is never stored), and the smoke advection (which
// applies the smoke part of the forces as it loads and the decay as it
// stores).  Both advections write into the other half of a ping-pong pair
// instead of the reference's buffer-then-copy-back (src/fluid.cu:614-617,
// 569-571): same values, half the traffic.  The advections stage a tile of
// the fields plus a 4-cell halo in LDS with coalesced loads; a back-trace
// that leaves the staged region reads HBM directly.
#include "advect_device.hpp"
#include "fluid_device.hpp"
#include "kernels.hpp"
#include "batch.hpp"
#include "projection.hpp"
#include "projection_device.hpp"
#include "step_device.hpp"

#pragma clang fp contract(off)

namespace osy {
namespace {

constexpr int kBX = 64, kBY = 4;

inline dim3 cell_grid(int W, int H) {
  return dim3((W + kBX - 1) / kBX, (H + kBY - 1) / kBY);
}

// ---- mask build: Fluid::init_device_memory (src/fluid.cu:111-142) -------

// is_solid of the reference.  The obstacle test there is
// sqrt(pow(dx,2) + pow(dy,2)) < radius in double.  For an integer d2 and a
// radius that is a float widened to double, RN(sqrt(d2)) < r holds exactly
// when r > 0 and d2 < r*r, both sides exact in double (d2 < 2^51 and r*r
// has at most 48 significant bits); the test restates it that way so the
// result does not depend on the device's square root.
__device__ bool solid_at(const MaskParams& m, int i, int j) {
  const double dx = (double)i - m.cx, dy = (double)j - m.cy;
  const double d2 = dx * dx + dy * dy;
  const bool obstacle = m.obstacle_enable && m.radius > 0.0 && d2 < m.radius * m.radius;
  const bool pipe = i < m.pipe_length &&
                    (j == m.H / 2 - m.pipe_height / 2 - 1 || j == m.H / 2 + m.pipe_height / 2 + 1);
  return i == 0 || j == 0 || j == m.H - 1 || (!m.enable_drain && i == m.W - 1) || obstacle ||
         pipe;
}

__global__ void k_build_flags(MaskParams m, uint8_t* flags) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= m.P || j >= m.H) return;
  if (i >= m.W) {  // row padding: never read as fluid
    flags[(long)j * m.P + i] = 0;
    return;
  }
  const bool s = solid_at(m, i, j);
  const bool l = i > 0 && !solid_at(m, i - 1, j);
  const bool r = i + 1 < m.W && !solid_at(m, i + 1, j);
  const bool b = j > 0 && !solid_at(m, i, j - 1);
  const bool t = j + 1 < m.H && !solid_at(m, i, j + 1);
  const int total_s = (int)l + (int)r + (int)b + (int)t;
  uint8_t f = (uint8_t)(total_s << CF_S_SHIFT);
  if (!s) {
    f |= CF_FLUID;
    if (i >= 1 && i <= m.W - 2 && j >= 1 && j <= m.H - 2) f |= CF_UPD;
    if (l) f |= CF_FU;
    if (b) f |= CF_FV;
  }
  flags[(long)j * m.P + i] = f;
}

// The projection's gate byte of every cell (gate_byte: its face gates,
// projected bit and quotient class), from its flags and those of its right
// and top neighbours; 0 outside the grid and in the row padding.
__global__ void k_build_gates(MaskParams m, const uint8_t* flags, uint8_t* gates) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= m.P || j >= m.H) return;
  const long k = (long)j * m.P + i;
  if (i >= m.W) {
    gates[k] = 0;
    return;
  }
  const uint32_t f = flags[k];
  const uint32_t fr = i + 1 < m.W ? flags[k + 1] : 0u;
  const uint32_t ft = j + 1 < m.H ? flags[k + m.P] : 0u;
  gates[k] = (uint8_t)gate_byte(f, fr, ft);
}

// ---- P1: forces over 64 x 16 tiles (forces_tile), four rows per thread --
constexpr int kForcesTY = 16;
template <typename T>
__global__ void __launch_bounds__(kBX * kBY)
    k_forces(StepParams sp, int do_uv, int do_smoke, T* u, T* v, T* smoke, float* p) {
  static_assert(kBX == kAdvTX && kForcesTY % kBY == 0, "forces tile shape");
  forces_tile<T, kForcesTY, kBY>(sp, do_uv, do_smoke, u, v, smoke, p, blockIdx.x, blockIdx.y);
}

// ---- P5: thrust::reduce min/max of pressure (src/fluid.cu:778-787) -------

__device__ inline float wave_min(float x) {
  for (int o = 32; o > 0; o >>= 1) x = fminf(x, __shfl_xor(x, o));
  return x;
}
__device__ inline float wave_max(float x) {
  for (int o = 32; o > 0; o >>= 1) x = fmaxf(x, __shfl_xor(x, o));
  return x;
}

__global__ void __launch_bounds__(256) k_minmax_partial(StepParams sp, const float* p, float* part) {
  float lo = INFINITY, hi = -INFINITY;
  for (int j = blockIdx.x; j < sp.H; j += gridDim.x)
    for (int i = threadIdx.x; i < sp.W; i += blockDim.x) {
      const float x = p[(long)j * sp.P + i];
      lo = fminf(lo, x);
      hi = fmaxf(hi, x);
    }
  __shared__ float slo[4], shi[4];
  lo = wave_min(lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) {
      lo = fminf(lo, slo[w]);
      hi = fmaxf(hi, shi[w]);
    }
    part[2 * blockIdx.x] = lo;
    part[2 * blockIdx.x + 1] = hi;
  }
}

__global__ void __launch_bounds__(256) k_minmax_final(const float* part, int nparts, float* out) {
  float lo = INFINITY, hi = -INFINITY;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
    lo = fminf(lo, part[2 * b]);
    hi = fmaxf(hi, part[2 * b + 1]);
  }
  __shared__ float slo[4], shi[4];
  lo = wave_min(lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) {
      lo = fminf(lo, slo[w]);
      hi = fmaxf(hi, shi[w]);
    }
    out[0] = lo;
    out[1] = hi;
  }
}

// ---- P6-P8: the advections over 64 x 32 tiles, 512 threads ---------------
constexpr int kAdvTY = 32, kAdvNY = 8;

// 1-D grids of ntx x nty tiles in the XCD-aware order (xcd_tile).
template <typename T, bool CS1>
__global__ void __launch_bounds__(64 * kAdvNY)
    k_advect_velocity(StepParams sp, GridView g, const T* __restrict__ u,
                      const T* __restrict__ v, T* __restrict__ u_out, T* __restrict__ v_out,
                      int ntx) {
  __shared__ AdvShared<kAdvTY> sh;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  advect_velocity_tile<T, CS1, kAdvTY, kAdvNY>(sp, g, u, v, u_out, v_out, tile % ntx, tile / ntx,
                                               sh);
}

template <typename T, bool CS1>
__global__ void __launch_bounds__(64 * kAdvNY)
    k_advect_smoke(StepParams sp, GridView g, const T* __restrict__ u, const T* __restrict__ v,
                   const T* __restrict__ smoke, T* __restrict__ smoke_out, int ntx) {
  __shared__ AdvShared<kAdvTY> sh;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  advect_smoke_tile<T, CS1, kAdvTY, kAdvNY>(sp, g, u, v, smoke, smoke_out, tile % ntx,
                                            tile / ntx, sh);
}

// ---- P2b: apply_diffusion (src/fluid.cu:167-190) as Jacobi rounds -------
// Every interior cell (1 <= i <= W-2, 1 <= j <= H-2, solid or not) becomes
// (u + a*(((u(i-1,j) + u(i+1,j)) + u(i,j-1)) + u(i,j+1))) / (1 + 4a) of the
// previous round's u; borders keep theirs; v is not diffused.  A launch runs
// up to kDiffHalo rounds on a 128 x 64 region (the owned tile plus a
// kDiffHalo-cell halo) staged in LDS, two buffers in turn: a value at region
// distance d from the region edge is exact for d rounds, so the owned tile
// is exact after kDiffHalo rounds.  Eight wavefronts take the region's rows
// in turn, each lane two columns 64 apart, so the 126 updated columns fill
// 126 of 128 lanes.  u goes from u_in to u_out; with copy_v the launch also
// moves v to v_out so that u and v stay in the same buffer of their pairs.
constexpr int kDiffRX = 128, kDiffRY = 64, kDiffHalo = 6;
static_assert(kDiffRY % 8 == 0, "eight wavefronts take the region's rows in turn");
constexpr int kDiffTX = kDiffRX - 2 * kDiffHalo, kDiffTY = kDiffRY - 2 * kDiffHalo;

template <typename T>
__global__ void __launch_bounds__(512)
    k_diffuse(StepParams sp, int rounds, const T* __restrict__ u_in, T* __restrict__ u_out,
              const T* __restrict__ v_in, T* __restrict__ v_out, int copy_v, int ntx) {
  __shared__ float buf[2][kDiffRY * kDiffRX];
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int x0 = (tile % ntx) * kDiffTX - kDiffHalo, y0 = (tile / ntx) * kDiffTY - kDiffHalo;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  // stage the region: every load of the thread (from the cell clamped into
  // the grid, so with no branch) before its first LDS write, so a wavefront
  // waits for HBM once rather than once per row; cells outside the grid
  // hold 0
  constexpr int kRowsPerWave = kDiffRY / 8;
  float staged[kRowsPerWave][2];
#pragma unroll
  for (int m = 0; m < kRowsPerWave; m++) {
    const int j = min(max(y0 + wave + 8 * m, 0), sp.H - 1);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int i = min(max(x0 + lane + 64 * h, 0), sp.W - 1);
      staged[m][h] = ldf(u_in, (long)j * sp.P + i);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < kRowsPerWave; m++) {
    const int r = wave + 8 * m, j = y0 + r;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int c = lane + 64 * h, i = x0 + c;
      const bool in = (unsigned)i < (unsigned)sp.W && (unsigned)j < (unsigned)sp.H;
      const float x = in ? staged[m][h] : 0.0f;
      buf[0][r * kDiffRX + c] = x;
      buf[1][r * kDiffRX + c] = x;
    }
  }
  __syncthreads();
  const float a = sp.visc_a, den = sp.visc_den;
  // columns 1 + lane and 65 + lane (the second only for lanes 0..61)
  const int c0 = 1 + lane, c1 = 65 + lane;
  const bool in0 = x0 + c0 >= 1 && x0 + c0 <= sp.W - 2;
  const bool in1 = c1 < kDiffRX - 1 && x0 + c1 >= 1 && x0 + c1 <= sp.W - 2;
  for (int it = 0; it < rounds; it++) {
    const float* src = buf[it & 1];
    float* dst = buf[(it & 1) ^ 1];
#pragma unroll
    for (int m = 0; m < kRowsPerWave; m++) {
      const int r = 1 + wave + 8 * m, j = y0 + r;
      if (r >= kDiffRY - 1 || j < 1 || j > sp.H - 2) continue;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (!(h ? in1 : in0)) continue;
        const int l = r * kDiffRX + (h ? c1 : c0);
        const float sum = ((src[l - 1] + src[l + 1]) + src[l - kDiffRX]) + src[l + kDiffRX];
        dst[l] = rnd<T>((src[l] + a * sum) / den);
      }
    }
    __syncthreads();
  }
  const float* res = buf[rounds & 1];
  for (int r = kDiffHalo + wave; r < kDiffHalo + kDiffTY; r += 8) {
    const int j = y0 + r;
    if (j >= sp.H) break;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int c = kDiffHalo + lane + 64 * h, i = x0 + c;
      if (c >= kDiffHalo + kDiffTX || i >= sp.W) continue;
      const long k = (long)j * sp.P + i;
      stf(u_out, k, res[r * kDiffRX + c]);
      if (copy_v) v_out[k] = v_in[k];
    }
  }
}

// `sp.n` rounds in ceil(n / kDiffHalo) launches; advances *uv by the launch
// count (v moves with u on the last launch when that count is odd).
template <typename T>
void enqueue_diffusion(const StepParams& sp, const FieldPtrs& f, int* uv, hipStream_t s) {
  if (sp.n <= 0) return;
  const int ntx = (sp.W + kDiffTX - 1) / kDiffTX, nty = (sp.H + kDiffTY - 1) / kDiffTY;
  const int launches = (sp.n + kDiffHalo - 1) / kDiffHalo;
  int cur = *uv;
  for (int q = 0; q < launches; q++) {
    const int rounds = q + 1 < launches ? kDiffHalo : sp.n - kDiffHalo * (launches - 1);
    const int copy_v = (q + 1 == launches) && (launches & 1);
    hipLaunchKernelGGL(k_diffuse<T>, dim3(ntx * nty), dim3(512), 0, s, sp, rounds,
                       (const T*)f.u[cur], (T*)f.u[1 - cur], (const T*)f.v[cur],
                       (T*)f.v[1 - cur], copy_v, ntx);
    cur = 1 - cur;
  }
  *uv = cur;
}

// ---- divergence norms (Fluid::get_divergence, src/fluid.cu:192-201) -------
__device__ inline double wave_sum_d(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ inline double wave_max_d(double x) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
  return x;
}

template <typename T>
__global__ void __launch_bounds__(256) k_div_partial(StepParams sp, GridView g, const T* u,
                                                     const T* v, double* part) {
  double ss = 0.0, mx = 0.0;
  for (int j = 1 + blockIdx.x; j <= sp.H - 2; j += gridDim.x)
    for (int i = 1 + threadIdx.x; i <= sp.W - 2; i += blockDim.x) {
      const long k = g.at(i, j);
      if (!(g.flags[k] & CF_FLUID)) continue;
      const double d = (((double)ldf(u, k + 1) - (double)ldf(u, k)) + (double)ldf(v, k + sp.P)) -
                       (double)ldf(v, k);
      ss += d * d;
      mx = fmax(mx, fabs(d));
    }
  __shared__ double sss[4], smx[4];
  ss = wave_sum_d(ss);
  mx = wave_max_d(mx);
  if ((threadIdx.x & 63) == 0) {
    sss[threadIdx.x >> 6] = ss;
    smx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) {
      ss += sss[w];
      mx = fmax(mx, smx[w]);
    }
    part[2 * blockIdx.x] = ss;
    part[2 * blockIdx.x + 1] = mx;
  }
}

__global__ void __launch_bounds__(64) k_div_final(const double* part, int nparts, double* out) {
  double ss = 0.0, mx = 0.0;
  for (int b = threadIdx.x; b < nparts; b += 64) {
    ss += part[2 * b];
    mx = fmax(mx, part[2 * b + 1]);
  }
  ss = wave_sum_d(ss);
  mx = wave_max_d(mx);
  if (threadIdx.x == 0) {
    out[0] = sqrt(ss);
    out[1] = mx;
  }
}

// ---- layout copies --------------------------------------------------------
// mode 0: device field (T) -> reference float; 1: reference float -> device
// field; 2: flags -> is_solid; 3: flags -> total_s; 4: float device field
// (pressure) -> reference; 5: reference -> pressure.
template <typename T>
__global__ void k_layout(int mode, int W, int H, int P, T* dev, float* pdev, const uint8_t* flags,
                         void* ref) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= W || j >= H) return;
  const long k = (long)j * P + i;
  const long r = (long)(H - 1 - j) * W + i;  // Fluid::indx (src/fluid.cu:163-165)
  switch (mode) {
    case 0: ((float*)ref)[r] = ldf(dev, k); break;
    case 1: stf(dev, k, ((const float*)ref)[r]); break;
    case 2: ((int32_t*)ref)[r] = (flags[k] & CF_FLUID) ? 0 : 1; break;
    case 3: ((int32_t*)ref)[r] = cf_total_s(flags[k]); break;
    case 4: ((float*)ref)[r] = pdev[k]; break;
    case 5: pdev[k] = ((const float*)ref)[r]; break;
  }
}

template <typename T, bool CS1>
void advect_t(const StepParams& sp, const GridView& g, const FieldPtrs& f, int cur, int* uv,
              int* sm, hipStream_t s) {
  const int ntx = (sp.W + kAdvTX - 1) / kAdvTX, nty = (sp.H + kAdvTY - 1) / kAdvTY;
  const dim3 grid(ntx * nty), block(64 * kAdvNY);
  const int nxt = 1 - cur;
  T* u2 = (T*)f.u[nxt];
  T* v2 = (T*)f.v[nxt];
  hipLaunchKernelGGL((k_advect_velocity<T, CS1>), grid, block, 0, s, sp, g, (const T*)f.u[cur],
                     (const T*)f.v[cur], u2, v2, ntx);
  *uv = nxt;
  if (sp.smoke_on) {
    hipLaunchKernelGGL((k_advect_smoke<T, CS1>), grid, block, 0, s, sp, g, (const T*)u2,
                       (const T*)v2, (const T*)f.smoke[*sm], (T*)f.smoke[1 - *sm], ntx);
    *sm = 1 - *sm;
  }
}

template <typename T>
void step_t(const StepParams& sp_in, const FieldPtrs& f, int* uv, int* sm,
            const ProjTuning& tuning, hipStream_t s) {
  const GridView g{sp_in.W, sp_in.H, sp_in.P, sp_in.cs, f.flags};
  StepParams sp = sp_in;
  // the first tiled projection launch applies the velocity forces (and
  // starts the pressure sums at 0) as it loads; the smoke advection applies
  // the smoke forces as it loads
  // (with diffusion the forces come first as a pass of their own, then the
  // diffusion launches, src/fluid.cu:771-776)
  const bool diffuse = sp.visc_on && sp.n > 0;
  const bool fuse_uv =
      !diffuse && sp.n > 0 && tuning.kernel != PROJ_COLOUR && tuning.valid(sp.pressure_on);
  const bool fuse_smoke = sp.smoke_on;
  sp.forces_in_projection = fuse_uv ? 1 : 0;
  if (!fuse_uv || !fuse_smoke)
    hipLaunchKernelGGL(k_forces<T>,
                       dim3((sp.W + kBX - 1) / kBX, (sp.H + kForcesTY - 1) / kForcesTY),
                       dim3(kBX * kBY), 0, s, sp,
                       fuse_uv ? 0 : 1, fuse_smoke ? 0 : 1, (T*)f.u[*uv], (T*)f.v[*uv],
                       (T*)f.smoke[*sm], f.p);
  int cur = *uv;
  if (diffuse) enqueue_diffusion<T>(sp, f, &cur, s);
  enqueue_projection<T>(sp, g, f, &cur, tuning, s);
  if (sp.pressure_on) {
    hipLaunchKernelGGL(k_minmax_partial, dim3(kReduceBlocks), dim3(256), 0, s, sp, f.p, f.scratch);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, f.scratch, kReduceBlocks,
                       f.prange);
  }
  if (sp.cs == 1)
    advect_t<T, true>(sp, g, f, cur, uv, sm, s);
  else
    advect_t<T, false>(sp, g, f, cur, uv, sm, s);
}

template <typename T>
void divergence_t(const StepParams& sp, const FieldPtrs& f, int uv, hipStream_t s) {
  const GridView g{sp.W, sp.H, sp.P, sp.cs, f.flags};
  hipLaunchKernelGGL(k_div_partial<T>, dim3(kReduceBlocks), dim3(256), 0, s, sp, g,
                     (const T*)f.u[uv], (const T*)f.v[uv], f.dscratch);
  hipLaunchKernelGGL(k_div_final, dim3(1), dim3(64), 0, s, f.dscratch, kReduceBlocks,
                     f.dscratch + 2 * kReduceBlocks);
}

template <typename T>
void layout_t(int field, bool to_ref, const StepParams& sp, const FieldPtrs& f, int uv, int sm,
              void* ref, hipStream_t s) {
  T* dev = nullptr;
  int mode = to_ref ? 0 : 1;
  switch (field) {
    case OSY_FIELD_VEL_X: dev = (T*)f.u[uv]; break;
    case OSY_FIELD_VEL_Y: dev = (T*)f.v[uv]; break;
    case OSY_FIELD_SMOKE: dev = (T*)f.smoke[sm]; break;
    case OSY_FIELD_PRESSURE: mode = to_ref ? 4 : 5; break;
    case OSY_FIELD_IS_SOLID: mode = 2; break;
    case OSY_FIELD_TOTAL_S: mode = 3; break;
  }
  hipLaunchKernelGGL(k_layout<T>, cell_grid(sp.W, sp.H), dim3(kBX, kBY), 0, s, mode, sp.W, sp.H,
                     sp.P, dev, f.p, f.flags, ref);
}

}  // namespace

void launch_build_flags(const MaskParams& m, uint8_t* flags, uint8_t* gates, hipStream_t s) {
  hipLaunchKernelGGL(k_build_flags, cell_grid(m.P, m.H), dim3(kBX, kBY), 0, s, m, flags);
  hipLaunchKernelGGL(k_build_gates, cell_grid(m.P, m.H), dim3(kBX, kBY), 0, s, m, flags, gates);
}

void enqueue_step(int storage, const StepParams& sp, const FieldPtrs& f, int* uv, int* sm,
                  const ProjTuning& tuning, hipStream_t s) {
  switch (storage) {
    case OSY_STORAGE_F16: step_t<__half>(sp, f, uv, sm, tuning, s); break;
    case OSY_STORAGE_BF16: step_t<__hip_bfloat16>(sp, f, uv, sm, tuning, s); break;
    default: step_t<float>(sp, f, uv, sm, tuning, s); break;
  }
}

void enqueue_projection_only(int storage, const StepParams& sp_in, const FieldPtrs& f, int* uv,
                             const ProjTuning& tuning, hipStream_t s) {
  const GridView g{sp_in.W, sp_in.H, sp_in.P, sp_in.cs, f.flags};
  StepParams sp = sp_in;
  sp.forces_in_projection = 0;
  switch (storage) {
    case OSY_STORAGE_F16: enqueue_projection<__half>(sp, g, f, uv, tuning, s); break;
    case OSY_STORAGE_BF16:
      enqueue_projection<__hip_bfloat16>(sp, g, f, uv, tuning, s);
      break;
    default: enqueue_projection<float>(sp, g, f, uv, tuning, s); break;
  }
}

namespace {
__global__ void __launch_bounds__(256) k_copy(uint4* __restrict__ dst,
                                              const uint4* __restrict__ src, size_t n) {
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
    dst[k] = src[k];
}
}  // namespace

void enqueue_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (uint4*)dst,
                     (const uint4*)src, n);
}

void enqueue_pressure_range(const StepParams& sp, const FieldPtrs& f, hipStream_t s) {
  hipLaunchKernelGGL(k_minmax_partial, dim3(kReduceBlocks), dim3(256), 0, s, sp, f.p, f.scratch);
  hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, f.scratch, kReduceBlocks, f.prange);
}

void enqueue_divergence(int storage, const StepParams& sp, const FieldPtrs& f, int uv,
                        hipStream_t s) {
  switch (storage) {
    case OSY_STORAGE_F16: divergence_t<__half>(sp, f, uv, s); break;
    case OSY_STORAGE_BF16: divergence_t<__hip_bfloat16>(sp, f, uv, s); break;
    default: divergence_t<float>(sp, f, uv, s); break;
  }
}

void enqueue_export(int field, int storage, const StepParams& sp, const FieldPtrs& f, int uv,
                    int sm, void* ref, hipStream_t s) {
  switch (storage) {
    case OSY_STORAGE_F16: layout_t<__half>(field, true, sp, f, uv, sm, ref, s); break;
    case OSY_STORAGE_BF16: layout_t<__hip_bfloat16>(field, true, sp, f, uv, sm, ref, s); break;
    default: layout_t<float>(field, true, sp, f, uv, sm, ref, s); break;
  }
}

void enqueue_import(int field, int storage, const StepParams& sp, const FieldPtrs& f, int uv,
                    int sm, const void* ref, hipStream_t s) {
  switch (storage) {
    case OSY_STORAGE_F16: layout_t<__half>(field, false, sp, f, uv, sm, (void*)ref, s); break;
    case OSY_STORAGE_BF16:
      layout_t<__hip_bfloat16>(field, false, sp, f, uv, sm, (void*)ref, s);
      break;
    default: layout_t<float>(field, false, sp, f, uv, sm, (void*)ref, s); break;
  }
}

}  // namespace osy
