// kernels_step.hip - the per-step phases of Fluid::update (src/fluid.cu:770-795)
// as HIP kernels for gfx950, plus mask build, reductions and layout copies.
//
// Phase kernels read and write HBM once per phase.  u/v and smoke advection
// write into the other half of a ping-pong pair instead of the reference's
// buffer-then-copy-back (src/fluid.cu:614-617, 569-571), which gives the same
// values with half the traffic.  All kernels launch 64-wide rows of lanes
// along i (one wavefront = 64 consecutive cells of a row, coalesced).
#include "fluid_device.hpp"
#include "kernels.hpp"
#include "projection.hpp"

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

// ---- P1: apply_external_forces_at (src/fluid.cu:308-349), zero_pressure ---

template <typename T>
__global__ void k_forces(StepParams sp, T* u, T* v, T* smoke, float* p) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= sp.W || j >= sp.H) return;
  const long k = (long)j * sp.P + i;
  float uu = ldf(u, k), vv = ldf(v, k);
  bool u_set = false, sm_set = false;
  float sm = 0.0f;
  if (i <= sp.smoke_length && i != 0 && j >= sp.band_lo && j <= sp.band_hi) {
    uu = sp.speed;
    u_set = true;
    if ((sp.smoke_count == 1 && j >= sp.sband_lo && j <= sp.sband_hi) ||
        (sp.smoke_count != 1 && (sp.smoke_top - j) % sp.smoke_mod < sp.smoke_height)) {
      sm = sp.wind_smoke;
      sm_set = true;
    }
  }
  if (sp.drag_on) {
    uu = uu * sp.damping;
    vv = vv * sp.damping;
    u_set = true;
  }
  if (sp.src_active) {
    const int dx = i - sp.src_x, dy = j - sp.src_y;
    if (dx * dx + dy * dy < 40 * 40) {
      if (sp.src_smoke != 0.0f) {
        sm = sp.src_smoke;
        sm_set = true;
      }
      uu = uu + sp.src_velocity * (float)dx;
      vv = vv + sp.src_velocity * (float)dy;
      u_set = true;
    }
  }
  vv = vv + sp.gdt;
  stf(v, k, vv);
  if (u_set) stf(u, k, uu);
  if (sm_set) stf(smoke, k, sm);
  if (sp.pressure_on) p[k] = 0.0f;
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

// ---- P6: apply_extrapolation (src/fluid.cu:720-747) ------------------------
// Threads 0..W-1 do the row phase for column i, threads W..W+H-1 the column
// phase for row j.  Reads and writes of the two phases are disjoint once the
// row phase skips column 1 (whose row-0 and row-(H-1) u values the column
// phase sets to 0) and the column phase uses 0 for v(1,1) and v(W-2,1) (which
// the row phase zeroes), so one launch gives the two-phase result.
template <typename T>
__global__ void k_extrapolate(StepParams sp, T* u, T* v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int W = sp.W, H = sp.H, P = sp.P;
  if (t < W) {
    const int i = t;
    if (i != 1) {
      stf(u, i, ldf(u, (long)P + i));                                // u(i,0) = u(i,1)
      stf(u, (long)(H - 1) * P + i, ldf(u, (long)(H - 2) * P + i));  // u(i,H-1) = u(i,H-2)
    }
    stf(v, (long)P + i, 0.0f);  // v(i,1) = 0
  } else if (t < W + H) {
    const int j = t - W;
    const long r = (long)j * P;
    stf(v, r, j == 1 ? 0.0f : ldf(v, r + 1));                   // v(0,j) = v(1,j)
    stf(u, r + 1, 0.0f);                                        // u(1,j) = 0
    stf(v, r + W - 1, j == 1 ? 0.0f : ldf(v, r + W - 2));       // v(W-1,j) = v(W-2,j)
  }
}

// ---- P7: apply_velocity_advection_at (src/fluid.cu:598-612) ---------------
template <typename T>
__global__ void k_advect_velocity(StepParams sp, GridView g, const T* u, const T* v, T* u_out,
                                  T* v_out) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= sp.W || j >= sp.H) return;
  const long k = g.at(i, j);
  const int cs = sp.cs;
  const float dt = sp.d_t;
  {  // u face at (i*cs, (j+0.5)*cs); get_vertical_edge_velocity (:364-389)
    const float px = (float)(i * cs);
    const float py = (float)(((double)j + 0.5) * (double)cs);
    const float vel_x = ldf(u, k);
    float avg_v = ldf(v, k);
    int count = 1;
    if (g.valid(i - 1, j + 1)) { avg_v = avg_v + ldf(v, g.at(i - 1, j + 1)); count++; }
    if (g.valid(i, j + 1)) { avg_v = avg_v + ldf(v, g.at(i, j + 1)); count++; }
    if (g.valid(i - 1, j)) { avg_v = avg_v + ldf(v, g.at(i - 1, j)); count++; }
    avg_v = avg_v / (float)count;
    const float x = px - vel_x * dt;
    const float y = py - avg_v * dt;
    stf(u_out, k, sample_u(g, u, x, y));
  }
  {  // v face at ((i+0.5)*cs, j*cs); get_horizontal_edge_velocity (:391-416)
    const float px = (float)(((double)i + 0.5) * (double)cs);
    const float py = (float)(j * cs);
    const float vel_y = ldf(v, k);
    float avg_u = ldf(u, k);
    int count = 1;
    if (g.valid(i + 1, j)) { avg_u = avg_u + ldf(u, g.at(i + 1, j)); count++; }
    if (g.valid(i, j - 1)) { avg_u = avg_u + ldf(u, g.at(i, j - 1)); count++; }
    if (g.valid(i + 1, j - 1)) { avg_u = avg_u + ldf(u, g.at(i + 1, j - 1)); count++; }
    avg_u = avg_u / (float)count;
    const float x = px - avg_u * dt;
    const float y = py - vel_y * dt;
    stf(v_out, k, sample_v(g, v, x, y));
  }
}

// ---- P8: apply_smoke_advection_at + decay_smoke_at (src/fluid.cu:560-567,
// 758-762), fused: the decay is applied to the advected value before it is
// stored, which is the reference's copy-back-then-decay for FP32 storage.
template <typename T>
__global__ void k_advect_smoke(StepParams sp, GridView g, const T* u, const T* v, const T* smoke,
                               T* smoke_out) {
  const int i = blockIdx.x * kBX + threadIdx.x;
  const int j = blockIdx.y * kBY + threadIdx.y;
  if (i >= sp.W || j >= sp.H) return;
  const int cs = sp.cs;
  const float cx = centre(i, cs), cy = centre(j, cs);
  const float vx = sample_u(g, u, cx, cy);
  const float vy = sample_v(g, v, cx, cy);
  const float x = cx - vx * sp.d_t;
  const float y = cy - vy * sp.d_t;
  float s = sample_smoke(g, smoke, x, y);
  if (sp.decay_on) s = (float)fmax((double)(s - sp.decay_rd), 0.0);
  stf(smoke_out, g.at(i, j), s);
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

template <typename T>
void step_t(const StepParams& sp, const FieldPtrs& f, int* uv, int* sm, const ProjTuning& tuning,
            hipStream_t s) {
  const GridView g{sp.W, sp.H, sp.P, sp.cs, f.flags};
  T* u = (T*)f.u[*uv];
  T* v = (T*)f.v[*uv];
  T* smk = (T*)f.smoke[*sm];
  const dim3 grid = cell_grid(sp.W, sp.H), block(kBX, kBY);
  hipLaunchKernelGGL(k_forces<T>, grid, block, 0, s, sp, u, v, smk, f.p);
  int cur = *uv;
  enqueue_projection<T>(sp, g, f, &cur, tuning, s);
  u = (T*)f.u[cur];
  v = (T*)f.v[cur];
  if (sp.pressure_on) {
    hipLaunchKernelGGL(k_minmax_partial, dim3(kReduceBlocks), dim3(256), 0, s, sp, f.p, f.scratch);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, s, f.scratch, kReduceBlocks,
                       f.prange);
  }
  hipLaunchKernelGGL(k_extrapolate<T>, dim3((sp.W + sp.H + 255) / 256), dim3(256), 0, s, sp, u, v);
  const int nxt = 1 - cur;
  T* u2 = (T*)f.u[nxt];
  T* v2 = (T*)f.v[nxt];
  hipLaunchKernelGGL(k_advect_velocity<T>, grid, block, 0, s, sp, g, u, v, u2, v2);
  *uv = nxt;
  if (sp.smoke_on) {
    T* s2 = (T*)f.smoke[1 - *sm];
    hipLaunchKernelGGL(k_advect_smoke<T>, grid, block, 0, s, sp, g, u2, v2, smk, s2);
    *sm = 1 - *sm;
  }
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

void launch_build_flags(const MaskParams& m, uint8_t* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_build_flags, cell_grid(m.P, m.H), dim3(kBX, kBY), 0, s, m, flags);
}

void enqueue_step(int storage, const StepParams& sp, const FieldPtrs& f, int* uv, int* sm,
                  const ProjTuning& tuning, hipStream_t s) {
  switch (storage) {
    case OSY_STORAGE_F16: step_t<__half>(sp, f, uv, sm, tuning, s); break;
    case OSY_STORAGE_BF16: step_t<__hip_bfloat16>(sp, f, uv, sm, tuning, s); break;
    default: step_t<float>(sp, f, uv, sm, tuning, s); break;
  }
}

void enqueue_projection_only(int storage, const StepParams& sp, const FieldPtrs& f, int* uv,
                             const ProjTuning& tuning, hipStream_t s) {
  const GridView g{sp.W, sp.H, sp.P, sp.cs, f.flags};
  switch (storage) {
    case OSY_STORAGE_F16: enqueue_projection<__half>(sp, g, f, uv, tuning, s); break;
    case OSY_STORAGE_BF16:
      enqueue_projection<__hip_bfloat16>(sp, g, f, uv, tuning, s);
      break;
    default: enqueue_projection<float>(sp, g, f, uv, tuning, s); break;
  }
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
