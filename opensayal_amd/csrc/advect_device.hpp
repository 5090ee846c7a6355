// advect_device.hpp - per-tile bodies of the forces and the two advections,
// shared by the per-grid kernels (kernels_step.hip) and the batched step
// kernel (kernels_batch.hip).
//
// A tile is 64 columns by TY rows, run by 64 * NY threads (thread t owns
// column t % 64 and rows t / 64 + NY * m).  The advections first stage the
// tile plus a kAdvHalo-cell halo of every field they sample in LDS with
// coalesced loads; a back-trace that leaves the staged region reads HBM
// directly.
#pragma once

#include "fluid_device.hpp"
#include "kernels.hpp"
#include "step_device.hpp"

#pragma clang fp contract(off)

namespace osy {

// The 4-cell halo keeps back-traces of up to about 3 cells inside the
// staged region: a wind-tunnel flow under gravity reaches |v| * d_t of 3-4
// cells, and a wavefront with one lane outside takes the HBM path.
constexpr int kAdvTX = 64, kAdvHalo = 4, kAdvRX = kAdvTX + 2 * kAdvHalo;

template <int TY>
struct AdvShared {
  static constexpr int RY = TY + 2 * kAdvHalo, N = RY * kAdvRX;
  float su[N], sv[N], ss[N];
  uint8_t sf[N];
};

template <int TY>
struct AdvRegion {
  int x0, y0;  // grid coordinates of staged cell (0, 0)
  __device__ bool in(int i, int j) const {
    return (unsigned)(i - x0) < (unsigned)kAdvRX &&
           (unsigned)(j - y0) < (unsigned)AdvShared<TY>::RY;
  }
  __device__ int at(int i, int j) const { return (j - y0) * kAdvRX + (i - x0); }
  // every cell within one of (i, j) is staged
  __device__ bool inner(int i, int j) const {
    return (unsigned)(i - x0 - 1) < (unsigned)(kAdvRX - 2) &&
           (unsigned)(j - y0 - 1) < (unsigned)(AdvShared<TY>::RY - 2);
  }
};

// Stages every cell of the region (rr < RY, cc < kAdvRX): load(rr, cc)
// reads the cell's raw values from HBM (from in-bounds addresses whatever
// (rr, cc), with no branch), put(rr, cc, vals) writes them to LDS.
// Wavefront w takes rows w, w + NY, ... with lane = column for the first 64
// columns (the row is wave-uniform, so row-only work runs on the scalar
// unit); then all threads share the remaining kAdvRX - 64 columns.  The
// loops are unrolled and every load of a thread is issued before its first
// LDS write, so a wavefront waits for HBM once rather than once per row.
// Slots past the region load a duplicate of its last cell and put nothing.
template <int TY, int NY, class Load, class Put>
__device__ __attribute__((always_inline)) inline void stage_region(int tid, const Load& load,
                                                                   const Put& put) {
  constexpr int RY = AdvShared<TY>::RY, EX = kAdvRX - 64;
  constexpr int NM = (RY + NY - 1) / NY, NE = (RY * EX + 64 * NY - 1) / (64 * NY);
  constexpr bool full_m = RY % NY == 0, full_e = (RY * EX) % (64 * NY) == 0;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  using V = decltype(load(0, 0));
  V a[NM], b[NE];
#pragma unroll
  for (int m = 0; m < NM; m++) a[m] = load(full_m ? wv + NY * m : min(wv + NY * m, RY - 1), lane);
#pragma unroll
  for (int e = 0; e < NE; e++) {
    const int q = full_e ? tid + 64 * NY * e : min(tid + 64 * NY * e, RY * EX - 1);
    b[e] = load(q / EX, 64 + q % EX);
  }
  __builtin_amdgcn_sched_barrier(0);  // keep every load ahead of the first write
  // the shared columns first: the compiler moves loads whose values only a
  // guarded write uses into the guarded block, which then comes right after
  // the row loads
#pragma unroll
  for (int e = 0; e < NE; e++) {
    const int q = tid + 64 * NY * e;
    if (full_e || q < RY * EX) put(q / EX, 64 + q % EX, b[e]);
  }
#pragma unroll
  for (int m = 0; m < NM; m++)
    if (full_m || wv + NY * m < RY) put(wv + NY * m, lane, a[m]);
}

// (i, j) clamped into the grid: the address a staged cell outside the grid
// loads from (its value is then replaced by 0).
__device__ inline long clamped_at(const GridView& g, int i, int j) {
  return g.at(min(max(i, 0), g.W - 1), min(max(j, 0), g.H - 1));
}

// Cell of a sample position (the index computation the samplers start with).
template <bool CS1>
__device__ inline int sample_cell(float x, int cs) {
  return to_index(div_cs<CS1>(x, cs));
}

// ---- P1: apply_external_forces_at (src/fluid.cu:308-349), zero_pressure,
// over the tile at (bx, by).  Only the parts no later phase applies while
// loading (do_uv: velocity part and pressure zeroing; do_smoke: inflow
// smoke).
template <typename T, int TY, int NY>
__device__ __attribute__((always_inline)) inline void forces_tile(const StepParams& sp, int do_uv, int do_smoke, T* u, T* v,
                                   T* smoke, float* p, int bx, int by) {
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int i = bx * kAdvTX + (tid & 63);
  if (i >= sp.W) return;
  // every load first (rows clamped into the grid), so a wavefront waits for
  // HBM once rather than once per row
  constexpr int M = TY / NY;
  float uu[M], vv[M], s0[M];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const long k = (long)min(by * TY + (tid >> 6) + NY * m, sp.H - 1) * sp.P + i;
    uu[m] = do_uv ? ldf(u, k) : 0.0f;
    vv[m] = do_uv ? ldf(v, k) : 0.0f;
    s0[m] = do_smoke ? ldf(smoke, k) : 0.0f;
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int j = by * TY + (tid >> 6) + NY * m;
    if (j >= sp.H) continue;
    const long k = (long)j * sp.P + i;
    if (do_uv) {
      force_uv(sp, i, j, uu[m], vv[m]);
      stf(u, k, uu[m]);
      stf(v, k, vv[m]);
      if (sp.pressure_on) p[k] = 0.0f;
    }
    if (do_smoke) {
      const float s1 = force_smoke(sp, i, j, s0[m]);
      if (__float_as_uint(s1) != __float_as_uint(s0[m])) stf(smoke, k, s1);
    }
  }
}

// ---- P6 + P7: apply_extrapolation then apply_velocity_advection_at
// (src/fluid.cu:720-747, 598-612) over the tile at (bx, by).  The staged u, v
// are the extrapolated values (u_extrapolated / v_extrapolated), which are
// also what the HBM fallback reads.  Ends with the LDS free for reuse.
template <typename T, bool CS1, int TY, int NY>
__device__ __attribute__((always_inline)) inline void advect_velocity_tile(const StepParams& sp, const GridView& g,
                                            const T* __restrict__ u, const T* __restrict__ v,
                                            T* __restrict__ u_out, T* __restrict__ v_out, int bx,
                                            int by, AdvShared<TY>& sh) {
  using S = AdvShared<TY>;
  float* su = sh.su;
  float* sv = sh.sv;
  uint8_t* sf = sh.sf;
  const AdvRegion<TY> R{bx * kAdvTX - kAdvHalo, by * TY - kAdvHalo};
  // the thread index is re-read per tile (not hoisted into a register that
  // stays live across a persistent kernel's loop)
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  // u_extrapolated / v_extrapolated with the reads made unconditional: the
  // remapped cell of (i, j) clamped into the grid
  struct Raw {
    float u, v;
    uint8_t f;
  };
  stage_region<TY, NY>(
      tid,
      [&](int rr, int cc) {
        const int i = min(max(R.x0 + cc, 0), g.W - 1), j = min(max(R.y0 + rr, 0), g.H - 1);
        const int ju = j == 0 ? 1 : (j == g.H - 1 ? g.H - 2 : j);
        const int iv = i == 0 ? 1 : (i == g.W - 1 ? g.W - 2 : i);
        return Raw{ldf(u, g.at(i, ju)), ldf(v, g.at(iv, j)), g.flags[g.at(i, j)]};
      },
      [&](int rr, int cc, const Raw& r) {
        const int q = rr * kAdvRX + cc;
        const int i = R.x0 + cc, j = R.y0 + rr;
        const bool in = g.in(i, j);
        su[q] = in && i != 1 ? r.u : 0.0f;
        sv[q] = in && j != 1 ? r.v : 0.0f;
        sf[q] = in ? r.f : (uint8_t)0;
      });
  __syncthreads();
  auto valid = [&](int i, int j) -> bool {
    return R.in(i, j) ? (sf[R.at(i, j)] & CF_FLUID) != 0 : g.valid(i, j);
  };
  auto uval = [&](int i, int j) -> float {
    return R.in(i, j) ? su[R.at(i, j)] : u_extrapolated(u, g, i, j);
  };
  auto vval = [&](int i, int j) -> float {
    return R.in(i, j) ? sv[R.at(i, j)] : v_extrapolated(v, g, i, j);
  };
  auto fluid = [&](int l) -> bool { return (sf[l] & CF_FLUID) != 0; };
  const int cs = sp.cs;
  const float dt = sp.d_t;
  const int i = bx * kAdvTX + (tid & 63);
#pragma unroll
  for (int m = 0; m < TY / NY; m++) {
    const int j = by * TY + (tid >> 6) + NY * m;
    if (i >= sp.W || j >= sp.H) continue;
    const int l = R.at(i, j);
    // the edge velocities read the 3x3 neighbourhood of a tile cell, which
    // is staged; the same terms in the same order as
    // get_vertical_edge_velocity (:364-389) and
    // get_horizontal_edge_velocity (:391-416), skipped terms adding -0
    // u face at (i*cs, (j+0.5)*cs)
    const float pux = (float)(CS1 ? i : i * cs);
    const float puy = CS1 ? (float)j + 0.5f : (float)(((double)j + 0.5) * (double)cs);
    float avg_v = sv[l];
    {  // (i-1, j+1), (i, j+1), (i-1, j)
      const int la = l - 1 + kAdvRX, lb = l + kAdvRX, lc = l - 1;
      const bool a = fluid(la), b = fluid(lb), c = fluid(lc);
      avg_v = add_if(avg_v, a, sv[la]);
      avg_v = add_if(avg_v, b, sv[lb]);
      avg_v = add_if(avg_v, c, sv[lc]);
      avg_v = div_count_bf(avg_v, 1 + (int)a + (int)b + (int)c);
    }
    const float xu = pux - su[l] * dt;
    const float yu = puy - avg_v * dt;
    // v face at ((i+0.5)*cs, j*cs)
    const float pvx = CS1 ? (float)i + 0.5f : (float)(((double)i + 0.5) * (double)cs);
    const float pvy = (float)(CS1 ? j : j * cs);
    float avg_u = su[l];
    {  // (i+1, j), (i, j-1), (i+1, j-1)
      const int la = l + 1, lb = l - kAdvRX, lc = l + 1 - kAdvRX;
      const bool a = fluid(la), b = fluid(lb), c = fluid(lc);
      avg_u = add_if(avg_u, a, su[la]);
      avg_u = add_if(avg_u, b, su[lb]);
      avg_u = add_if(avg_u, c, su[lc]);
      avg_u = div_count_bf(avg_u, 1 + (int)a + (int)b + (int)c);
    }
    const float xv = pvx - avg_u * dt;
    const float yv = pvy - sv[l] * dt;
    float nu, nv;
    const bool staged = R.inner(sample_cell<CS1>(xu, cs), sample_cell<CS1>(yu, cs)) &&
                        R.inner(sample_cell<CS1>(xv, cs), sample_cell<CS1>(yv, cs));
    if (__all(staged)) {
      nu = lds_sample_u<CS1, kAdvRX>(cs, xu, yu, su, sf, R.x0, R.y0);
      nv = lds_sample_v<CS1, kAdvRX>(cs, xv, yv, sv, sf, R.x0, R.y0);
    } else {
      nu = sample_u<CS1>(cs, xu, yu, valid, uval);
      nv = sample_v<CS1>(cs, xv, yv, valid, vval);
    }
    stf(u_out, g.at(i, j), nu);
    stf(v_out, g.at(i, j), nv);
  }
  __syncthreads();
}

// ---- P8: apply_smoke_advection_at + decay_smoke_at (src/fluid.cu:560-567,
// 758-762) over the tile at (bx, by).  The staged smoke is the smoke after
// the forces' inflow band and source (force_smoke); the decay is applied to
// the advected value before it is stored, which is the reference's
// copy-back-then-decay.  Ends with the LDS free for reuse.
template <typename T, bool CS1, int TY, int NY>
__device__ __attribute__((always_inline)) inline void advect_smoke_tile(const StepParams& sp, const GridView& g,
                                         const T* __restrict__ u, const T* __restrict__ v,
                                         const T* __restrict__ smoke, T* __restrict__ smoke_out,
                                         int bx, int by, AdvShared<TY>& sh) {
  using S = AdvShared<TY>;
  float* su = sh.su;
  float* sv = sh.sv;
  float* ss = sh.ss;
  uint8_t* sf = sh.sf;
  const AdvRegion<TY> R{bx * kAdvTX - kAdvHalo, by * TY - kAdvHalo};
  // the thread index is re-read per tile (not hoisted into a register that
  // stays live across a persistent kernel's loop)
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  struct Raw {
    float u, v, s;
    uint8_t f;
  };
  stage_region<TY, NY>(
      tid,
      [&](int rr, int cc) {
        const long k = clamped_at(g, R.x0 + cc, R.y0 + rr);
        return Raw{ldf(u, k), ldf(v, k), ldf(smoke, k), g.flags[k]};
      },
      [&](int rr, int cc, const Raw& r) {
        const int q = rr * kAdvRX + cc;
        const int i = R.x0 + cc, j = R.y0 + rr;
        const bool in = g.in(i, j);
        su[q] = in ? r.u : 0.0f;
        sv[q] = in ? r.v : 0.0f;
        ss[q] = in ? rnd<T>(force_smoke(sp, i, j, r.s)) : 0.0f;
        sf[q] = in ? r.f : (uint8_t)0;
      });
  __syncthreads();
  auto valid = [&](int i, int j) -> bool {
    return R.in(i, j) ? (sf[R.at(i, j)] & CF_FLUID) != 0 : g.valid(i, j);
  };
  auto uval = [&](int i, int j) -> float { return R.in(i, j) ? su[R.at(i, j)] : ldf(u, g.at(i, j)); };
  auto vval = [&](int i, int j) -> float { return R.in(i, j) ? sv[R.at(i, j)] : ldf(v, g.at(i, j)); };
  auto sval = [&](int i, int j) -> float {
    return R.in(i, j) ? ss[R.at(i, j)] : rnd<T>(force_smoke(sp, i, j, ldf(smoke, g.at(i, j))));
  };

  const int cs = sp.cs;
  const int i = bx * kAdvTX + (tid & 63);
#pragma unroll
  for (int m = 0; m < TY / NY; m++) {
    const int j = by * TY + (tid >> 6) + NY * m;
    if (i >= sp.W || j >= sp.H) continue;
    const float cx = centre<CS1>(i, cs), cy = centre<CS1>(j, cs);
    // the velocity at a cell centre samples cells within one of (i, j),
    // which are staged
    float vx, vy;
    if constexpr (CS1) {
      vx = lds_centre_u<kAdvRX>(su, sf, R.at(i, j));
      vy = lds_centre_v<kAdvRX>(sv, sf, R.at(i, j));
    } else {
      vx = lds_sample_u<CS1, kAdvRX>(cs, cx, cy, su, sf, R.x0, R.y0);
      vy = lds_sample_v<CS1, kAdvRX>(cs, cx, cy, sv, sf, R.x0, R.y0);
    }
    const float x = cx - vx * sp.d_t;
    const float y = cy - vy * sp.d_t;
    float s;
    if (__all(R.inner(sample_cell<CS1>(x, cs), sample_cell<CS1>(y, cs))))
      s = lds_sample_smoke<CS1, kAdvRX>(cs, x, y, ss, sf, R.x0, R.y0);
    else
      s = sample_smoke<CS1>(cs, x, y, valid, sval);
    s = rnd<T>(s);  // the advection's store, before the decay reads it back
    if (sp.decay_on) s = (float)fmax((double)(s - sp.decay_rd), 0.0);
    stf(smoke_out, g.at(i, j), s);
  }
  __syncthreads();
}

}  // namespace osy
