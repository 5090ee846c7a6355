// projection_device.hpp - device code of the tiled projections shared by the
// per-grid kernels (kernels_project.hip) and the batched step kernel
// (kernels_batch.hip).
#pragma once

#include <type_traits>

#include "fluid_device.hpp"
#include "kernels.hpp"
#include "step_device.hpp"

#pragma clang fp contract(off)

namespace osy {

// Per-cell gate byte of the tiled kernel, built in LDS from the flags.  The
// four face gates are set only for projected cells, so a cell that is not
// projected changes nothing when it runs the update.
enum : uint32_t {
  G_L = 1u << 0,     // u(i,j) moves   (left neighbour fluid)
  G_R = 1u << 1,     // u(i+1,j) moves (right neighbour fluid)
  G_B = 1u << 2,     // v(i,j) moves   (bottom neighbour fluid)
  G_T = 1u << 3,     // v(i,j+1) moves (top neighbour fluid)
  G_UPD = 1u << 4,   // cell is projected
  G_L2_SHIFT = 5,    // bits 5..6: log2(total_s) when total_s is 1, 2 or 4
  G_SLOW = 1u << 7,  // projected with total_s 0 or 3: divide
};

// d = o * (div / s) for the cell with gate byte g.  For s = 1, 2, 4 the
// quotient is div * 2^-log2(s), exact; for s = 0 or 3 (cells against a
// wall or the obstacle) the lanes that need it divide.  Exhaustively checked
// against IEEE division over all binary32 inputs (tests/test_quotient.py).
__device__ inline float relaxed_quotient(float div, uint32_t g, float o) {
  const float y = __uint_as_float(0x3F800000u - (((g >> G_L2_SHIFT) & 3u) << 23));
  float q = div * y;
  if (g & G_SLOW) q = div / (float)__builtin_popcount(g & 15u);
  return o * q;
}

// 2^-log2(s) of a cell with gate byte g (for s = 1, 2, 4).
__device__ inline float gate_scale(uint32_t g) {
  return __uint_as_float(0x3F800000u - (((g >> G_L2_SHIFT) & 3u) << 23));
}

// The same quotient for the register-resident kernel, from the cell's
// gate_scale y: `slow` (wavefront uniform) says whether any lane of this
// row may need the division; only such rows divide, every lane, and select.
__device__ inline float row_quotient(float div, float y, uint32_t g, float o, bool slow) {
  float q = div * y;
  if (slow) {
    const float qs = div / (float)__builtin_popcount(g & 15u);
    q = (g & G_SLOW) ? qs : q;
  }
  return o * q;
}

// x + gate(d): d when the gate bit is set, else -0 (x + -0 == x for every x,
// signed zeros and NaN included), so the face keeps its bits when closed.
__device__ inline float gated_add(float x, float d, uint32_t g, int bit) {
  const uint32_t m = (uint32_t)((int32_t)(g << (31 - bit)) >> 31);
  return x + __uint_as_float((__float_as_uint(d) & m) | (~m & 0x80000000u));
}
// x - gate(d): d when set, else +0 (x - +0 == x for every x).
__device__ inline float gated_sub(float x, float d, uint32_t g, int bit) {
  const uint32_t m = (uint32_t)((int32_t)(g << (31 - bit)) >> 31);
  return x - __uint_as_float(__float_as_uint(d) & m);
}

// Two horizontally adjacent elements (cells 2q, 2q+1 of a row).
template <typename T> struct PairOf;
template <> struct PairOf<float> { using type = float2; };
template <> struct PairOf<__half> { using type = __half2; };
template <> struct PairOf<__hip_bfloat16> { using type = __hip_bfloat162; };

template <typename T>
__device__ inline void load_pair(const T* __restrict__ a, long k, bool in0, bool in1, float& x0,
                                 float& x1) {
  if (in0 && in1) {  // k is even and rows start on 64-element boundaries
    const typename PairOf<T>::type pr = *reinterpret_cast<const typename PairOf<T>::type*>(a + k);
    x0 = Store<T>::ld(pr.x);
    x1 = Store<T>::ld(pr.y);
  } else {
    x0 = in0 ? ldf(a, k) : 0.0f;
    x1 = in1 ? ldf(a, k + 1) : 0.0f;
  }
}

template <typename T>
__device__ inline void store_pair(T* __restrict__ a, long k, bool in0, bool in1, float x0,
                                  float x1) {
  if (in0 && in1) {
    typename PairOf<T>::type pr;
    pr.x = Store<T>::st(x0);
    pr.y = Store<T>::st(x1);
    *reinterpret_cast<typename PairOf<T>::type*>(a + k) = pr;
  } else {
    if (in0) stf(a, k, x0);
    if (in1) stf(a, k + 1, x1);
  }
}

// Accesses of the resident projection's halo exchange, coherent across the
// GPU without cache maintenance: relaxed agent-scope atomics, which the
// compiler issues as loads and stores that bypass the caches that are not
// coherent between CUs and XCDs (a CU's L1, an XCD's L2).  The bits of a
// pair travel as one 2-element word.
template <typename T>
__device__ inline void coherent_load_pair(const T* a, long k, float& x0, float& x1) {
  using W = typename std::conditional<sizeof(T) == 4, unsigned long long, unsigned>::type;
  const W w = __hip_atomic_load(reinterpret_cast<const W*>(a + k), __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  typename PairOf<T>::type pr;
  __builtin_memcpy(&pr, &w, sizeof(W));
  x0 = Store<T>::ld(pr.x);
  x1 = Store<T>::ld(pr.y);
}

template <typename T>
__device__ inline void coherent_store_one(T* a, long k, float x) {
  using W = typename std::conditional<sizeof(T) == 4, unsigned, unsigned short>::type;
  const T e = Store<T>::st(x);
  W w;
  __builtin_memcpy(&w, &e, sizeof(W));
  __hip_atomic_store(reinterpret_cast<W*>(a + k), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ inline void coherent_store_pair(T* a, long k, bool in0, bool in1, float x0,
                                           float x1) {
  if (in0 && in1) {  // k is even and rows start on 64-element boundaries
    using W = typename std::conditional<sizeof(T) == 4, unsigned long long, unsigned>::type;
    typename PairOf<T>::type pr;
    pr.x = Store<T>::st(x0);
    pr.y = Store<T>::st(x1);
    W w;
    __builtin_memcpy(&w, &pr, sizeof(W));
    __hip_atomic_store(reinterpret_cast<W*>(a + k), w, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  } else {
    if (in0) coherent_store_one(a, k, x0);
    if (in1) coherent_store_one(a, k + 1, x1);
  }
}

// ---- register-resident kernel ------------------------------------------------
//
// Eight wavefronts stacked vertically; wavefront w holds region rows
// [w*R, w*R + R), lane l holds cells 2l and 2l+1 of each of them: u, v and
// the gate byte of both cells, all in VGPRs for the whole launch.  The
// region is 128 x 8R cells and the owned tile (128 - 4k) x (8R - 4k).
//
// In a colour phase each row has one active cell per lane (parity b).  For
// b = 0 every face the cell touches is in the lane's own registers.  For
// b = 1 the right face u(2l+2) belongs to lane l+1: it is read and updated
// through DPP lane shifts.  The top face of a wavefront's last row belongs
// to the wavefront above: its value is published through LDS before the
// phase and the update sent back after it (two barriers per phase).

__device__ inline float lane_from_right(float x) {  // lane l gets lane l+1's x (0 for lane 63)
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, true));
}
__device__ inline float lane_from_left(float x) {  // lane l gets lane l-1's x (0 for lane 0)
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}

// gate of d: d when the bit is set, else -0 (which an addition leaves out
// bit for bit)
__device__ inline float gate_neg(float d, uint32_t g, int bit) {
  const uint32_t m = (uint32_t)((int32_t)(g << (31 - bit)) >> 31);
  return __uint_as_float((__float_as_uint(d) & m) | (~m & 0x80000000u));
}

// gate of d: d when the bit is set, else +0
__device__ inline float gate_pos(float d, uint32_t g, int bit) {
  const uint32_t m = (uint32_t)((int32_t)(g << (31 - bit)) >> 31);
  return __uint_as_float(__float_as_uint(d) & m);
}

// Gate byte of a cell from its flag f, the right neighbour's flag fr and
// the top neighbour's flag ft (same encoding as the LDS-tiled kernel).
__device__ inline uint32_t gate_byte(uint32_t f, uint32_t fr, uint32_t ft) {
  if (!(f & CF_UPD)) return 0;
  uint32_t g = G_UPD;
  if (f & CF_FU) g |= G_L;
  if (fr & CF_FU) g |= G_R;
  if (f & CF_FV) g |= G_B;
  if (ft & CF_FV) g |= G_T;
  const uint32_t s = (f >> CF_S_SHIFT) & 7u;
  if (s == 0 || s == 3)
    g |= G_SLOW;
  else
    g |= (s >> 1) << G_L2_SHIFT;
  return g;
}

struct RegArgs {
  int tw, th, k;  // owned tile and sweeps per launch (halo 2k)
  int ntx;        // tiles per grid row
  int sweeps;     // sweeps this launch (<= k)
  int forces;     // apply the velocity forces on load, pressure sums from 0
};

__device__ inline bool lc_in_tile(int lc, int K, int tw) { return lc >= 2 * K && lc < 2 * K + tw; }

// The divisor of update_pressure_at: d_t, and recip = 1.0 / (double)d_t
// (the double nearest 1/d_t) when 0 < |d_t| < 2, else 0.
struct PressureDivisor {
  float d_t;
  double recip;
};

// update_pressure_at (src/fluid.cu:221-227): ((d * density) * cell_size) / d_t.
// With recip set the quotient is (float)(x * recip) in double, which is the
// IEEE float quotient x / d_t for every binary32 x: the double is within
// 2^-52 of x / d_t; a quotient of two binary32 numbers lies at least 2^-49
// of itself from every binary32 rounding midpoint in the normal range, and
// in the subnormal range it can only equal a midpoint if the divisor's
// lowest set bit weighs 2 or more, which |d_t| < 2 excludes.  (Exhaustive
// device check for d_t = 0.05 and 1/60: osy_arith_check, modes 5 and 6.)
__device__ inline float pressure_term(float d, float density, float csf,
                                      const PressureDivisor& pd) {
  const float x = (d * density) * csf;
  return pd.recip != 0.0 ? (float)((double)x * pd.recip) : x / pd.d_t;
}

// Register state of one lane of the register-resident kernel.
template <int R>
struct RegTile {
  float u0[R], u1[R], v0[R], v1[R];  // faces of cells 2l and 2l+1, rows 0..R-1
  float s0[R], s1[R];                // pressure sums (pressure runs only)
  uint32_t gw[R];                    // gate bytes: cell 2l in bits 0-7, 2l+1 in 8-15
  uint32_t m0[R], m1[R];             // face masks of cells 2l and 2l+1 (reg_face_masks)
  uint32_t slow;                     // wavefront-uniform: bit 2 rho + c set when some
                                     // lane's cell c of row rho needs the general
                                     // quotient (row_quotient, mask_quotient)
};

// Whether a tile keeps each cell's four face gates as byte masks (one VGPR
// per cell, byte k 0xFF when face bit k is set, else 0) rather than as bits
// of the gate byte: a gated face update is then one AND that selects and
// sign-extends the mask byte in its operand, and one add.  Pressure runs
// keep the gate byte, whose projected bit gates the sum.
template <int R, bool PRESSURE> constexpr bool reg_face_masks() { return !PRESSURE; }

// The face masks of gate byte g: bit k of g (k = 0..3) spread to bit 8k by
// one multiply (the 16 partial products land on distinct bits), then to
// the whole byte.
__device__ inline uint32_t face_masks(uint32_t g) {
  return (((g & 15u) * 0x00204081u) & 0x01010101u) * 0xFFu;
}

// d where face K of the masks m is open, else +0.
template <int K>
__device__ inline float mask_face(float d, uint32_t m) {
  const uint32_t mk = (uint32_t)(int32_t)(int8_t)(uint8_t)(m >> (8 * K));
  return __uint_as_float(__float_as_uint(d) & mk);
}

// div / s for a cell with face masks m (popcount(m) = 8s).  A cell with
// s = 0 moves no face, so its quotient is never used, and almost every
// other cell has s = 4: a row whose lanes all have s = 0 or 4 takes
// div * 0.25.  The rest (`mixed`, wavefront uniform) take div * 2^-log2(s)
// for s = 1, 2, 4 (log2(s) = 8s >> 4), exact, and third_quotient for s = 3
// (the IEEE quotient div / 3 for every binary32 div).
__device__ inline float mask_quotient(float div, uint32_t m, bool mixed) {
  float q = div * 0.25f;
  if (mixed) {
    const uint32_t c8 = (uint32_t)__builtin_popcount(m);
    const float qp = __builtin_amdgcn_ldexpf(div, -(int)(c8 >> 4));
    q = c8 == 24u ? third_quotient(div) : qp;
  }
  return q;
}

// Whether some lane's face masks m need mask_quotient's general branch.
__device__ inline bool mask_mixed(uint32_t m) {
  const uint32_t c8 = (uint32_t)__builtin_popcount(m);
  return __any(c8 != 0u && c8 != 32u);
}

// apply_projection_at for the active cell of row rho (parity B) of every
// lane.  Cells near the region edge compute with stale neighbours; their
// wrong values move one cell per phase and never reach the tile within a
// launch, so no lane or row is masked.
template <typename T, int R, int B, bool PRESSURE>
__device__ __attribute__((always_inline)) inline void reg_row(RegTile<R>& t, const int rho,
                                                              float vt0, float vt1, float& dtop,
                                                              float o, float density, float csf,
                                                              const PressureDivisor& pd) {
  const int up = rho + 1 < R ? rho + 1 : rho;  // row above (when inside the strip)
  if constexpr (reg_face_masks<R, PRESSURE>()) {
    // x + gate(d) as x - mask(-d) and x - gate(d) as x - mask(d): an open
    // face gets x + d (x - (-d) is x + d in IEEE arithmetic), a closed one
    // x - (+0), which is x for every x
    if constexpr (B == 0) {
      const uint32_t m = t.m0[rho];
      const float uu = t.u0[rho], ur = t.u1[rho], vv = t.v0[rho];
      const float vt = (rho + 1 < R) ? t.v0[up] : vt0;
      const float div = ((ur - uu) + vt) - vv;
      const float q = mask_quotient(div, m, (t.slow >> (2 * rho)) & 1u);
      const float d = o * q, nd = -d;
      float nu0 = uu - mask_face<0>(nd, m), nu1 = ur - mask_face<1>(d, m);
      rnd2<T>(nu0, nu1);
      t.u0[rho] = nu0;
      t.u1[rho] = nu1;
      if (rho + 1 < R) {
        float nv = vv - mask_face<2>(nd, m), nt = vt - mask_face<3>(d, m);
        rnd2<T>(nv, nt);
        t.v0[rho] = nv;
        t.v0[up] = nt;
      } else {
        t.v0[rho] = rnd<T>(vv - mask_face<2>(nd, m));
        dtop = mask_face<3>(d, m);
      }
    } else {
      const uint32_t m = t.m1[rho];
      const float uu = t.u1[rho], ur = lane_from_right(t.u0[rho]), vv = t.v1[rho];
      const float vt = (rho + 1 < R) ? t.v1[up] : vt1;
      const float div = ((ur - uu) + vt) - vv;
      const float q = mask_quotient(div, m, (t.slow >> (2 * rho + 1)) & 1u);
      const float d = o * q, nd = -d;
      float nu1 = uu - mask_face<0>(nd, m), nu0 = t.u0[rho] - lane_from_left(mask_face<1>(d, m));
      rnd2<T>(nu1, nu0);
      t.u1[rho] = nu1;
      t.u0[rho] = nu0;
      if (rho + 1 < R) {
        float nv = vv - mask_face<2>(nd, m), nt = vt - mask_face<3>(d, m);
        rnd2<T>(nv, nt);
        t.v1[rho] = nv;
        t.v1[up] = nt;
      } else {
        t.v1[rho] = rnd<T>(vv - mask_face<2>(nd, m));
        dtop = mask_face<3>(d, m);
      }
    }
  } else if constexpr (B == 0) {
    // active cell 2l: faces u0, u1, v0[rho], v0[rho+1] (all in this lane)
    const uint32_t g = t.gw[rho];
    const float uu = t.u0[rho], ur = t.u1[rho], vv = t.v0[rho];
    const float vt = (rho + 1 < R) ? t.v0[up] : vt0;
    const float div = ((ur - uu) + vt) - vv;
    const float d = row_quotient(div, gate_scale(g), g, o, (t.slow >> (2 * rho)) & 1u);
    if constexpr (PRESSURE) {
      // a cell that is not projected adds -0, which leaves its sum's bits
      t.s0[rho] = t.s0[rho] + gate_neg(pressure_term(d, density, csf, pd), g, 4);
    }
    float nu0 = gated_add(uu, d, g, 0), nu1 = gated_sub(ur, d, g, 1);
    rnd2<T>(nu0, nu1);
    t.u0[rho] = nu0;
    t.u1[rho] = nu1;
    if (rho + 1 < R) {
      float nv = gated_add(vv, d, g, 2), nt = gated_sub(vt, d, g, 3);
      rnd2<T>(nv, nt);
      t.v0[rho] = nv;
      t.v0[up] = nt;
    } else {
      t.v0[rho] = rnd<T>(gated_add(vv, d, g, 2));
      dtop = gate_pos(d, g, 3);
    }
  } else {
    // active cell 2l+1: faces u1, u0 of lane l+1, v1[rho], v1[rho+1]
    const uint32_t g = t.gw[rho] >> 8;
    const float uu = t.u1[rho], ur = lane_from_right(t.u0[rho]), vv = t.v1[rho];
    const float vt = (rho + 1 < R) ? t.v1[up] : vt1;
    const float div = ((ur - uu) + vt) - vv;
    const float d = row_quotient(div, gate_scale(g), g, o, (t.slow >> (2 * rho + 1)) & 1u);
    if constexpr (PRESSURE) {
      t.s1[rho] = t.s1[rho] + gate_neg(pressure_term(d, density, csf, pd), g, 4);
    }
    float nu1 = gated_add(uu, d, g, 0), nu0 = t.u0[rho] - lane_from_left(gate_pos(d, g, 1));
    rnd2<T>(nu1, nu0);
    t.u1[rho] = nu1;
    t.u0[rho] = nu0;
    if (rho + 1 < R) {
      float nv = gated_add(vv, d, g, 2), nt = gated_sub(vt, d, g, 3);
      rnd2<T>(nv, nt);
      t.v1[rho] = nv;
      t.v1[up] = nt;
    } else {
      t.v1[rho] = rnd<T>(gated_add(vv, d, g, 2));
      dtop = gate_pos(d, g, 3);
    }
  }
}

// LDS the register-resident tile exchanges rows through, two slots of each
// (with one barrier per phase consecutive phases use different slots).
struct RegShared {
  float2 vrow[2][kRegWaves][64];  // row 0 (v0, v1) of each wavefront
  float dsend[2][kRegWaves][64];  // top-row update for the wavefront above
};

// Rows RHO..R-1 of one phase, unrolled at compile time.
template <typename T, int R, int PB, bool PRESSURE, int RHO>
__device__ __attribute__((always_inline)) inline void reg_rows_from(RegTile<R>& t, float vt0,
                                                                    float vt1, float& dtop,
                                                                    float o, float density,
                                                                    float csf, const PressureDivisor& pd) {
  if constexpr (RHO < R) {
    reg_row<T, R, ((PB + RHO) & 1), PRESSURE>(t, RHO, vt0, vt1, dtop, o, density, csf, pd);
    reg_rows_from<T, R, PB, PRESSURE, RHO + 1>(t, vt0, vt1, dtop, o, density, csf, pd);
  }
}

// One colour phase.  PB is the parity of the active cell in row 0 of this
// wavefront (compile time: phases come in pairs with PB and 1 - PB).  A
// wavefront none of whose rows can still reach the tile skips the phase;
// the others run it as straight-line code over registers.
//
// (vt0, vt1) is row 0 of the wavefront above as it stands when the phase
// starts: the top faces of this wavefront's last row.  The last row moves
// those faces by dtop, which the wavefront above applies to its row 0.
// Two ways to exchange both through LDS:
//  * two barriers (ONE_SYNC false): each wavefront publishes its row 0 and
//    reads (vt0, vt1) after the first; sends dtop and applies the one from
//    below after the second.
//  * one barrier (ONE_SYNC true): each wavefront publishes its row 0 after
//    computing (this phase's own updates, before the one from below) beside
//    dtop; after the barrier it applies the update from below to its row 0,
//    and the wavefront below applies the same update, with the same
//    rounding, to the copy it reads, which is then the row as the wavefront
//    above holds it.  The slots alternate between phases (slot b), so none
//    is rewritten before every reader of its previous contents has passed
//    the next barrier.  The caller publishes and reads the rows as loaded
//    before the first phase.
// The two give the same bits.  The persistent batch kernel measured faster
// with one barrier; the per-grid kernel measured no faster at 8 rows and
// slower at 4 (its extra live registers cost a wavefront per SIMD).
template <typename T, int R, int PB, bool PRESSURE, bool ONE_SYNC>
__device__ __attribute__((always_inline)) inline void reg_phase(RegTile<R>& t, RegShared& sh,
                                                                float& vt0, float& vt1, int lane,
                                                                int w, int rbase, int s,
                                                                int nphase, int K, int th,
                                                                float o, float density,
                                                                float csf, const PressureDivisor& pd) {
  const int e = nphase - 1 - s;
  const int lo = 2 * K - 1 - e;
  const int rhi = 2 * K + th + e;
  const int rlo = lo - rbase, rhi_rel = rhi - rbase;  // box rows of this wavefront
  const int b = ONE_SYNC ? (s + 1) & 1 : 0;
  if constexpr (!ONE_SYNC) {
    if (w > 0) sh.vrow[0][w][lane] = make_float2(t.v0[0], t.v1[0]);
    __syncthreads();
    vt0 = vt1 = 0.0f;  // not carried from the previous phase
    if (w < kRegWaves - 1) {
      const float2 x = sh.vrow[0][w + 1][lane];
      vt0 = x.x;
      vt1 = x.y;
    }
  }
  float dtop = 0.0f;
  if (rlo < R && rhi_rel > 0) {  // wavefront-uniform: some row is in the box
    // the gate words do not change during the launch; hiding that from the
    // compiler keeps it from hoisting the 8 face masks per row out of the
    // phase loop into registers
#pragma unroll
    for (int rho = 0; rho < R; rho++) {
      if constexpr (reg_face_masks<R, PRESSURE>()) {
        asm volatile("" : "+v"(t.m0[rho]));
        asm volatile("" : "+v"(t.m1[rho]));
      } else {
        asm volatile("" : "+v"(t.gw[rho]));
      }
    }
    reg_rows_from<T, R, PB, PRESSURE, 0>(t, vt0, vt1, dtop, o, density, csf, pd);
  }
  if constexpr (ONE_SYNC) {
    if (w > 0) sh.vrow[b][w][lane] = make_float2(t.v0[0], t.v1[0]);
  }
  if (w < kRegWaves - 1) sh.dsend[b][w][lane] = dtop;
  __syncthreads();
  if (w > 0) {
    // the wavefront below moved the top faces of its last row: they are
    // this wavefront's row-0 v of parity 1 - PB
    const float dd = sh.dsend[b][w - 1][lane];
    if constexpr (PB == 0)
      t.v1[0] = rnd<T>(t.v1[0] - dd);
    else
      t.v0[0] = rnd<T>(t.v0[0] - dd);
  }
  if constexpr (ONE_SYNC) {
    if (w < kRegWaves - 1) {
      const float2 x = sh.vrow[b][w + 1][lane];
      if constexpr (PB == 0) {
        vt0 = x.x;
        vt1 = rnd<T>(x.y - dtop);
      } else {
        vt0 = rnd<T>(x.x - dtop);
        vt1 = x.y;
      }
    }
  }
}

// Where a workgroup's tile lies: K (halo 2K), the region origin (gx0 is
// even) and this wavefront's first region row.  Wavefront-uniform.
struct RegPlace {
  int K, gx0, gy0, w, rbase;
};

template <int R>
__device__ inline RegPlace reg_place(const RegArgs& a, int tile) {
  RegPlace pl;
  pl.K = a.k;
  const int tx = tile % a.ntx, ty = tile / a.ntx;
  pl.gx0 = tx * a.tw - 2 * a.k;
  pl.gy0 = ty * a.th - 2 * a.k;
  pl.w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  pl.rbase = pl.w * R;
  return pl;
}

// This lane's index, re-read from the thread index so that it is not kept
// live across the sweeps.
__device__ inline int reg_lane() {
  int tid = (int)threadIdx.x;
  asm volatile("" : "+v"(tid));
  return tid & 63;
}

// Loads the region of u, v (and the pressure sums) and the gate bytes of
// the tile into registers: every load first, from the even column clamped
// into the row (so a pair stays inside its row of P cells) and the row
// clamped into the grid, with no branch, so a wavefront waits for HBM once
// rather than once per row.  Cells outside the grid are then set to 0.
template <typename T, bool PRESSURE, int R>
__device__ __attribute__((always_inline)) inline void reg_load(RegTile<R>& t, const StepParams& sp,
                                                               const uint8_t* __restrict__ gates,
                                                               const T* __restrict__ u_in,
                                                               const T* __restrict__ v_in,
                                                               const float* __restrict__ p,
                                                               const RegArgs& a,
                                                               const RegPlace& pl) {
  const int lane = reg_lane();
  const int W = sp.W, H = sp.H, P = sp.P;
  const int gi = pl.gx0 + 2 * lane;
  const bool col0 = (unsigned)gi < (unsigned)W, col1 = (unsigned)(gi + 1) < (unsigned)W;
  t.slow = 0;
  const int gic = min(max(gi, 0), P - 2);
  uint32_t graw[R];
#pragma unroll
  for (int rho = 0; rho < R; rho++) {
    const long k = (long)min(max(pl.gy0 + pl.rbase + rho, 0), H - 1) * P + gic;
    load_pair(u_in, k, true, true, t.u0[rho], t.u1[rho]);
    load_pair(v_in, k, true, true, t.v0[rho], t.v1[rho]);
    graw[rho] = *reinterpret_cast<const uint16_t*>(gates + k);
    if constexpr (PRESSURE) {
      const float2 pp = *reinterpret_cast<const float2*>(p + k);
      t.s0[rho] = pp.x;
      t.s1[rho] = pp.y;
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int rho = 0; rho < R; rho++) {
    const int gj = pl.gy0 + pl.rbase + rho;
    const bool row_in = (unsigned)gj < (unsigned)H;
    const bool in0 = row_in && col0, in1 = row_in && col1;
    t.u0[rho] = in0 ? t.u0[rho] : 0.0f;
    t.v0[rho] = in0 ? t.v0[rho] : 0.0f;
    t.u1[rho] = in1 ? t.u1[rho] : 0.0f;
    t.v1[rho] = in1 ? t.v1[rho] : 0.0f;
    if (a.forces) {  // the forces' results as stored (rounded to T)
      if (in0) {
        force_uv(sp, gi, gj, t.u0[rho], t.v0[rho]);
        t.u0[rho] = rnd<T>(t.u0[rho]);
        t.v0[rho] = rnd<T>(t.v0[rho]);
      }
      if (in1) {
        force_uv(sp, gi + 1, gj, t.u1[rho], t.v1[rho]);
        t.u1[rho] = rnd<T>(t.u1[rho]);
        t.v1[rho] = rnd<T>(t.v1[rho]);
      }
    }
    if constexpr (PRESSURE) {  // the forces zero the pressure
      t.s0[rho] = (in0 && !a.forces) ? t.s0[rho] : 0.0f;
      t.s1[rho] = (in1 && !a.forces) ? t.s1[rho] : 0.0f;
    } else {
      t.s0[rho] = t.s1[rho] = 0.0f;
    }
    // the gate bytes of cells 2l and 2l+1 (static, built with the flags)
    t.gw[rho] = (in0 ? (graw[rho] & 0xFFu) : 0u) | (in1 ? (graw[rho] & 0xFF00u) : 0u);
    if constexpr (reg_face_masks<R, PRESSURE>()) {
      t.m0[rho] = face_masks(t.gw[rho]);
      t.m1[rho] = face_masks(t.gw[rho] >> 8);
      t.slow |= (mask_mixed(t.m0[rho]) ? 1u : 0u) << (2 * rho);
      t.slow |= (mask_mixed(t.m1[rho]) ? 1u : 0u) << (2 * rho + 1);
    } else {
      t.slow |= (__any((t.gw[rho] & G_SLOW) != 0) ? 1u : 0u) << (2 * rho);
      t.slow |= (__any((t.gw[rho] & (G_SLOW << 8)) != 0) ? 1u : 0u) << (2 * rho + 1);
    }
  }
}

// `sweeps` sweeps over the registers: colour 0 (even cells) first; row 0's
// active parity in a colour-0 phase is (gy0 + rbase) & 1 and flips every
// phase.
template <typename T, bool PRESSURE, int R, bool ONE_SYNC>
__device__ __attribute__((always_inline)) inline void reg_sweeps(RegTile<R>& t, RegShared& sh,
                                                                 const StepParams& sp,
                                                                 const RegPlace& pl, int th,
                                                                 int sweeps) {
  const int lane = reg_lane();
  const int w = pl.w, rbase = pl.rbase, K = pl.K;
  const float o = sp.o;
  const float csf = (float)sp.cs;
  const PressureDivisor pdiv{sp.d_t, sp.pdt_recip};
  const int nphase = 2 * sweeps;
  float vt0 = 0.0f, vt1 = 0.0f;
  if constexpr (ONE_SYNC) {  // row 0 of the wavefront above, as loaded
    if (w > 0) sh.vrow[0][w][lane] = make_float2(t.v0[0], t.v1[0]);
    __syncthreads();
    if (w < kRegWaves - 1) {
      const float2 x = sh.vrow[0][w + 1][lane];
      vt0 = x.x;
      vt1 = x.y;
    }
  }
  if (((pl.gy0 + rbase) & 1) == 0) {
    for (int s = 0; s < nphase; s += 2) {
      reg_phase<T, R, 0, PRESSURE, ONE_SYNC>(t, sh, vt0, vt1, lane, w, rbase, s, nphase, K, th, o,
                                             sp.density, csf, pdiv);
      reg_phase<T, R, 1, PRESSURE, ONE_SYNC>(t, sh, vt0, vt1, lane, w, rbase, s + 1, nphase, K,
                                             th, o, sp.density, csf, pdiv);
    }
  } else {
    for (int s = 0; s < nphase; s += 2) {
      reg_phase<T, R, 1, PRESSURE, ONE_SYNC>(t, sh, vt0, vt1, lane, w, rbase, s, nphase, K, th, o,
                                             sp.density, csf, pdiv);
      reg_phase<T, R, 0, PRESSURE, ONE_SYNC>(t, sh, vt0, vt1, lane, w, rbase, s + 1, nphase, K,
                                             th, o, sp.density, csf, pdiv);
    }
  }
}

// Writes back the owned tile (RING: only its cells within 2K of the tile's
// edge, which are the halos of the neighbouring tiles, and no pressure).
template <typename T, bool PRESSURE, int R, bool RING>
__device__ __attribute__((always_inline)) inline void reg_store(const RegTile<R>& t,
                                                                T* __restrict__ u_out,
                                                                T* __restrict__ v_out,
                                                                float* __restrict__ p,
                                                                const StepParams& sp,
                                                                const RegArgs& a,
                                                                const RegPlace& pl) {
  const int lane = reg_lane();
  const int K = pl.K, W = sp.W, H = sp.H, P = sp.P;
  const int gi = pl.gx0 + 2 * lane;
  bool own0 = lc_in_tile(2 * lane, K, a.tw) && (unsigned)gi < (unsigned)W;
  bool own1 = lc_in_tile(2 * lane + 1, K, a.tw) && (unsigned)(gi + 1) < (unsigned)W;
  // RING: columns within 2K of the tile's left or right edge (pairs never
  // straddle that bound: 2K and the tile width are even)
  const int oc = 2 * lane - 2 * K;
  const bool ring_col = oc < 2 * K || oc >= a.tw - 2 * K;
#pragma unroll
  for (int rho = 0; rho < R; rho++) {
    const int r = pl.rbase + rho;
    const int gj = pl.gy0 + r;
    if (r < 2 * K || r >= 2 * K + a.th || gj >= H) continue;
    bool s0 = own0, s1 = own1;
    if constexpr (RING) {
      const int orow = r - 2 * K;
      const bool ring_row = orow < 2 * K || orow >= a.th - 2 * K;
      s0 = s0 && (ring_row || ring_col);
      s1 = s1 && (ring_row || ring_col);
    }
    const long gk = (long)gj * P + gi;
    if constexpr (RING) {
      coherent_store_pair(u_out, gk, s0, s1, t.u0[rho], t.u1[rho]);
      coherent_store_pair(v_out, gk, s0, s1, t.v0[rho], t.v1[rho]);
    } else {
      store_pair(u_out, gk, s0, s1, t.u0[rho], t.u1[rho]);
      store_pair(v_out, gk, s0, s1, t.v0[rho], t.v1[rho]);
    }
    if constexpr (PRESSURE && !RING) {
      if (s0) p[gk] = t.s0[rho];
      if (s1) p[gk + 1] = t.s1[rho];
    }
  }
}

// Reloads u and v of the halo (every register cell outside the owned tile)
// from the neighbours' stored rings; the owned cells keep their registers.
// Loads first (masked to the halo lanes), one wait, then cells outside the
// grid are set to 0, as at the first load.
template <typename T, int R>
__device__ __attribute__((always_inline)) inline void reg_reload_halo(RegTile<R>& t,
                                                                      const T* __restrict__ u_in,
                                                                      const T* __restrict__ v_in,
                                                                      const StepParams& sp,
                                                                      const RegArgs& a,
                                                                      const RegPlace& pl) {
  const int lane = reg_lane();
  const int K = pl.K, W = sp.W, H = sp.H, P = sp.P;
  const int gi = pl.gx0 + 2 * lane;
  const bool col0 = (unsigned)gi < (unsigned)W, col1 = (unsigned)(gi + 1) < (unsigned)W;
  const bool halo_lane = lane < K || lane >= 64 - K;
  const int gic = min(max(gi, 0), P - 2);
  float nu0[R], nu1[R], nv0[R], nv1[R];
#pragma unroll
  for (int rho = 0; rho < R; rho++) {
    const int r = pl.rbase + rho;
    const bool halo_row = r < 2 * K || r >= 2 * K + a.th;
    nu0[rho] = t.u0[rho];
    nu1[rho] = t.u1[rho];
    nv0[rho] = t.v0[rho];
    nv1[rho] = t.v1[rho];
    if (halo_row || halo_lane) {
      const long k = (long)min(max(pl.gy0 + r, 0), H - 1) * P + gic;
      coherent_load_pair(u_in, k, nu0[rho], nu1[rho]);
      coherent_load_pair(v_in, k, nv0[rho], nv1[rho]);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int rho = 0; rho < R; rho++) {
    const int gj = pl.gy0 + pl.rbase + rho;
    const bool row_in = (unsigned)gj < (unsigned)H;
    const bool in0 = row_in && col0, in1 = row_in && col1;
    t.u0[rho] = in0 ? nu0[rho] : 0.0f;
    t.v0[rho] = in0 ? nv0[rho] : 0.0f;
    t.u1[rho] = in1 ? nu1[rho] : 0.0f;
    t.v1[rho] = in1 ? nv1[rho] : 0.0f;
  }
}

// One launch of the register-resident projection for tile `tile` (row-major
// over the grid's ntx tiles per row), run by a workgroup of 64 * kRegWaves
// threads.
template <typename T, bool PRESSURE, int R, bool ONE_SYNC>
__device__ __attribute__((always_inline)) inline void project_regs_tile(
    const StepParams& sp, const uint8_t* __restrict__ gates, const T* __restrict__ u_in,
    const T* __restrict__ v_in, T* __restrict__ u_out, T* __restrict__ v_out,
    float* __restrict__ p, const RegArgs& a, int tile, RegShared& sh) {
  const RegPlace pl = reg_place<R>(a, tile);
  RegTile<R> t;
  reg_load<T, PRESSURE, R>(t, sp, gates, u_in, v_in, p, a, pl);
  reg_sweeps<T, PRESSURE, R, ONE_SYNC>(t, sh, sp, pl, a.th, a.sweeps);
  reg_store<T, PRESSURE, R, false>(t, u_out, v_out, p, sp, a, pl);
}

}  // namespace osy
