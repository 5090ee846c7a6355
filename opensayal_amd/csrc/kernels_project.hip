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
#include "step_device.hpp"

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

struct TiledArgs {
  int tw, th, k;  // owned tile, sweeps per launch (halo 2k)
  int pw, rh;     // region: pw cell pairs (2*pw cells) x rh rows
  int wpr;        // wavefronts per region row
  int ntx;        // tiles per grid row
  int sweeps;     // sweeps this launch (<= k)
  int forces;     // apply the velocity forces on load, pressure sums from 0
};

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

  const int tx = blockIdx.x % a.ntx, ty = blockIdx.x / a.ntx;
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
    if (a.forces) {
      if (in0) force_uv(sp, gi, gj, a0, b0);
      if (in1) force_uv(sp, gi + 1, gj, a1, b1);
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
        U[own] = gated_add(uu, d, g, 0);
        U[right] = gated_sub(ur, d, g, 1);
        V[own] = gated_add(vv, d, g, 2);
        V[top] = gated_sub(vt, d, g, 3);
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
__device__ inline uint32_t lane_from_right_u(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x130, 0xf, 0xf, true);
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

// update_pressure_at (src/fluid.cu:221-227): ((d * density) * cell_size) / d_t
__device__ inline float pressure_term(float d, float density, float csf, float d_t) {
  return ((d * density) * csf) / d_t;
}

// Register state of one lane of the register-resident kernel.
template <int R>
struct RegTile {
  float u0[R], u1[R], v0[R], v1[R];  // faces of cells 2l and 2l+1, rows 0..R-1
  float s0[R], s1[R];                // pressure sums (pressure runs only)
  uint32_t gw[R];                    // gate bytes: cell 2l in bits 0-7, 2l+1 in 8-15
};

// apply_projection_at for the active cell of row rho (parity B) of every
// lane.  mr masks rows outside the box, m0/m1 lanes outside it.
template <int R, int B, bool PRESSURE>
__device__ inline void reg_row(RegTile<R>& t, const int rho, uint32_t mr, uint32_t m0,
                               uint32_t m1, float vt0, float vt1, float& dtop, float o,
                               float density, float csf, float d_t) {
  const int up = rho + 1 < R ? rho + 1 : rho;  // row above (when inside the strip)
  if constexpr (B == 0) {
    // active cell 2l: faces u0, u1, v0[rho], v0[rho+1] (all in this lane)
    const uint32_t g = t.gw[rho] & m0 & mr;
    const float uu = t.u0[rho], ur = t.u1[rho], vv = t.v0[rho];
    const float vt = (rho + 1 < R) ? t.v0[up] : vt0;
    const float div = ((ur - uu) + vt) - vv;
    const float d = relaxed_quotient(div, g, o);
    if constexpr (PRESSURE) {
      if (g & G_UPD) t.s0[rho] = t.s0[rho] + pressure_term(d, density, csf, d_t);
    }
    t.u0[rho] = gated_add(uu, d, g, 0);
    t.u1[rho] = gated_sub(ur, d, g, 1);
    t.v0[rho] = gated_add(vv, d, g, 2);
    if (rho + 1 < R)
      t.v0[up] = gated_sub(vt, d, g, 3);
    else
      dtop = gate_pos(d, g, 3);
  } else {
    // active cell 2l+1: faces u1, u0 of lane l+1, v1[rho], v1[rho+1]
    const uint32_t g = (t.gw[rho] >> 8) & m1 & mr;
    const float uu = t.u1[rho], ur = lane_from_right(t.u0[rho]), vv = t.v1[rho];
    const float vt = (rho + 1 < R) ? t.v1[up] : vt1;
    const float div = ((ur - uu) + vt) - vv;
    const float d = relaxed_quotient(div, g, o);
    if constexpr (PRESSURE) {
      if (g & G_UPD) t.s1[rho] = t.s1[rho] + pressure_term(d, density, csf, d_t);
    }
    t.u1[rho] = gated_add(uu, d, g, 0);
    t.u0[rho] = t.u0[rho] - lane_from_left(gate_pos(d, g, 1));
    t.v1[rho] = gated_add(vv, d, g, 2);
    if (rho + 1 < R)
      t.v1[up] = gated_sub(vt, d, g, 3);
    else
      dtop = gate_pos(d, g, 3);
  }
}

// Rows RHO..R-1 of one phase, unrolled at compile time.
template <int R, int PB, bool PRESSURE, int RHO>
__device__ inline void reg_rows_from(RegTile<R>& t, int rlo, int rhi_rel, uint32_t m0, uint32_t m1,
                                     float vt0, float vt1, float& dtop, float o, float density,
                                     float csf, float d_t) {
  if constexpr (RHO < R) {
    const uint32_t mr = (RHO >= rlo && RHO < rhi_rel) ? 0xffu : 0u;  // wavefront-uniform
    reg_row<R, ((PB + RHO) & 1), PRESSURE>(t, RHO, mr, m0, m1, vt0, vt1, dtop, o, density, csf,
                                           d_t);
    reg_rows_from<R, PB, PRESSURE, RHO + 1>(t, rlo, rhi_rel, m0, m1, vt0, vt1, dtop, o, density,
                                            csf, d_t);
  }
}

// One colour phase.  PB is the parity of the active cell in row 0 of this
// wavefront (compile time: phases come in pairs with PB and 1 - PB).  Rows
// outside the box of cells that still reach the tile are masked rather
// than skipped, so the phase is straight-line code over registers.
template <int R, int PB, bool PRESSURE>
__device__ inline void reg_phase(RegTile<R>& t, float2 (*vrow)[64], float (*dsend)[64],
                                 int lane, int w, int rbase, int s, int nphase, int K, int tw,
                                 int th, float o, float density, float csf, float d_t) {
  const int e = nphase - 1 - s;
  const int lo = 2 * K - 1 - e;
  const int chi = 2 * K + tw + e;
  const int rhi = 2 * K + th + e;
  const int lc = 2 * lane;
  const uint32_t m0 = (lc >= lo && lc < chi) ? 0xffu : 0u;
  const uint32_t m1 = (lc + 1 >= lo && lc + 1 < chi) ? 0xffu : 0u;
  const int rlo = lo - rbase, rhi_rel = rhi - rbase;  // box rows of this wavefront
  // publish row 0 for the wavefront below, read the row above
  if (w > 0) vrow[w][lane] = make_float2(t.v0[0], t.v1[0]);
  __syncthreads();
  float vt0 = 0.0f, vt1 = 0.0f;
  if (w < kRegWaves - 1) {
    const float2 x = vrow[w + 1][lane];
    vt0 = x.x;
    vt1 = x.y;
  }
  float dtop = 0.0f;
  if (rlo < R && rhi_rel > 0)  // wavefront-uniform: some row is in the box
    reg_rows_from<R, PB, PRESSURE, 0>(t, rlo, rhi_rel, m0, m1, vt0, vt1, dtop, o, density, csf,
                                      d_t);
  if (w < kRegWaves - 1) dsend[w][lane] = dtop;
  __syncthreads();
  if (w > 0) {
    // the wavefront below moved the top faces of its last row: they are
    // this wavefront's row-0 v of parity 1 - PB
    const float dd = dsend[w - 1][lane];
    if constexpr (PB == 0)
      t.v1[0] = t.v1[0] - dd;
    else
      t.v0[0] = t.v0[0] - dd;
  }
}

template <typename T, bool PRESSURE, int R>
__global__ void __launch_bounds__(64 * kRegWaves)
    k_project_regs(StepParams sp, const uint8_t* __restrict__ flags, const T* __restrict__ u_in,
                   const T* __restrict__ v_in, T* __restrict__ u_out, T* __restrict__ v_out,
                   float* __restrict__ p, RegArgs a) {
  __shared__ float2 vrow[kRegWaves][64];  // row 0 (v0, v1) of each wavefront
  __shared__ float dsend[kRegWaves][64];  // top-row update for the wavefront above
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int K = a.k;
  const int tx = blockIdx.x % a.ntx, ty = blockIdx.x / a.ntx;
  const int x0 = tx * a.tw, y0 = ty * a.th;
  const int gx0 = x0 - 2 * K, gy0 = y0 - 2 * K;  // region origin; gx0 is even
  const int W = sp.W, H = sp.H, P = sp.P;
  const int rbase = w * R;  // first region row of this wavefront
  const int gi = gx0 + 2 * lane;
  const bool col0 = (unsigned)gi < (unsigned)W, col1 = (unsigned)(gi + 1) < (unsigned)W;

  RegTile<R> t;
  // flags: rows rbase .. rbase+R (one above for the top faces), columns
  // 2l, 2l+1 of this lane and 2l+2 (lane l+1's first column; lane 63 loads
  // its own)
  uint32_t fl0[R + 1], fl1[R + 1], fl2[R + 1];
#pragma unroll
  for (int rho = 0; rho <= R; rho++) {
    const int gj = gy0 + rbase + rho;
    const bool row_in = (unsigned)gj < (unsigned)H;
    const long gk = (long)gj * P + gi;
    fl0[rho] = (row_in && col0) ? flags[gk] : 0u;
    fl1[rho] = (row_in && col1) ? flags[gk + 1] : 0u;
    const uint32_t from_right = lane_from_right_u(fl0[rho]);
    fl2[rho] = (lane == 63) ? ((row_in && (unsigned)(gi + 2) < (unsigned)W) ? flags[gk + 2] : 0u)
                            : from_right;
  }
#pragma unroll
  for (int rho = 0; rho < R; rho++) {
    const int gj = gy0 + rbase + rho;
    const bool row_in = (unsigned)gj < (unsigned)H;
    const long gk = (long)gj * P + gi;
    load_pair(u_in, gk, row_in && col0, row_in && col1, t.u0[rho], t.u1[rho]);
    load_pair(v_in, gk, row_in && col0, row_in && col1, t.v0[rho], t.v1[rho]);
    if (a.forces) {
      if (row_in && col0) force_uv(sp, gi, gj, t.u0[rho], t.v0[rho]);
      if (row_in && col1) force_uv(sp, gi + 1, gj, t.u1[rho], t.v1[rho]);
    }
    if constexpr (PRESSURE) {
      t.s0[rho] = (row_in && col0 && !a.forces) ? p[gk] : 0.0f;
      t.s1[rho] = (row_in && col1 && !a.forces) ? p[gk + 1] : 0.0f;
    } else {
      t.s0[rho] = t.s1[rho] = 0.0f;
    }
    t.gw[rho] = gate_byte(fl0[rho], fl1[rho], fl0[rho + 1]) |
                (gate_byte(fl1[rho], fl2[rho], fl1[rho + 1]) << 8);
  }

  const float o = sp.o;
  const float csf = (float)sp.cs;
  const int nphase = 2 * a.sweeps;
  // colour 0 (even cells) first; row 0's active parity in a colour-0 phase
  // is (gy0 + rbase) & 1 and flips every phase
  if (((gy0 + rbase) & 1) == 0) {
    for (int s = 0; s < nphase; s += 2) {
      reg_phase<R, 0, PRESSURE>(t, vrow, dsend, lane, w, rbase, s, nphase, K, a.tw, a.th, o,
                                sp.density, csf, sp.d_t);
      reg_phase<R, 1, PRESSURE>(t, vrow, dsend, lane, w, rbase, s + 1, nphase, K, a.tw, a.th, o,
                                sp.density, csf, sp.d_t);
    }
  } else {
    for (int s = 0; s < nphase; s += 2) {
      reg_phase<R, 1, PRESSURE>(t, vrow, dsend, lane, w, rbase, s, nphase, K, a.tw, a.th, o,
                                sp.density, csf, sp.d_t);
      reg_phase<R, 0, PRESSURE>(t, vrow, dsend, lane, w, rbase, s + 1, nphase, K, a.tw, a.th, o,
                                sp.density, csf, sp.d_t);
    }
  }

  // write back the owned tile
  const bool own0 = lc_in_tile(2 * lane, K, a.tw) && col0;
  const bool own1 = lc_in_tile(2 * lane + 1, K, a.tw) && col1;
#pragma unroll
  for (int rho = 0; rho < R; rho++) {
    const int r = rbase + rho;
    const int gj = gy0 + r;
    if (r < 2 * K || r >= 2 * K + a.th || gj >= H) continue;
    const long gk = (long)gj * P + gi;
    store_pair(u_out, gk, own0, own1, t.u0[rho], t.u1[rho]);
    store_pair(v_out, gk, own0, own1, t.v0[rho], t.v1[rho]);
    if constexpr (PRESSURE) {
      if (own0) p[gk] = t.s0[rho];
      if (own1) p[gk + 1] = t.s1[rho];
    }
  }
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
                       0, s, sp, g.flags, (const T*)f.u[*cur], (const T*)f.v[*cur], (T*)f.u[nxt],
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

template <typename T>
void enqueue_projection(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                        const ProjTuning& t, hipStream_t s) {
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
