// kernels_project.hip - over-relaxed red-black projection (src/fluid.cu:229-295).
//
// A sweep is the even colour phase ((i+j) even) then the odd one, over
// interior fluid cells.  Same-colour cells share no face, so every colour
// phase is a set of independent cell updates and any schedule that finishes
// one phase before the next starts reproduces the reference bit for bit.
//
// Two kernels:
//  * k_project_colour: one launch per colour phase, in place (2n launches);
//    kept for tiny grids and as the simple restatement.
//  * k_project_tiled: temporal blocking.  A workgroup loads a region of
//    u, v and cell flags into LDS, runs up to k sweeps there, and writes the
//    tile it owns to the other half of a ping-pong pair.  A colour phase
//    changes only faces of the cells it updates, so a value can be wrong
//    after phase s only within s cells of the region edge; a halo of 2k
//    cells therefore keeps every owned face exact through 2k phases.  Each
//    phase updates only the cells whose results still reach the owned tile
//    (a box that shrinks by one cell per phase), so the redundant halo work
//    falls off as the sweeps proceed.
#include "projection.hpp"

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

// Per-cell gate byte of the tiled kernel, built in LDS from the flags.
enum : uint32_t {
  G_UPD = 1u << 0,  // cell is projected
  G_L = 1u << 1,    // u(i,j) moves   (left neighbour fluid)
  G_R = 1u << 2,    // u(i+1,j) moves (right neighbour fluid)
  G_B = 1u << 3,    // v(i,j) moves   (bottom neighbour fluid)
  G_T = 1u << 4,    // v(i,j+1) moves (top neighbour fluid)
  G_S_SHIFT = 5,    // bits 5..7: total_s
};

struct TiledArgs {
  int pw, rh, k;      // region: 2*pw x rh cells, k sweeps per launch
  int tw, th;         // owned tile
  int ntx;            // tiles per row
  int sweeps;         // sweeps this launch (<= k)
};

// LDS layout: U and V hold rh rows of two half-rows each, the even columns
// of the region row then the odd ones, so the 64 lanes of a wave that update
// one colour of a row touch 64 consecutive words (no bank conflicts).
template <typename T, bool PRESSURE>
__global__ void __launch_bounds__(kTileThreads)
    k_project_tiled(StepParams sp, const uint8_t* __restrict__ flags, const T* __restrict__ u_in,
                    const T* __restrict__ v_in, T* __restrict__ u_out, T* __restrict__ v_out,
                    float* __restrict__ p, TiledArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int pw = a.pw, rh = a.rh, K = a.k;
  const int pwp = pw + 1;
  const int rs = 2 * pwp;      // LDS row stride (floats)
  const int rw = 2 * pw;       // region width in cells
  const int fs = rw + 1;       // flag row stride (bytes)
  float* U = (float*)smem;
  float* V = U + rh * rs;
  uint8_t* F = (uint8_t*)(V + rh * rs);

  const int tx = blockIdx.x % a.ntx, ty = blockIdx.x / a.ntx;
  const int x0 = tx * a.tw, y0 = ty * a.th;
  const int gx0 = x0 - 2 * K, gy0 = y0 - 2 * K;  // region origin (even: tw, 2K even)
  const int W = sp.W, H = sp.H, P = sp.P;
  const int nitems = pw * rh;
  const int M = (nitems + kTileThreads - 1) / kTileThreads;

  // flags of the region plus one extra column and row (right/top neighbours)
  for (int q = threadIdx.x; q < (rh + 1) * fs; q += kTileThreads) {
    const int r = q / fs, c = q - r * fs;
    const int gi = gx0 + c, gj = gy0 + r;
    F[q] = ((unsigned)gi < (unsigned)W && (unsigned)gj < (unsigned)H) ? flags[(long)gj * P + gi]
                                                                      : (uint8_t)0;
  }
  // u, v of the region, de-interleaved by column parity
  for (int q = threadIdx.x; q < nitems; q += kTileThreads) {
    const int r = q / pw, pp = q - r * pw;
    const int gj = gy0 + r;
    const bool row_in = (unsigned)gj < (unsigned)H;
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int gi = gx0 + 2 * pp + b;
      const bool in = row_in && (unsigned)gi < (unsigned)W;
      const long gk = (long)gj * P + gi;
      U[r * rs + b * pwp + pp] = in ? ldf(u_in, gk) : 0.0f;
      V[r * rs + b * pwp + pp] = in ? ldf(v_in, gk) : 0.0f;
    }
  }
  __syncthreads();

  // per-item constants: position, gate bytes of the item's two cells, and
  // the pressure accumulators of owned cells
  int rp[kTileMaxItems];
  uint32_t gates[kTileMaxItems];
  float pacc[PRESSURE ? kTileMaxItems : 1][2];
#pragma unroll
  for (int m = 0; m < kTileMaxItems; m++) {
    rp[m] = 0;
    gates[m] = 0;
    if (m < M) {
      const int q = threadIdx.x + m * kTileThreads;
      if (q < nitems) {
        const int r = q / pw, pp = q - r * pw;
        rp[m] = (r << 16) | pp;
        uint32_t g2 = 0;
#pragma unroll
        for (int b = 0; b < 2; b++) {
          const int lc = 2 * pp + b;
          const uint8_t f = F[r * fs + lc];
          const uint8_t fr = F[r * fs + lc + 1];
          const uint8_t ft = F[(r + 1) * fs + lc];
          uint32_t g = 0;
          if (f & CF_UPD) g |= G_UPD;
          if (f & CF_FU) g |= G_L;
          if (fr & CF_FU) g |= G_R;
          if (f & CF_FV) g |= G_B;
          if (ft & CF_FV) g |= G_T;
          g |= (uint32_t)cf_total_s(f) << G_S_SHIFT;
          g2 |= g << (8 * b);
          if constexpr (PRESSURE) {
            const int gi = gx0 + lc, gj = gy0 + r;
            const bool owned = lc >= 2 * K && lc < 2 * K + a.tw && r >= 2 * K &&
                               r < 2 * K + a.th && gi < W && gj < H;
            pacc[m][b] = owned ? p[(long)gj * P + gi] : 0.0f;
          }
        }
        gates[m] = g2;
      }
    }
  }

  const float o = sp.o;
  const int nphase = 2 * a.sweeps;
  for (int s = 0; s < nphase; s++) {
    const int c = s & 1;                 // 0: even cells first, as the reference
    const int e = nphase - 1 - s;        // phases still to come after this one
    const int lo = 2 * K - 1 - e;        // needed box, region-local cells
    const int chi = 2 * K + a.tw + e;
    const int rhi = 2 * K + a.th + e;
#pragma unroll
    for (int m = 0; m < kTileMaxItems; m++) {
      if (m >= M) break;
      const uint32_t gg = gates[m];
      const int r = rp[m] >> 16, pp = rp[m] & 0xffff;
      const int b = (c + gy0 + r) & 1;   // the pair's cell of colour c (gx0 is even)
      const int lc = 2 * pp + b;
      const uint32_t g = (gg >> (8 * b)) & 0xffu;
      if (!(g & G_UPD) || r < lo || r >= rhi || lc < lo || lc >= chi) continue;
      const int own = r * rs + b * pwp + pp;
      const int right = b ? own - pwp + 1 : own + pwp;  // u(i+1): next pair's even cell
      const int top = own + rs;
      const float uu = U[own], ur = U[right], vv = V[own], vt = V[top];
      const float div = ((ur - uu) + vt) - vv;
      const float d = o * (div / (float)(g >> G_S_SHIFT));
      if constexpr (PRESSURE) {
        const float t = ((d * sp.density) * (float)sp.cs) / sp.d_t;
        pacc[m][b] = pacc[m][b] + t;
      }
      if (g & G_L) U[own] = uu + d;
      if (g & G_R) U[right] = ur - d;
      if (g & G_B) V[own] = vv + d;
      if (g & G_T) V[top] = vt - d;
    }
    __syncthreads();
  }

  // write back the owned tile
#pragma unroll
  for (int m = 0; m < kTileMaxItems; m++) {
    if (m >= M) break;
    const int q = threadIdx.x + m * kTileThreads;
    if (q >= nitems) break;
    const int r = rp[m] >> 16, pp = rp[m] & 0xffff;
    if (r < 2 * K || r >= 2 * K + a.th) continue;
    const int gj = gy0 + r;
    if (gj >= H) continue;
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int lc = 2 * pp + b;
      const int gi = gx0 + lc;
      if (lc < 2 * K || lc >= 2 * K + a.tw || gi >= W) continue;
      const long gk = (long)gj * P + gi;
      stf(u_out, gk, U[r * rs + b * pwp + pp]);
      stf(v_out, gk, V[r * rs + b * pwp + pp]);
      if constexpr (PRESSURE) p[gk] = pacc[m][b];
    }
  }
}

template <typename T, bool PRESSURE>
void launch_tiled(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                  const ProjTuning& t, hipStream_t s) {
  TiledArgs a;
  a.pw = t.pw;
  a.rh = t.rh;
  a.k = t.k;
  a.tw = t.tile_w();
  a.th = t.tile_h();
  a.ntx = (sp.W + a.tw - 1) / a.tw;
  const int nty = (sp.H + a.th - 1) / a.th;
  const size_t lds = t.lds_bytes();
  for (int done = 0; done < sp.n; done += t.k) {
    a.sweeps = (sp.n - done < t.k) ? sp.n - done : t.k;
    const int nxt = 1 - *cur;
    hipLaunchKernelGGL((k_project_tiled<T, PRESSURE>), dim3(a.ntx * nty), dim3(kTileThreads), lds,
                       s, sp, g.flags, (const T*)f.u[*cur], (const T*)f.v[*cur], (T*)f.u[nxt],
                       (T*)f.v[nxt], f.p, a);
    *cur = nxt;
  }
}

template <typename T>
hipError_t allow_large_lds() {
  const int bytes = 160 * 1024;
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
  if (t.kernel == PROJ_TILED) {
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
