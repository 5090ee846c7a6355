// kernels_project.hip - over-relaxed red-black projection (src/fluid.cu:229-295).
//
// A sweep is the even colour phase ((i+j) even) then the odd one, over
// interior fluid cells.  Same-colour cells share no face, so every colour
// phase is a set of independent cell updates and any schedule that finishes
// one phase before the next starts reproduces the reference bit for bit.
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

}  // namespace

template <typename T>
void enqueue_projection(const StepParams& sp, const GridView& g, const FieldPtrs& f, int* cur,
                        int kernel, hipStream_t s) {
  (void)kernel;
  T* u = (T*)f.u[*cur];
  T* v = (T*)f.v[*cur];
  const dim3 grid((sp.W / 2 + 63) / 64 + 1, (sp.H + 3) / 4), block(64, 4);
  for (int sweep = 0; sweep < sp.n; sweep++) {
    hipLaunchKernelGGL(k_project_colour<T>, grid, block, 0, s, sp, g.flags, u, v, f.p, 0);
    hipLaunchKernelGGL(k_project_colour<T>, grid, block, 0, s, sp, g.flags, u, v, f.p, 1);
  }
}

template void enqueue_projection<float>(const StepParams&, const GridView&, const FieldPtrs&, int*,
                                        int, hipStream_t);
template void enqueue_projection<__half>(const StepParams&, const GridView&, const FieldPtrs&,
                                         int*, int, hipStream_t);
template void enqueue_projection<__hip_bfloat16>(const StepParams&, const GridView&,
                                                 const FieldPtrs&, int*, int, hipStream_t);

}  // namespace osy
