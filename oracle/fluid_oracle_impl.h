/*
 * fluid_oracle_impl.h - body of the CPU oracle, included twice by
 * fluid_oracle.c: once with REAL=float (the reference's arithmetic) and once
 * with REAL=double (the FP64 run used to state precision-mode errors).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; tests, the smoke check and the benchmark's CPU-baseline leg do.
 *
 * Every function restates one function of the reference's src/fluid.cu with
 * the same operation order.  With REAL=float every operation is an IEEE
 * binary32 operation in source order (the file is compiled with
 * -ffp-contract=off, so no fused multiply-add), and the sub-expressions the
 * reference writes in double (cell_size / 2.0, (i + 0.5) * cell_size,
 * 1.0 / (distance + 1e-6), max(..., 0.0)) are evaluated in double and then
 * rounded, as C++ does.  Field arrays use the reference layout
 * indx(i, j) = (H - 1 - j) * W + i (src/fluid.cu:163-165).
 */

#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define FN(name) ORC_CAT(name, SFX)

/* int i = x / cell_size (src/fluid.cu:419-420, 480-481, 645-646).  The
 * reference converts a float to int directly, which is undefined for NaN and
 * out-of-range values; this restatement clamps the quotient to +-2^24 first
 * (NaN goes to -2^24), so every non-finite or far-away sample position lands
 * outside the grid.  For every in-range value this is plain truncation. */
static inline int FN(orc_to_index)(REAL q) {
  if (!(q >= (REAL)-16777216.0)) q = (REAL)-16777216.0;
  if (q > (REAL)16777216.0) q = (REAL)16777216.0;
  return (int)q;
}

/* get_general_velocity_x (src/fluid.cu:479-539): masked, non-renormalised
 * bilinear interpolation of u at (x, y). */
REAL FN(orc_gx)(const orc_grid* g, const REAL* u, REAL x, REAL y) {
  const int cs = g->cs;
  const int i = FN(orc_to_index)(x / (REAL)cs);
  const int j = FN(orc_to_index)(y / (REAL)cs);
  if (!orc_valid(g, i, j)) return 0;
  const REAL in_x = x - (REAL)(i * cs);
  const REAL in_y = y - (REAL)(j * cs);
  const double h = cs / 2.0;
  REAL acc = 0;
  if ((double)in_y <= h) { /* take the bottom row pair (src/fluid.cu:493-513) */
    const REAL d_y = (REAL)(h - (double)in_y);
    const REAL w_x = (REAL)1 - in_x / (REAL)cs;
    const REAL w_y = (REAL)1 - d_y / (REAL)cs;
    if (orc_valid(g, i, j)) acc = acc + (w_y * w_x) * u[orc_idx(g, i, j)];
    if (orc_valid(g, i + 1, j))
      acc = acc + (w_y * ((REAL)1 - w_x)) * u[orc_idx(g, i + 1, j)];
    if (orc_valid(g, i, j - 1))
      acc = acc + (((REAL)1 - w_y) * w_x) * u[orc_idx(g, i, j - 1)];
    if (orc_valid(g, i + 1, j - 1))
      acc = acc +
            (((REAL)1 - w_y) * ((REAL)1 - w_x)) * u[orc_idx(g, i + 1, j - 1)];
  } else { /* take the top row pair (src/fluid.cu:516-536) */
    const REAL d_y = (REAL)((double)in_y - h);
    const REAL w_x = (REAL)1 - in_x / (REAL)cs;
    const REAL w_y = (REAL)1 - d_y / (REAL)cs;
    if (orc_valid(g, i, j)) acc = acc + (w_y * w_x) * u[orc_idx(g, i, j)];
    if (orc_valid(g, i, j + 1))
      acc = acc + (((REAL)1 - w_y) * w_x) * u[orc_idx(g, i, j + 1)];
    if (orc_valid(g, i + 1, j))
      acc = acc + (w_y * ((REAL)1 - w_x)) * u[orc_idx(g, i + 1, j)];
    if (orc_valid(g, i + 1, j + 1))
      acc = acc +
            (((REAL)1 - w_y) * ((REAL)1 - w_x)) * u[orc_idx(g, i + 1, j + 1)];
  }
  return acc;
}

/* get_general_velocity_y (src/fluid.cu:418-477); note the strict `<`. */
REAL FN(orc_gy)(const orc_grid* g, const REAL* v, REAL x, REAL y) {
  const int cs = g->cs;
  const int i = FN(orc_to_index)(x / (REAL)cs);
  const int j = FN(orc_to_index)(y / (REAL)cs);
  if (!orc_valid(g, i, j)) return 0;
  const REAL in_x = x - (REAL)(i * cs);
  const REAL in_y = y - (REAL)(j * cs);
  const double h = cs / 2.0;
  REAL acc = 0;
  if ((double)in_x < h) { /* left column pair (src/fluid.cu:432-452) */
    const REAL d_x = (REAL)(h - (double)in_x);
    const REAL w_x = (REAL)1 - d_x / (REAL)cs;
    const REAL w_y = (REAL)1 - in_y / (REAL)cs;
    if (orc_valid(g, i, j)) acc = acc + (w_y * w_x) * v[orc_idx(g, i, j)];
    if (orc_valid(g, i - 1, j))
      acc = acc + (w_y * ((REAL)1 - w_x)) * v[orc_idx(g, i - 1, j)];
    if (orc_valid(g, i - 1, j + 1))
      acc = acc +
            (((REAL)1 - w_y) * ((REAL)1 - w_x)) * v[orc_idx(g, i - 1, j + 1)];
    if (orc_valid(g, i, j + 1))
      acc = acc + (((REAL)1 - w_y) * w_x) * v[orc_idx(g, i, j + 1)];
  } else { /* right column pair (src/fluid.cu:454-474) */
    const REAL d_x = (REAL)((double)in_x - h);
    const REAL w_x = (REAL)1 - d_x / (REAL)cs;
    const REAL w_y = (REAL)1 - in_y / (REAL)cs;
    if (orc_valid(g, i, j)) acc = acc + (w_y * w_x) * v[orc_idx(g, i, j)];
    if (orc_valid(g, i, j + 1))
      acc = acc + (((REAL)1 - w_y) * w_x) * v[orc_idx(g, i, j + 1)];
    if (orc_valid(g, i + 1, j + 1))
      acc = acc +
            (((REAL)1 - w_y) * ((REAL)1 - w_x)) * v[orc_idx(g, i + 1, j + 1)];
    if (orc_valid(g, i + 1, j))
      acc = acc + (w_y * ((REAL)1 - w_x)) * v[orc_idx(g, i + 1, j)];
  }
  return acc;
}

/* get_center_position (src/fluid.cu:547-550): double product, then rounded. */
static inline REAL FN(orc_center)(int a, int cs) {
  return (REAL)(((double)a + 0.5) * (double)cs);
}

/* get_distance (inc/helper.cuh:77-81) in REAL arithmetic. */
static inline REAL FN(orc_distance)(REAL ax, REAL ay, REAL bx, REAL by) {
  const REAL dx = ax - bx;
  const REAL dy = ay - by;
  return ORC_SQRT(dx * dx + dy * dy);
}

/* interpolate_smoke (src/fluid.cu:644-716): inverse-distance weights over
 * the sample's cell and the three cells of its nearest quadrant; terms only
 * from valid fluid cells, weights not renormalised. */
REAL FN(orc_idw)(const orc_grid* g, const REAL* smoke, REAL x, REAL y) {
  const int cs = g->cs;
  const int i = FN(orc_to_index)(x / (REAL)cs);
  const int j = FN(orc_to_index)(y / (REAL)cs);
  const REAL in_x = x - (REAL)(i * cs);
  const REAL in_y = y - (REAL)(j * cs);
  const double h = cs / 2.0;
  int ci[4], cj[4];
  ci[0] = i;
  cj[0] = j;
  if ((double)in_x < h && (double)in_y < h) {
    ci[1] = i - 1; cj[1] = j;
    ci[2] = i;     cj[2] = j - 1;
    ci[3] = i - 1; cj[3] = j - 1;
  } else if ((double)in_x < h) {
    ci[1] = i - 1; cj[1] = j;
    ci[2] = i;     cj[2] = j + 1;
    ci[3] = i - 1; cj[3] = j + 1;
  } else if ((double)in_y < h) {
    ci[1] = i + 1; cj[1] = j;
    ci[2] = i;     cj[2] = j - 1;
    ci[3] = i + 1; cj[3] = j - 1;
  } else {
    ci[1] = i + 1; cj[1] = j;
    ci[2] = i;     cj[2] = j + 1;
    ci[3] = i + 1; cj[3] = j + 1;
  }
  REAL inv[4];
  for (int k = 0; k < 4; k++) {
    const REAL d = FN(orc_distance)(x, y, FN(orc_center)(ci[k], cs),
                                    FN(orc_center)(cj[k], cs));
    inv[k] = (REAL)(1.0 / ((double)d + 1e-6));
  }
  const REAL sum_inv = ((inv[0] + inv[1]) + inv[2]) + inv[3];
  REAL acc = 0;
  for (int k = 0; k < 4; k++) {
    const REAL w = inv[k] / sum_inv;
    if (orc_valid(g, ci[k], cj[k]))
      acc = acc + w * smoke[orc_idx(g, ci[k], cj[k])];
  }
  return acc;
}

/* apply_external_forces_at (src/fluid.cu:308-349) over every cell, plus
 * zero_pressure (src/fluid.cu:203-219) when pressure is enabled. */
void FN(orc_forces)(const osy_config* c, const orc_grid* g, REAL* u, REAL* v,
                    REAL* smoke, REAL* p, const osy_source* src, REAL d_t,
                    int nthreads) {
  const int W = g->W, H = g->H;
  const int ph = c->pipe_height, sl = c->smoke_length;
  const int sc = c->smoke_count, sh = c->smoke_height;
  /* smoke_spacing (src/fluid.cu:312-315); the reference divides by zero when
   * smoke_count == 1 but never reads the value then. */
  const int spacing = (sc != 1) ? (ph - sc * sh) / (sc - 1) : 0;
  const REAL speed = (REAL)c->wind_speed;
  const REAL wsmoke = (REAL)c->wind_smoke;
  const REAL drag = (REAL)c->drag_coeff;
  const REAL damping = (REAL)ORC_EXP(-drag * d_t);
  const REAL gdt = (REAL)c->g * d_t;
  const int active = src && src->active;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int j = 0; j < H; j++) {
    for (int i = 0; i < W; i++) {
      const int k = orc_idx(g, i, j);
      if (i <= sl && i != 0 && j >= H / 2 - ph / 2 && j <= H / 2 + ph / 2) {
        u[k] = speed;
        if ((sc == 1 && j >= H / 2 - sh / 2 && j <= H / 2 + sh / 2) ||
            (sc != 1 && (H / 2 + ph / 2 - j) % (spacing + sh) < sh)) {
          smoke[k] = wsmoke;
        }
      }
      if (drag != 0) {
        u[k] = u[k] * damping;
        v[k] = v[k] * damping;
      }
      if (active) {
        const int dx = i - src->position_x, dy = j - src->position_y;
        if (dx * dx + dy * dy < 40 * 40) {
          if (src->smoke != 0) smoke[k] = (REAL)src->smoke;
          u[k] = u[k] + (REAL)src->velocity * (REAL)dx;
          v[k] = v[k] + (REAL)src->velocity * (REAL)dy;
        }
      }
      v[k] = v[k] + gdt;
      if (p && c->enable_pressure) p[k] = 0;
    }
  }
}

/* apply_projection_at + update_pressure_at (src/fluid.cu:221-262). */
static inline void FN(orc_project_cell)(const osy_config* c, const orc_grid* g,
                                        REAL* u, REAL* v, REAL* p, REAL d_t,
                                        int i, int j) {
  if (g->solid[orc_idx(g, i, j)]) return;
  const int k = orc_idx(g, i, j);
  const int kr = orc_idx(g, i + 1, j);
  const int kt = orc_idx(g, i, j + 1);
  const REAL uu = u[k], vv = v[k], top_v = v[kt], right_u = u[kr];
  const REAL divergence = ((right_u - uu) + top_v) - vv;
  const int s = g->total_s[k];
  const REAL vd = (REAL)c->projection_o * (divergence / (REAL)s);
  if (c->enable_pressure)
    p[k] = p[k] + (((vd * (REAL)c->density) * (REAL)g->cs) / d_t);
  if (!g->solid[orc_idx(g, i - 1, j)]) u[k] = u[k] + vd;
  if (!g->solid[kr]) u[kr] = u[kr] - vd;
  if (!g->solid[orc_idx(g, i, j - 1)]) v[k] = v[k] + vd;
  if (!g->solid[kt]) v[kt] = v[kt] - vd;
}

/* One colour phase: apply_projection_even_kernel (colour 0, (i+j) even) or
 * apply_projection_odd_kernel (colour 1) (src/fluid.cu:264-280), interior
 * cells 1 <= i <= W-2, 1 <= j <= H-2.  Same-colour cells share no face, so
 * the row-parallel loop is deterministic. */
void FN(orc_projection_phase)(const osy_config* c, const orc_grid* g, REAL* u,
                              REAL* v, REAL* p, REAL d_t, int colour,
                              int nthreads) {
  const int W = g->W, H = g->H;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int j = 1; j <= H - 2; j++) {
    for (int i = 1 + ((j + colour + 1) & 1); i <= W - 2; i += 2)
      FN(orc_project_cell)(c, g, u, v, p, d_t, i, j);
  }
}

/* apply_projection (src/fluid.cu:282-295): `sweeps` times even then odd. */
void FN(orc_projection)(const osy_config* c, const orc_grid* g, REAL* u,
                        REAL* v, REAL* p, REAL d_t, int sweeps, int nthreads) {
  const long n = (long)g->W * g->H;
  for (int s = 0; s < sweeps; s++) {
    for (int colour = 0; colour < 2; colour++) {
      FN(orc_projection_phase)(c, g, u, v, p, d_t, colour, nthreads);
      FN(orc_round)(u, n, c->storage);
      FN(orc_round)(v, n, c->storage);
    }
  }
}

/* apply_extrapolation (src/fluid.cu:720-747) in a fixed two-phase order.
 * Row phase: u(i,0) = u(i,1), v(i,1) = 0, u(i,H-1) = u(i,H-2).  Column
 * phase, reading row-phase results: v(0,j) = v(1,j), u(1,j) = 0,
 * v(W-1,j) = v(W-2,j).  This is the reference's outcome whenever the block
 * of row 0 runs before the blocks of rows 1; the only cells where the
 * reference races are v(0,1) and v(W-1,1), here 0. */
void FN(orc_extrapolation)(const orc_grid* g, REAL* u, REAL* v) {
  const int W = g->W, H = g->H;
  for (int i = 0; i < W; i++) {
    u[orc_idx(g, i, 0)] = u[orc_idx(g, i, 1)];
    v[orc_idx(g, i, 1)] = 0;
    u[orc_idx(g, i, H - 1)] = u[orc_idx(g, i, H - 2)];
  }
  for (int j = 0; j < H; j++) {
    v[orc_idx(g, 0, j)] = v[orc_idx(g, 1, j)];
    u[orc_idx(g, 1, j)] = 0;
    v[orc_idx(g, W - 1, j)] = v[orc_idx(g, W - 2, j)];
  }
}

/* apply_velocity_advection_at + update (src/fluid.cu:598-642): Jacobi
 * back-trace of every u face and v face into (u_out, v_out). */
void FN(orc_advect_velocity)(const orc_grid* g, const REAL* u, const REAL* v,
                             REAL* u_out, REAL* v_out, REAL d_t,
                             int nthreads) {
  const int W = g->W, H = g->H, cs = g->cs;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int j = 0; j < H; j++) {
    for (int i = 0; i < W; i++) {
      const int k = orc_idx(g, i, j);
      /* u face: get_u_position, get_vertical_edge_velocity
       * (src/fluid.cu:552-554, 364-389) */
      {
        const REAL px = (REAL)(i * cs);
        const REAL py = (REAL)(((double)j + 0.5) * (double)cs);
        const REAL vel_x = u[k];
        REAL avg_v = v[k];
        int count = 1;
        if (orc_valid(g, i - 1, j + 1)) {
          avg_v = avg_v + v[orc_idx(g, i - 1, j + 1)];
          count++;
        }
        if (orc_valid(g, i, j + 1)) {
          avg_v = avg_v + v[orc_idx(g, i, j + 1)];
          count++;
        }
        if (orc_valid(g, i - 1, j)) {
          avg_v = avg_v + v[orc_idx(g, i - 1, j)];
          count++;
        }
        avg_v = avg_v / (REAL)count;
        const REAL x = px - vel_x * d_t;
        const REAL y = py - avg_v * d_t;
        u_out[k] = FN(orc_gx)(g, u, x, y);
      }
      /* v face: get_v_position, get_horizontal_edge_velocity
       * (src/fluid.cu:556-558, 391-416) */
      {
        const REAL px = (REAL)(((double)i + 0.5) * (double)cs);
        const REAL py = (REAL)(j * cs);
        const REAL vel_y = v[k];
        REAL avg_u = u[k];
        int count = 1;
        if (orc_valid(g, i + 1, j)) {
          avg_u = avg_u + u[orc_idx(g, i + 1, j)];
          count++;
        }
        if (orc_valid(g, i, j - 1)) {
          avg_u = avg_u + u[orc_idx(g, i, j - 1)];
          count++;
        }
        if (orc_valid(g, i + 1, j - 1)) {
          avg_u = avg_u + u[orc_idx(g, i + 1, j - 1)];
          count++;
        }
        avg_u = avg_u / (REAL)count;
        const REAL x = px - avg_u * d_t;
        const REAL y = py - vel_y * d_t;
        v_out[k] = FN(orc_gy)(g, v, x, y);
      }
    }
  }
}

/* apply_smoke_advection_at + update (src/fluid.cu:560-596): back-trace of
 * every cell centre through the (new) velocity, IDW gather into smoke_out. */
void FN(orc_advect_smoke)(const orc_grid* g, const REAL* u, const REAL* v,
                          const REAL* smoke, REAL* smoke_out, REAL d_t,
                          int nthreads) {
  const int W = g->W, H = g->H, cs = g->cs;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int j = 0; j < H; j++) {
    for (int i = 0; i < W; i++) {
      const REAL cx = FN(orc_center)(i, cs);
      const REAL cy = FN(orc_center)(j, cs);
      const REAL vx = FN(orc_gx)(g, u, cx, cy);
      const REAL vy = FN(orc_gy)(g, v, cx, cy);
      const REAL x = cx - vx * d_t;
      const REAL y = cy - vy * d_t;
      smoke_out[orc_idx(g, i, j)] = FN(orc_idw)(g, smoke, x, y);
    }
  }
}

/* decay_smoke_at (src/fluid.cu:758-762): max(smoke - rate*d_t, 0.0) with
 * the max taken in double (CUDA's max(float, double)). */
void FN(orc_decay)(const orc_grid* g, REAL* smoke, REAL rate, REAL d_t,
                   int nthreads) {
  const long n = (long)g->W * g->H;
  const REAL rd = rate * d_t;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long k = 0; k < n; k++) smoke[k] = (REAL)fmax((double)(smoke[k] - rd), 0.0);
}

/* apply_diffusion (src/fluid.cu:167-190), restated as Jacobi: `sweeps`
 * rounds in which every interior cell (1 <= i <= W-2, 1 <= j <= H-2; solid
 * or not, as the reference's kernel does) becomes
 *   (u + a * (((u(i-1,j) + u(i+1,j)) + u(i,j-1)) + u(i,j+1))) / (1 + 4a)
 * from the previous round's values; borders keep theirs; v is untouched.
 * a = viscosity * d_t / (cs * cs).  The reference updates u in place from
 * concurrent threads (reads may see either value) and counts its loop with
 * an uninitialised variable; the round count here is the loop's evident
 * bound, projection n, and every read sees the previous round.  `tmp`
 * holds W*H REAL values. */
void FN(orc_diffusion)(const osy_config* c, const orc_grid* g, REAL* u, REAL* tmp, REAL d_t,
                       int sweeps, int nthreads) {
  const int W = g->W, H = g->H;
  const REAL a = (REAL)c->viscosity * d_t / (REAL)(g->cs * g->cs);
  const REAL den = (REAL)1 + (REAL)4 * a;
  const long n = (long)W * H;
  for (int it = 0; it < sweeps; it++) {
    memcpy(tmp, u, sizeof(REAL) * n);
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int j = 1; j <= H - 2; j++) {
      for (int i = 1; i <= W - 2; i++) {
        const REAL sum = ((tmp[orc_idx(g, i - 1, j)] + tmp[orc_idx(g, i + 1, j)]) +
                          tmp[orc_idx(g, i, j - 1)]) +
                         tmp[orc_idx(g, i, j + 1)];
        u[orc_idx(g, i, j)] = (tmp[orc_idx(g, i, j)] + a * sum) / den;
      }
    }
    FN(orc_round)(u, n, c->storage);
  }
}

/* Fluid::update (src/fluid.cu:770-795).  `scratch` holds 3*W*H REAL values
 * for the advection outputs.  Writes the pressure range when pressure is
 * enabled (thrust::reduce min/max, src/fluid.cu:778-787). */
void FN(orc_update)(const osy_config* c, const orc_grid* g, REAL* u, REAL* v,
                    REAL* smoke, REAL* p, REAL* scratch, REAL* prange,
                    const osy_source* src, REAL d_t, int nthreads) {
  const long n = (long)g->W * g->H;
  /* the state as the storage type holds it, then the forces' stores */
  FN(orc_round)(u, n, c->storage);
  FN(orc_round)(v, n, c->storage);
  FN(orc_round)(smoke, n, c->storage);
  FN(orc_forces)(c, g, u, v, smoke, p, src, d_t, nthreads);
  FN(orc_round)(u, n, c->storage);
  FN(orc_round)(v, n, c->storage);
  FN(orc_round)(smoke, n, c->storage);
  if (c->viscosity != 0) FN(orc_diffusion)(c, g, u, scratch, d_t, c->projection_n, nthreads);
  FN(orc_projection)(c, g, u, v, p, d_t, c->projection_n, nthreads);
  if (c->enable_pressure && prange) {
    REAL lo = (REAL)INFINITY, hi = (REAL)-INFINITY;
    for (long k = 0; k < n; k++) {
      lo = ORC_FMIN(lo, p[k]);
      hi = ORC_FMAX(hi, p[k]);
    }
    prange[0] = lo;
    prange[1] = hi;
  }
  FN(orc_extrapolation)(g, u, v);
  REAL* u2 = scratch;
  REAL* v2 = scratch + n;
  FN(orc_advect_velocity)(g, u, v, u2, v2, d_t, nthreads);
  FN(orc_round)(u2, n, c->storage);
  FN(orc_round)(v2, n, c->storage);
  memcpy(u, u2, sizeof(REAL) * n);
  memcpy(v, v2, sizeof(REAL) * n);
  if (c->enable_smoke && c->wind_smoke != 0) {
    REAL* s2 = scratch + 2 * n;
    FN(orc_advect_smoke)(g, u, v, smoke, s2, d_t, nthreads);
    FN(orc_round)(s2, n, c->storage);
    memcpy(smoke, s2, sizeof(REAL) * n);
    if (c->enable_decay) {
      FN(orc_decay)(g, smoke, (REAL)c->decay_rate, d_t, nthreads);
      FN(orc_round)(smoke, n, c->storage);
    }
  }
}

/* Fluid::get_divergence (src/fluid.cu:192-201) over interior fluid cells,
 * accumulated in double: out[0] = sqrt(sum div^2), out[1] = max |div|. */
void FN(orc_divergence_norm)(const orc_grid* g, const REAL* u, const REAL* v,
                             double* out) {
  double ss = 0.0, mx = 0.0;
  for (int j = 1; j <= g->H - 2; j++) {
    for (int i = 1; i <= g->W - 2; i++) {
      if (g->solid[orc_idx(g, i, j)]) continue;
      const double d = (((double)u[orc_idx(g, i + 1, j)] -
                         (double)u[orc_idx(g, i, j)]) +
                        (double)v[orc_idx(g, i, j + 1)]) -
                       (double)v[orc_idx(g, i, j)];
      ss += d * d;
      if (fabs(d) > mx) mx = fabs(d);
    }
  }
  out[0] = sqrt(ss);
  out[1] = mx;
}

#undef FN
