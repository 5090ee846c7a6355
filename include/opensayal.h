/*
 * opensayal.h - C ABI of the MI355X (gfx950) OpenSayal per-step fluid solver.
 *
 * The reference exposes the solver as the C++ class `Fluid`
 * (inc/fluid.cuh:15-116): construction from a parsed `Config`
 * (src/fluid.cu:41-71, src/config_parser.cpp:15-184), one step per call to
 * `Fluid::update(Source, float d_t)` (src/fluid.cu:770-795), and public
 * device pointers plus `min_pressure`/`max_pressure` host members
 * (inc/fluid.cuh:61-72).  This header replaces that class with an opaque
 * handle and plain C entry points: no C++ types, no exceptions, every call
 * returns an `osy_status` (0 = ok, negative = error) and `osy_last_error()`
 * gives the message of the calling thread's last failure.
 *
 * Field layout at this boundary is the reference's: a W x H array indexed by
 * indx(i, j) = (H - 1 - j) * W + i (src/fluid.cu:163-165), memory row 0 being
 * the top row.  u(i, j) sits on the left face of cell (i, j), v(i, j) on its
 * bottom face (src/fluid.cu:552-558).
 *
 * Threading: a handle is not thread-safe; distinct handles may be used from
 * distinct threads.  All handles of one process share the current HIP device.
 */
#ifndef OPENSAYAL_H
#define OPENSAYAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSY_ABI_VERSION 7

typedef enum osy_status {
  OSY_OK = 0,
  OSY_E_INVALID = -1,     /* bad argument or configuration value            */
  OSY_E_UNSUPPORTED = -2, /* a configuration the solver does not implement  */
  OSY_E_HIP = -3,         /* a HIP runtime call failed                      */
  OSY_E_NOMEM = -4,       /* device or host allocation failed               */
  OSY_E_CONFIG = -5,      /* configuration file unreadable or malformed     */
  OSY_E_BUDGET = -6,      /* request does not fit the device-memory budget  */
  OSY_E_STATE = -7        /* call not valid in the object's current state   */
} osy_status;

/* Storage type of u, v and smoke.  Pressure and all arithmetic stay FP32. */
typedef enum osy_storage {
  OSY_STORAGE_F32 = 0,
  OSY_STORAGE_F16 = 1,
  OSY_STORAGE_BF16 = 2
} osy_storage;

/* Field identifiers for get/set.  Element type in parentheses. */
typedef enum osy_field {
  OSY_FIELD_VEL_X = 0,    /* u, d_vel_x (float)                            */
  OSY_FIELD_VEL_Y = 1,    /* v, d_vel_y (float)                            */
  OSY_FIELD_SMOKE = 2,    /* d_smoke (float)                               */
  OSY_FIELD_PRESSURE = 3, /* d_pressure (float)                            */
  OSY_FIELD_IS_SOLID = 4, /* d_is_solid (int32, 0/1); read-only            */
  OSY_FIELD_TOTAL_S = 5   /* d_total_s (int32, 0..4); read-only            */
} osy_field;

/*
 * Solver configuration: the solver-relevant part of the reference's `Config`
 * (inc/config_parser.hpp:9-116).  Names follow the OpenSayal.conf.json keys.
 * Keys that only affect rendering, the CUDA launch shape or OpenMP are
 * accepted by the JSON reader and kept here for completeness; they do not
 * change the solver's results.
 */
typedef struct osy_config {
  /* sim.* */
  int32_t width;              /* sim.width, default 1920                      */
  int32_t height;             /* sim.height, default 1080                     */
  float cell_size;            /* sim.cell_size, default 1 (truncated to int)  */
  int32_t cell_pixel_size;    /* sim.cell_pixel_size (render only)            */
  int32_t enable_drain;       /* sim.enable_drain, default 1                  */
  int32_t enable_pressure;    /* sim.enable_pressure, default 0               */
  int32_t enable_smoke;       /* sim.enable_smoke, default 1                  */
  int32_t enable_interactive; /* sim.enable_interactive (mouse; no effect)    */
  int32_t projection_n;       /* sim.projection.n, default 50                 */
  float projection_o;         /* sim.projection.o, default 1.9                */
  int32_t pipe_height;        /* sim.wind_tunnel.pipe_height, default H/4     */
  int32_t pipe_length;        /* not configurable in the reference: always 0  */
  int32_t smoke_length;       /* sim.wind_tunnel.smoke_length, default 1      */
  int32_t smoke_height;       /* sim.wind_tunnel.smoke_height, default H/4    */
  int32_t smoke_count;        /* sim.wind_tunnel.smoke_count, default 1       */
  float wind_speed;           /* sim.wind_tunnel.speed, default 0             */
  float wind_smoke;           /* sim.wind_tunnel.smoke, default 1             */
  float g;                    /* sim.physics.g, default 0                     */
  float d_t;                  /* sim.time.d_t, default 0.05                   */
  int32_t enable_real_time;   /* sim.time.enable_real_time (headless: no-op)  */
  float real_time_multiplier; /* sim.time.real_time_multiplier                */
  int32_t enable_decay;       /* sim.smoke.enable_decay, default 0            */
  float decay_rate;           /* sim.smoke.decay_rate, default 0.05           */
  int32_t obstacle_enable;    /* sim.obstacle.enable, default 1               */
  int32_t obstacle_center_x;  /* sim.obstacle.center_x, default W/2           */
  int32_t obstacle_center_y;  /* sim.obstacle.center_y, default H/2           */
  float obstacle_radius;      /* sim.obstacle.radius, default min(W,H)/30     */
  /* fluid.* */
  float density;              /* fluid.density, default 1                     */
  float drag_coeff;           /* fluid.drag_coeff, default 0                  */
  float viscosity;            /* fluid.viscosity, default 0.001; must be 0    */
  /* thread.* (accepted, no effect) */
  int32_t block_size_x;       /* thread.cuda.block_size_x, default 64         */
  int32_t block_size_y;       /* thread.cuda.block_size_y, default 1          */
  int32_t omp_thread_count;   /* thread.openMP.thread_count                   */
  /* this implementation's own key "sim.storage": "f32" | "f16" | "bf16"    */
  int32_t storage;            /* osy_storage, default OSY_STORAGE_F32         */
} osy_config;

/* Mouse source of the reference (inc/fluid.cuh:8-13).  Headless callers pass
 * NULL or active = 0. */
typedef struct osy_source {
  int32_t active;
  float smoke;
  float velocity;
  int32_t position_x;
  int32_t position_y;
} osy_source;

/* Geometry and device storage of a created solver. */
typedef struct osy_fluid_info {
  int32_t width, height, cell_size, projection_n;
  int32_t pitch;        /* elements per device row                        */
  int32_t storage;      /* osy_storage of u, v, smoke                      */
  int32_t projection_kernel; /* 0: one launch per colour phase;
                                1: LDS-tiled, several sweeps per launch;
                                2: register tiles, several sweeps per launch;
                                3: register tiles resident for all sweeps,
                                   one launch                              */
  int32_t sweeps_per_launch; /* sweeps one projection launch runs          */
  size_t device_bytes;  /* device memory held by this handle               */
} osy_fluid_info;

/* Launch shape of the projection, this implementation's own knob (the
 * reference's thread.cuda.block_size_x/y has no counterpart here).  Results
 * are bit-identical for every valid shape; only speed changes. */
typedef struct osy_tuning {
  int32_t projection_kernel; /* 0: one launch per colour phase; 1: tiled;
                                2: register tiles; 3: register tiles
                                resident for all sweeps (one launch; every
                                tile's workgroup must fit on the device at
                                once, and both tile sides be >= 2 x k)     */
  int32_t tile_width;        /* tiled: cells a workgroup owns along i (even) */
  int32_t tile_height;       /* tiled: cells a workgroup owns along j       */
  int32_t sweeps_per_launch; /* tiled: sweeps per launch (halo = 2x this);
                                3: sweeps between halo exchanges           */
} osy_tuning;

typedef struct osy_fluid osy_fluid;

/* Projection launch shape of a batch: register-tile rows per wavefront (8 or
 * 12; the tile is (128 - 4k) x (8 * rows - 4k) cells with a 2k-cell halo),
 * sweeps per tile k, and occupancy: 1 = the compiler's register allocation
 * (rows 8 only), 2 = registers capped for two workgroups per CU; 0 means 1. */
typedef struct osy_batch_tuning {
  int32_t rows;
  int32_t sweeps_per_tile;
  int32_t occupancy;
} osy_batch_tuning;

typedef struct osy_batch_info {
  int32_t count;         /* member handles                                  */
  int32_t groups;        /* persistent launches per step (storage x pressure) */
  int64_t work_items;    /* (stage, tile) items per step, all groups        */
  int64_t cells;         /* sum of W*H over members                         */
  uint64_t budget_bytes; /* 0 = unlimited                                   */
  uint64_t used_bytes;   /* device bytes of the member handles              */
} osy_batch_info;

typedef struct osy_batch osy_batch;

/* One work item of a batch's last step, as osy_batch_item_times reports it:
 * the item's member (index in the order handles were added), stage kind
 * (0 forces, 1 projection launch, 2 velocity advection, 3 smoke advection),
 * stage index within the member's step and tile; the wall-clock ticks at
 * which a workgroup claimed it, found its dependencies complete and finished
 * it; and that workgroup's hardware id (__smid(): XCD, shader engine, CU). */
typedef struct osy_batch_item_time {
  int32_t member;
  int32_t kind;
  int32_t stage;
  int32_t tile;
  uint64_t claim;
  uint64_t ready;
  uint64_t done;
  uint64_t hw_id;
} osy_batch_item_time;

/* ---- configuration: replaces ConfigParser (src/config_parser.cpp:11-184) */

/* Fill `cfg` with the reference's defaults (src/config_parser.cpp:21-181),
 * with viscosity forced to 0 (the reference's diffusion is not supported). */
int osy_config_defaults(osy_config* cfg);
/* ConfigParser(path).parse(): dotted or nested keys, get_or semantics
 * (inc/config_parser.hpp:131-150). */
int osy_config_parse_file(const char* path, osy_config* cfg);
/* Same, from JSON text in memory. */
int osy_config_parse_string(const char* json_text, osy_config* cfg);

/* ---- the solver: replaces class Fluid (inc/fluid.cuh:15-116) ---------- */

/* Fluid::Fluid(Config) (src/fluid.cu:41-161): allocates device state, builds
 * the solid mask and neighbour counts, zeroes u, v, smoke and pressure. */
int osy_fluid_create(const osy_config* cfg, osy_fluid** out);
/* Fluid(ConfigParser(path).parse()). */
int osy_fluid_create_from_json(const char* path, osy_fluid** out);
/* Fluid::~Fluid (src/fluid.cu:73-84).  NULL is accepted. */
int osy_fluid_destroy(osy_fluid* fluid);

/* Fluid::update(Source, d_t) (src/fluid.cu:770-795): one step, synchronous.
 * `src` may be NULL (inactive source). */
int osy_fluid_update(osy_fluid* fluid, const osy_source* src, float d_t);
/* `steps` calls of Fluid::update with an inactive source, replayed from a
 * captured HIP graph; synchronous. */
int osy_fluid_step(osy_fluid* fluid, int32_t steps, float d_t);
/* As osy_fluid_step but enqueued on `hip_stream` (a hipStream_t, NULL = the
 * handle's own stream) and returning without waiting. */
int osy_fluid_step_async(osy_fluid* fluid, int32_t steps, float d_t,
                         void* hip_stream);
/* Wait for all work enqueued on the handle's stream. */
int osy_fluid_synchronize(osy_fluid* fluid);

/* Fluid::apply_projection (src/fluid.cu:282-295) alone: `sweeps` red-black
 * sweeps on the current u, v (and pressure accumulation when enabled; the
 * pressure is not zeroed first).  Enqueued on `hip_stream` (NULL = the
 * handle's stream) without waiting.  For timing and testing the dominant
 * phase by itself. */
int osy_fluid_project(osy_fluid* fluid, int32_t sweeps, void* hip_stream);

/* Copy a field to / from host memory in the reference's indx order.
 * `bytes` must equal W*H*4.  IS_SOLID and TOTAL_S are read-only. */
int osy_fluid_get_field(osy_fluid* fluid, int32_t field, void* dst,
                        size_t bytes);
int osy_fluid_set_field(osy_fluid* fluid, int32_t field, const void* src,
                        size_t bytes);

/* Fluid::min_pressure / max_pressure after the last step (thrust reduce at
 * src/fluid.cu:778-787).  Both are 0 when pressure is disabled. */
int osy_fluid_pressure_range(osy_fluid* fluid, float* min_pressure,
                             float* max_pressure);
/* L2 and max-abs norms, accumulated in FP64, of the divergence
 * ((u(i+1,j) - u(i,j)) + v(i,j+1)) - v(i,j) over interior fluid cells, the
 * formula of Fluid::get_divergence (src/fluid.cu:192-201). */
int osy_fluid_divergence_norm(osy_fluid* fluid, double* l2, double* linf);

int osy_fluid_get_info(const osy_fluid* fluid, osy_fluid_info* info);

/* ---- what the reference's GraphicsHandler reads from a Fluid each frame --- */

/* The frame of update_fluid_pixels_kernel (src/graphics_handler.cu:258-285)
 * and its copy to host (GraphicsHandler::update_fluid_pixels, :287-300): one
 * RGBA8888 pixel per cell, r << 24 | g << 16 | b << 8 | a (map_rgba,
 * src/helper.cu:45-47), in the reference's indx order (row 0 = the top row
 * of the screen).  Solid cells are (80, 80, 80, 255); with pressure and smoke
 * on the hue is (1 - norm) * 120 degrees, norm the pressure over the last
 * update's min / max (hsv_to_rgb, src/helper.cu:3-43), value the smoke; with
 * pressure only the value is 1; with smoke only (255, c, c, 255),
 * c = 255 - 255 * smoke; with neither the pixel is 0.  Float -> byte
 * conversions truncate and saturate, NaN giving 0.  `bytes` must equal
 * W*H*4.  Synchronous. */
int osy_fluid_render_rgba(osy_fluid* fluid, uint32_t* dst, size_t bytes);

/* Path lines of update_traces_kernel (src/graphics_handler.cu:384-396) with
 * Fluid::trace (src/fluid.cu:16-36) and the copies to host (:406-412): seeds
 * at (ti * distance, tj * distance), ti < W / distance, tj < H / distance;
 * seed (ti, tj) owns entries [o, o + length) of traces_x and traces_y with
 * o = ((H/distance - 1 - tj) * (W/distance) + ti) * length (indx_traces,
 * src/graphics_handler.cu:7-11).  Point k is the cell centre advanced k times
 * by the general velocity times d_t, stored as (round(x), H - 1 - round(y)),
 * round half away from zero, saturating, NaN giving 0.  A seed on a solid
 * cell gets -1 in all its entries (the reference sets entry 0, which its
 * drawing loop tests).  `bytes` is the size of each of traces_x and
 * traces_y and must equal (W/distance) * (H/distance) * length * 4;
 * distance >= 1, 1 <= length <= 65536.  Synchronous. */
int osy_fluid_traces(osy_fluid* fluid, int32_t distance, int32_t length, float d_t,
                     int32_t* traces_x, int32_t* traces_y, size_t bytes);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc after set-up):
 * osy_fluid_traces with the same arguments, checks and results, except that
 * the device staging of the two arrays is a fixed library-owned device
 * array (2 MiB per GPU, held until the process ends) instead of a buffer
 * osy_fluid_traces allocates and grows per handle.  Each of traces_x and
 * traces_y may hold at most 262144 entries ((W/distance)*(H/distance)*
 * length); larger calls return OSY_E_INVALID.  All calls on one GPU share
 * the staging array: calls that overlap on one GPU (from several threads)
 * race.  Synchronous. */
int osy_fluid_traces_noalloc(osy_fluid* fluid, int32_t distance, int32_t length, float d_t,
                             int32_t* traces_x, int32_t* traces_y, size_t bytes);

/* Select the projection launch shape (OSY_E_INVALID if the shape does not
 * fit the kernel's limits).  Captured step graphs are rebuilt on next use. */
int osy_fluid_set_tuning(osy_fluid* fluid, const osy_tuning* tuning);
int osy_fluid_get_tuning(const osy_fluid* fluid, osy_tuning* tuning);

/* Bound of one neighbour wait of the resident projection (projection_kernel
 * 3), in microseconds of the device's wall clock (default 2,000,000).  A
 * wait that exceeds it sets the handle's error word and the launch drains
 * without waiting; the next step, synchronize or field read then returns
 * OSY_E_STATE.  No counterpart in the reference: for tests of that path (a
 * limit of 0) and for callers that want a stalled step to fail sooner.
 * OSY_E_INVALID for a negative limit.  Captured step graphs are rebuilt on
 * next use. */
int osy_fluid_set_wait_limit(osy_fluid* fluid, int64_t microseconds);

/* Autotuner (this implementation's own; no counterpart in the reference):
 * times every supported projection launch shape on this grid, projecting a
 * copy of the handle's current u and v (the handle's fields are restored
 * afterwards), keeps the fastest and writes it to `chosen` (may be NULL)
 * with its time per projection in `ms` (may be NULL).  The choice is cached
 * per (width, height, n, pressure, storage) for the process, so later
 * handles of the same shape can adopt it with osy_fluid_autotune_cached. */
int osy_fluid_autotune(osy_fluid* fluid, osy_tuning* chosen, double* ms);
/* Adopt a cached autotuner choice for this handle's shape.  Returns 1 when
 * one was found and applied, 0 when none is cached, negative on error. */
int osy_fluid_autotune_cached(osy_fluid* fluid);

/* Device bytes osy_fluid_create would allocate for `cfg` (for admission
 * under a memory budget before creating). */
int osy_fluid_device_bytes(const osy_config* cfg, uint64_t* bytes);

/* ---- batches (no counterpart in the reference, which steps one Fluid) ----
 * A batch steps many independent handles of any sizes together: one
 * persistent kernel launch per step and storage group, replayed from a
 * captured HIP graph.  Each member's fields after k batch steps are
 * bit-identical to k osy_fluid_step calls on that handle alone.  While a
 * handle is a member, osy_fluid_update/step/step_async/project/autotune and
 * set_tuning return OSY_E_STATE; get_field, set_field, pressure_range and
 * divergence_norm first wait for the batch's queued steps.  Each member
 * steps with its own configuration's time.d_t. */

/* `budget_bytes` bounds the sum of the members' device bytes (0 = none).
 * The batch starts with the shape rows 12, sweeps_per_tile 10, occupancy 2. */
int osy_batch_create(uint64_t budget_bytes, osy_batch** out);
/* Members leave the batch (they stay valid).  NULL is accepted. */
int osy_batch_destroy(osy_batch* batch);
/* Add a handle; OSY_E_BUDGET when its device bytes exceed what is left of
 * the budget, OSY_E_STATE when it is already in a batch. */
int osy_batch_add(osy_batch* batch, osy_fluid* fluid);
int osy_batch_remove(osy_batch* batch, osy_fluid* fluid);
/* `steps` steps of every member; synchronous.  OSY_E_STATE if a
 * dependency wait inside the kernel timed out (results then invalid). */
int osy_batch_step(osy_batch* batch, int32_t steps);
/* As osy_batch_step, enqueued on `hip_stream` (NULL = the batch's stream)
 * without waiting. */
int osy_batch_step_async(osy_batch* batch, int32_t steps, void* hip_stream);
int osy_batch_synchronize(osy_batch* batch);
int osy_batch_set_tuning(osy_batch* batch, const osy_batch_tuning* tuning);
int osy_batch_get_tuning(const osy_batch* batch, osy_batch_tuning* tuning);
/* Autotuner: times each supported shape on the current members (their
 * state is saved first and restored afterwards), keeps the fastest; writes
 * it to `chosen` and its time per batch step to `ms` (either may be NULL).
 * Members that step on their own graphs (osy_batch_set_per_grid_cells) get
 * their projection shape from osy_fluid_autotune first. */
int osy_batch_autotune(osy_batch* batch, osy_batch_tuning* chosen, double* ms);
int osy_batch_get_info(const osy_batch* batch, osy_batch_info* info);
/* Members of at least `cells` cells (W*H) fill the GPU on their own: each
 * batch step runs them through their own per-grid step graphs, on the
 * batch's stream ahead of the persistent launch, instead of as its work
 * items (a lone large grid then steps as fast as alone; the results are the
 * same bits either way).  Default 2^20; 0 keeps every member in the
 * persistent launch.  No counterpart in the reference. */
int osy_batch_set_per_grid_cells(osy_batch* batch, int64_t cells);
/* Per-item timing of the batch's persistent launches (no counterpart in the
 * reference; a profiling aid).  Enabled, every launch stamps each of its
 * work items (four stores per item); the results are unchanged. */
int osy_batch_set_timing(osy_batch* batch, int32_t enable);
/* Waits for the batch, then writes the records of its last step's items
 * (every storage group's launch, in group order) to `out` (room for
 * `capacity`), their number to `count` and the wall-clock tick in
 * nanoseconds to `tick_ns`.  OSY_E_STATE when timing is off or no step ran
 * since it was enabled; OSY_E_INVALID when `capacity` is too small (`count`
 * then holds the number needed). */
int osy_batch_item_times(osy_batch* batch, osy_batch_item_time* out, int64_t capacity,
                         int64_t* count, double* tick_ns);

/* ---- streams and events of the library's HIP runtime (no counterpart in
 * the reference).  For callers that order or time the solver's work without
 * a second HIP runtime in the process (a framework that bundles its own
 * libamdhip64 would otherwise hand the library foreign stream handles).
 * Streams are non-blocking with the default priority. */
int osy_stream_create(void** stream);
int osy_stream_destroy(void* stream);
int osy_stream_synchronize(void* stream);
int osy_event_create(void** event);
int osy_event_destroy(void* event);
/* Record `event` on `stream` (NULL = the null stream). */
int osy_event_record(void* event, void* stream);
int osy_event_synchronize(void* event);
/* Milliseconds between two recorded, completed events. */
int osy_event_elapsed_ms(void* start, void* end, float* ms);
/* `reps` device-to-device copies of `bytes` bytes (a multiple of 16) with
 * the library's own copy kernel, between two fresh buffers; synchronous.
 * For scaling hardware byte counters against a known traffic. */
int osy_calibration_copy(uint64_t bytes, int32_t reps);

/* Exhaustive device checks of the arithmetic the solver uses in place of
 * the plain operations (no counterpart in the reference; the proof that the
 * results are unchanged, run on the GPU that runs the solver).  Each mode
 * visits all 2^32 binary32 patterns:
 *   0: inv_distance's reciprocal path against 1.0 / (s + 1e-6) in double
 *      rounded to float, bit for bit (src/fluid.cu:690-693);
 *   1: the refined double reciprocal of every finite nonzero value, within
 *      a relative 2^-50;
 *   2: the weights' quotient against IEEE binary32 division, on 2^32 pairs;
 *   3, 4: the projection's paired rounding to FP16 / BF16 storage (one
 *      packed conversion) against one scalar conversion per value;
 *   5, 6: the pressure sums' quotient x / d_t as a double multiply by the
 *      double nearest 1 / d_t, against IEEE division, for d_t = 0.05 and
 *      d_t = 1/60;
 *   7: the staged smoke weights' square root (the hardware root corrected
 *      from fused residuals) against the correctly rounded sqrtf, on 0 and
 *      every finite value from 2^-96 up.
 * Writes the number of inputs that fail and the smallest failing pattern
 * (0xFFFFFFFF when none); synchronous. */
int osy_arith_check(int32_t mode, uint64_t* mismatches, uint32_t* first_input);

/* Message of the calling thread's last failed call ("" if none). */
const char* osy_last_error(void);
/* OSY_ABI_VERSION of the loaded library. */
int osy_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* OPENSAYAL_H */
