// batch.hpp - tables and launch interface of the batched step kernel.
//
// A batch steps many independent simulations (of any sizes) with one
// persistent kernel launch per step and storage group.  The host flattens
// every simulation's step into a chain of stages (forces, projection launches
// of up to k sweeps, velocity advection, smoke advection); each stage is cut
// into tiles, and every (stage, tile) pair is a work item.  Items are listed
// stage level by stage level across simulations, so every item's
// dependencies precede it in the list.  Workgroups claim items in list order
// from an atomic counter and wait (bounded) for their dependencies before
// working, so the kernel cannot deadlock whatever number of workgroups is
// resident.  A projection item depends only on the tiles of the previous
// projection (or forces) stage that its region overlaps, read or written;
// an advection item, whose back-traces may read anywhere in the grid, waits
// for the whole previous stage.
#pragma once

#include "kernels.hpp"

namespace osy {

enum BatchStageKind : int {
  BK_FORCES = 0,     // forces not applied by a later load (n == 0 or smoke off)
  BK_PROJ = 1,       // one register-tile projection launch of `sweeps` sweeps
  BK_ADV_UV = 2,     // extrapolation + velocity advection
  BK_ADV_SMOKE = 3,  // smoke advection + decay
};

// Tile of the forces and advection stages in a batch: 64 x 64 cells, one
// row group per wavefront.
constexpr int kBatchAdvTX = 64, kBatchAdvTY = 64;
constexpr int kBatchThreads = 64 * kRegWaves;  // 512

// Per simulation: parameters, buffers, and the buffer parity rule.  At launch
// number e after the last rebuild, u and v sit in buffer uv0 ^ (e & flip_uv
// & 1) and smoke in sm0 ^ (e & flip_sm & 1).
struct BatchSim {
  StepParams sp;
  void* u[2];
  void* v[2];
  void* smoke[2];
  float* p;
  const uint8_t* flags;
  const uint8_t* gates;
  int uv0, sm0;
  int flip_uv, flip_sm;
};

struct BatchStage {
  int sim;
  int kind;
  int q;       // BK_PROJ: projection stage index; advections: projection stage count
  int sweeps;  // BK_PROJ
  int forces;  // BK_PROJ: apply velocity forces on load; BK_FORCES: velocity part
  int smoke;   // BK_FORCES: smoke part
  int dep;     // stage that must be complete first, -1 for none
  int dep_tiles;
  int ntx;     // tiles per row of this stage's tiling
  int tile_dep;  // 1: wait only for the item's own list of predecessor tiles
};

struct BatchArgs {
  const BatchSim* sims;
  const BatchStage* stages;
  const uint4* items;  // (stage, tile, first dependency, dependency count)
  const unsigned* deps;  // item indices the tile-level waits read in tdone
  int nitems;
  int grid;                  // workgroups launched
  unsigned long long* head;  // claim counter; each launch advances it by nitems + grid
  unsigned long long* done;  // per stage: tiles completed since the last rebuild
  unsigned long long* tdone;  // per item: launches that completed it since the last rebuild
  int* error;                // set when a dependency wait exceeds spin_ticks
  long long spin_ticks;      // bound of one dependency wait, wall-clock ticks
  int tw, th, k;             // register-tile projection shape
  // per item, when not null: the wall-clock ticks at which a workgroup
  // claimed it, found its dependencies complete and finished it, and the
  // workgroup's __smid() (XCD, shader engine, CU)
  unsigned long long* stamps;
};

// Register-tile rows per wavefront supported by the batch kernel, and its
// occupancy variants: 1 = registers as the compiler allocates them, 2 = capped
// for two workgroups per CU (rows 8 both ways, rows 12 only capped).
inline bool batch_rows_supported(int rows) { return rows == 8 || rows == 12; }
inline bool batch_occupancy_supported(int rows, int occupancy) {
  return (occupancy == 1 && rows == 8) || occupancy == 2;
}

// Occupancy-derived workgroup count for the batch kernel of this storage /
// pressure / rows combination.
hipError_t batch_grid_size(int storage, bool pressure, int rows, int occupancy, int* grid);

// Enqueue one batched step (one launch).
void enqueue_batch_step(int storage, bool pressure, int rows, int occupancy, const BatchArgs& a,
                        hipStream_t s);

// [min, max] of the pressure into f.prange (thrust::reduce in
// src/fluid.cu:778-787), for simulations stepped by a batch.
void enqueue_pressure_range(const StepParams& sp, const FieldPtrs& f, hipStream_t s);

}  // namespace osy
