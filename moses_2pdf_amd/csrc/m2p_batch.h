// m2p_batch.h -- batched ensembles: many scenarios of uneven grids stepped by one set of
// launches per time step (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include "m2p_kernels.h"

namespace m2p {

// One scenario of a batch as the batched kernels see it: its grid, physics, buffers and
// halting state, and where its blocks start in each batched launch.
struct Member {
  Grid g;
  Phys ph;
  DevParams dp;
  int mode, nloc, nsamp, nbx;   // nbx: CFL box blocks per row
  double phiS0;
  void* U0;
  void* U1;
  void* ONE;
  const void* MT;
  const double* M;
  unsigned char* LIM;
  DevState* st;
  const double* tf;
  const double* itime;
  const double* table;
  const double* loc;
  const int* dire;
  DryMap dry;
  int ghost0, box0, ring0, stage0;   // first block of this member in each launch
  int box_n, ring_n, stage_n;        // and how many
};

// Launch description of a batch: the member table and, per launch, the member of each
// block (blocks of one member are contiguous).
struct BatchDev {
  const Member* mem;
  const int* mapGhost;
  const int* mapBox;
  const int* mapRing;
  const int* mapStage;
  double* part;   // per-block partial maxima: CFL box [0, nBox), ring [nBox, +nRing),
                  // corrector tiles [nBox + nRing, +nStage)
  int nmem, nGhost, nBox, nRing, nStage;
  int remap;   // stage blocks dealt to the 8 XCDs in contiguous bands
  int cfl_blocks;   // most box blocks of the CFL launch; each strides over the boxes
  int vmax;    // some member has a stop speed: reduce corrector speeds after each step
  // dry-tile skipping (some member has it on): the stage plans (8 bands whatever remap
  // says) and their PLAN_COUNTERS counters, zeroed by block 0 of the dt kernel
  int plan_on;
  // the planned stage launch has one block per list slot for the first plan_cap slots of
  // each list (dry blocks past a list's end cost a block launch each, which for a large
  // batch's sparse lists dominated the stage); kb_stage_more takes any entries past them
  int plan_cap;
  // every member passed the uniform-metric check: the stage and CFL kernels form the
  // inverse Jacobian instead of loading it (FP64: the same values bit for bit; FP32: in
  // float from the rounded svec and cvalue)
  int uni;
  TilePlan plan[2];
  int* plan_cnt;
};

// Tile shapes and occupancy caps with a compiled batched stage kernel, per scalar type.
struct StageVariant {
  int ty, w;
};
int stage_variants(int prec, const StageVariant** out);

constexpr int GHOST_LINES = 128;   // ghost-fill lines per block
constexpr int BATCH_CFL_BLOCKS = 4096;   // default of BatchDev::cfl_blocks
constexpr int BATCH_PLAN_CAP = 4096;     // default of BatchDev::plan_cap
constexpr int STAGE_MORE_BLOCKS = 512;   // blocks of kb_stage_more (a multiple of 8)
int ring_blocks(const Grid& g);
int box_blocks_x(const Grid& g);
int box_blocks_y(const Grid& g);

// Enqueues the plan (with dry-tile skipping) and the stage kernel of stage io for every
// member; false when no kernel is compiled for (ty, w).
template <typename T>
bool launch_batch_stage(const BatchDev& B, int ty, int w, int io, hipStream_t s);

// Enqueues `nsteps` time steps of every member (halted members return at once).
template <typename T>
hipError_t launch_batch_steps(const BatchDev& B, int ty, int w, bool any_inflow, int nsteps,
                              hipStream_t s);

}  // namespace m2p
