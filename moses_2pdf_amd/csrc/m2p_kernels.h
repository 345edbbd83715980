// m2p_kernels.h -- device state blocks and kernel launchers of the adNOC_2Pb engine.
#pragma once

#include <hip/hip_runtime.h>

#include "m2p_numerics.h"

namespace m2p {

// Loop scalars kept in device memory so that a step needs no host round trip (the
// reference copies dt and TotalTime to the host every step, L/R:978-979).
struct DevState {
  double time;        // TotalTime (scaled)
  double dtval;       // dt of the step after clamping and output clipping
  double dt_pred;     // dtval/2 (CFL4Kernel, io = 0)
  double dt_corr;     // dtval   (io = 1)
  double dmax;        // float-rounded max wave speed
  double step_start;  // TotalTime at the start of the last step
  double t_end;       // > 0: the step that reaches it is shortened to end on it, then the
                      //      scenario halts
  double vstop2;      // > 0: halt once the squared max flow speed falls below it (after
                      //      it has been reached once)
  int iter;           // next output index (L/R:390)
  int schreiben;      // output due after this step
  int nstop;
  int inflowCount;
  int nt;             // steps done
  int halt_at;        // step count after which every kernel of this scenario returns at once
  int skip;           // halted (set by the dt kernel for the kernels after it in a step)
  double v2;          // squared max flow speed of the corrector of step v2_nt
  int moved;          // the flow speed has reached the stop threshold at least once
  int v2_nt;          // step whose corrector v2 comes from (-1: none)
  // Halting rules of one m2p_run_to call, kept apart from the caller's (t_end, halt_at),
  // which stay as they were when the call returns.
  double call_t_end;  // > 0: end time of the call in progress
  int call_halt;      // step count at which the call's budget is spent (INT_MAX: none)
  // A CFL pass met a non-finite wave speed on a face (the max for dt ignores it, as the
  // reference's fmaxf does): the state after nonfinite_nt steps holds a NaN or an Inf.
  int nonfinite;
  int nonfinite_nt;   // -1: never
  // dt/dx and dt/dy of the predictor and the corrector (Flux18Kernel's dt/dx, dt/dy,
  // L/K:2641-2652), formed once by the dt kernel in FP64 instead of in every thread
  double dtdx_pred, dtdy_pred, dtdx_corr, dtdy_corr;
};

// Dry-tile skipping.  A stage or CFL tile whose every input cell is zero in every
// component produces exactly +0 states, +0 limited momenta, a zero flow speed and the
// 1e-10 wave-speed floor, so it can be skipped without changing a bit of the results.
// Interior cells are grouped in row segments of 63 cells (the stage tile's width):
// segment s of row r holds columns [MD + 63 s, MD + 63 s + 62].  z0[(r - MD) * nseg + s]
// is 0 only if every component of every cell of that segment is zero in U0 (u^n after the
// inflow), z1 likewise for U1 (the predictor's output); a nonzero flag promises nothing.
// Every stage tile that is computed or cleared rewrites the flags of its rows (a skipped
// tile's rows stay zero, as its flags say), so after one step they are exact; any other
// write to the state sets them all to 1.  nseg == 0 switches the skipping off.
struct DryMap {
  unsigned char* z0;
  unsigned char* z1;
  int nseg;
  // Ghost-fill flags of U0 and U1, 7 * (nyd + nxd) each: g[k * (nyd + nxd) + line] is 0 only
  // if the last fill of that buffer wrote +0 to every ghost cell of plane k on that line
  // (ghost_line_skip); set to 1 with the segment flags.
  unsigned char* g0;
  unsigned char* g1;
};
inline int dry_segments(const Grid& g) { return (g.nx + 62) / 63; }
// The dry test infers the ghost cells from the interior cells two deep along the boundary,
// so it needs at least two interior rows and columns.
inline bool dry_capable(const Grid& g) { return g.nx >= 2 && g.ny >= 2; }
inline long long dry_flag_count(const Grid& g) { return (long long)g.ny * dry_segments(g); }
inline long long ghost_flag_count(const Grid& g) { return 7LL * (g.nyd + g.nxd); }
// Sets every flag of both maps to 1 (nothing may be skipped in the next step).
void launch_dry_fill(const Grid& g, const DryMap& dm, hipStream_t s);
// Sets both ghost-fill flag arrays to 1 (the next fills write every line).
void launch_ghost_flags_fill(const Grid& g, const DryMap& dm, hipStream_t s);
// Sets both maps from the state itself, z0 from U0 and z1 from U1, for dm.nseg segments per
// row: a segment's flag is 0 when all 7 components of its cells are +0 bit for bit, else 1
// (a -0 or a NaN counts as wet, so such cells are computed as without skipping).  phi is
// tested too: a skipped tile keeps its phi, where a computed dry cell gets +0.  So an
// initial landslide state (phi = phiS0 on every cell) has no dry segment, while a state
// that has been stepped has its exact flags again.
template <typename T>
void launch_dry_scan(const Grid& g, const DryMap& dm, const T* U0, const T* U1, hipStream_t s);

// A list of device-to-device copies done by one launch (the batch autotuner's snapshots of
// every member).  `items` is in device memory; copies whose ends are both 16-byte aligned
// move 16 bytes per lane, the others one byte.
struct CopyItem {
  const unsigned char* src;
  unsigned char* dst;
  size_t n;
};
constexpr int COPY_BLOCKS_PER_ITEM = 64;
void launch_copy_list(const CopyItem* items, int nitems, hipStream_t s);
// Copies n scenario states (*src[k] -> dst[k]) in one launch; src is in device memory.
void launch_gather_states(const DevState* const* src, DevState* dst, int n, hipStream_t s);

// Tile plan of one stage launch under dry-tile skipping.  A plan kernel tests every tile
// (PLAN_LANES = 2 lanes per tile) and lists, in one of 8 lists, the tiles the stage must
// visit: those to compute and, for the predictor, dry tiles whose U1 and ONE must be
// cleared (entry | PLAN_ZERO).  Block b of the stage runs on XCD b % 8 and takes entry
// b / 8 of list b % 8; blocks past the list's end return at once, so a dry tile costs one
// block launch, and the hardware's block dispatch balances the wet ones (a bounded grid
// striding over the lists was measured slower: mixed tile costs left it unbalanced).
// Tiles go to the lists in chunks of PLAN_CHUNK consecutive tiles (tile t to list
// (t / PLAN_CHUNK) % 8): a chunk's tiles share halos in one XCD's L2, and a compact wet
// region still spreads over all 8 XCDs (with contiguous bands of tiles it fell on one or
// two, and a batch's large members each on one).  The list lengths are zeroed by the dt
// kernel of the step.
constexpr int PLAN_ZERO = 1 << 30;
constexpr int PLAN_COUNTERS = 16;   // the 8 list lengths of each of the two stages
#ifndef M2P_PLAN_CHUNK
#define M2P_PLAN_CHUNK 32
#endif
constexpr int PLAN_CHUNK = M2P_PLAN_CHUNK;

// Waves per SIMD of the CFL kernels: 128 VGPRs in FP64 (no spills), 64 in FP32 (12-28
// bytes of spills, cheaper than the lost occupancy).
template <typename T>
constexpr int CFL_WAVES = sizeof(T) == 4 ? 8 : 4;
struct TilePlan {
  int* list;   // 8 * per entries
  int* cnt;    // 8 list lengths
  int per;     // list capacity, plan_per(tiles)
};
// The most tiles one list can receive: every chunk of it full.
inline int plan_per(int tiles) {
  const int span = 8 * PLAN_CHUNK;
  return PLAN_CHUNK * ((tiles + span - 1) / span > 0 ? (tiles + span - 1) / span : 1);
}
// Blocks of a planned stage launch: one per list slot.
inline int plan_blocks(int tiles) { return 8 * plan_per(tiles); }
__host__ __device__ __forceinline__ int plan_list_of(int t) { return (t / PLAN_CHUNK) & 7; }

// A scenario is halted for the kernels that run before the dt kernel of a step
// (inflow, ghost fill, CFL) once it has done halt_at (or call_halt) steps, and for those
// after it once the dt kernel has set `skip`.
__device__ __forceinline__ bool halted_pre(const DevState* st) {
  return st->nt >= st->halt_at || st->nt >= st->call_halt;
}
__device__ __forceinline__ bool halted_post(const DevState* st) { return st->skip != 0; }

struct DevParams {
  double cfl, dt_min, dt_max;
  double dx, dy;   // cell size (scaled), for the stages' dt/dx and dt/dy
  int ntf;     // entries of tf (0: no output clipping)
  int nsamp;   // hydrograph samples (0: landslide)
};

// The 18 face-average planes (MC_FX..MC_COUNT of the metric layout) into F, from the cell
// planes of M; F is M + MC_FX * P in FP64, the float copy's face planes in FP32.
template <typename TO>
void launch_face_metric(const Grid& g, const double* M, TO* F, hipStream_t s);
// The uniform-metric check: *ok (device int, set to 1 by the caller) is cleared unless the
// stored FP64 metric M equals, bit for bit, what FM_UNIFORM forms (face averages on faces
// [1, N-3]^2 in both directions, the inverse Jacobian of cells [3, N-4]^2), so that the
// step kernels may form it instead of loading it.
void launch_metric_uniform_check(const Grid& g, const double* M, int* ok, hipStream_t s);
// The cell-level form of that check, for the FP32 variant (which keeps no FP64 face planes):
// *ok is cleared unless, on cells [1, N-2]^2, the coordinate-derivative planes are exactly
// 1 and +0 and the inverse-Jacobian planes equal invj_uniform of svec and cvalue bit for bit.
void launch_metric_uniform_cells_check(const Grid& g, const double* M, int* ok, hipStream_t s);
void launch_metric_setup(const Grid& g, int mode, const double* topo, const double* depth,
                         const double* phicell, double phiS0, double* M, double* U,
                         hipStream_t s);
template <typename T>
void launch_convert(const double* a, T* b, long long n, hipStream_t s);
template <typename T>
void launch_inflow(const Grid& g, T* U, const double* M, const double* table, const double* loc,
                   const int* dire, int nloc, int nsamp, double phiS0, const DevState* st,
                   unsigned char* LIM, const DryMap& dm, hipStream_t s);
// Ghost fill of Tg (post 0: U0 before the CFL pass; post 1: U1 before the corrector).  With
// dry-tile skipping, a landslide's line whose source cells are all in dry segments (z of
// Tg) and whose ghosts the last fill left at +0 (gf) is left alone: the fill would write
// the same +0.
template <typename T>
void launch_ghost(const Grid& g, int mode, T* Tg, const T* prev, const DevState* st, int post,
                  const unsigned char* z, unsigned char* gf, int nseg, hipStream_t s);
// A landslide's ghost cells of the momenta (planes 1, 2, 4, 5) hold 0 at all times: every
// ghost fill would write 0 there (Boundary7/8), nothing else writes a landslide's ghost
// cells, and a host-side write of the state clears them with this kernel.  So the ghost
// fills leave those planes alone.  (Inflow strips write momenta into ghost cells, so an
// inflow fill writes all 7 planes.)
template <typename T>
void launch_zero_ghost_momenta(const Grid& g, T* U, hipStream_t s);
// Per-block partial maxima of one step (CFL box blocks, CFL ring blocks, corrector tiles),
// reduced by the dt kernel: partial_count() doubles per scenario.
size_t partial_count(const Grid& g, int prec);
// uni (FP64 only): the grid passed the uniform-metric check (launch_metric_uniform_check).
template <typename T>
void launch_cfl(const Grid& g, const Phys& ph, const T* U, const T* M, DevState* st, double* part,
                const DryMap& dm, bool uni, hipStream_t s);
// plan_cnt: the PLAN_COUNTERS plan counters of the step's two stages (zeroed here), or null.
void launch_dt(const DevParams& dp, DevState* st, const double* tf, const double* itime,
               const Grid& g, int prec, const double* part, int* plan_cnt, hipStream_t s);
// Offset of the corrector tiles' partials in a scenario's partial array (after the CFL ones).
size_t stage_partials_offset(const Grid& g);
// The most tiles of a single-scenario stage launch (planned launches, and the plan kernel
// before them, take 3-row tiles; dense ones plain_ty's height).
int plain_stage_tiles(const Grid& g, int prec);
template <typename T>
void launch_plan(int io, const Grid& g, const T* Uin, const DryMap& dm, const TilePlan& pl,
                 double* part, const DevState* st, hipStream_t s);
// After a corrector (planned: with dry-tile skipping): the squared max flow speed of its
// tiles into DevState (stop speed).
void launch_vmax(DevState* st, const Grid& g, int prec, bool planned, const double* part,
                 hipStream_t s);
template <typename T>
void launch_stage(int io, const Grid& g, const Phys& ph, const T* Uin, const T* Uzero, T* Uout,
                  T* ONE, const T* M, unsigned char* LIM, DevState* st, double* part,
                  const DryMap& dm, const TilePlan& pl, bool uni, hipStream_t s);
// Compensated mass totals: each of mass_blocks() blocks writes 4 partials (solid hi, lo,
// fluid hi, lo) of sum u0*Det and u3*Det over its share of the interior cells into
// out[4 b .. 4 b + 3], every slot every call.  det is the Detmin plane.
int mass_blocks();
template <typename T>
void launch_mass(const Grid& g, const T* U, const double* det, double* out, hipStream_t s);
// The double-double sum of the nb blocks' partials in block order, out[0] = solid,
// out[1] = fluid: the same operations, in the same order, as m2p_mass's host sum.
void launch_mass_sum(const double* part, int nb, double* out, hipStream_t s);
hipError_t probe_kernels();

}  // namespace m2p
