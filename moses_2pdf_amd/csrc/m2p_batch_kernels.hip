// m2p_batch_kernels.hip -- the batched per-step kernels: one launch per phase of the time
// step covers every member of a batch.  A block finds its member through the launch's
// block map, then runs the same device body as the single-scenario kernels (m2p_tiles.h),
// so a member's results do not depend on the batch it is in, its position or the tile
// shape (each face and cell is computed once, by the same expressions).
// Compiled once per scalar type: M2P_BATCH_T is double or float.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "m2p_batch.h"
#include "m2p_tiles.h"

#ifndef M2P_BATCH_T
#error "M2P_BATCH_T must be double or float"
#endif

namespace m2p {

template <typename T>
__global__ void kb_inflow(BatchDev B) {
  const Member& m = B.mem[blockIdx.x];
  if (threadIdx.x != 0 || m.mode != 1 || halted_pre(m.st)) return;
  inflow_body(m.g, (T*)m.U0, m.M, m.table, m.loc, m.dire, m.nloc, m.nsamp, m.phiS0,
              m.st->inflowCount, m.LIM, m.dry);
}

// post = 0: fill U0 before the CFL pass (previous-stage source U1); post = 1: fill U1
// before the corrector (source U0).
template <typename T>
__global__ void __launch_bounds__(GHOST_LINES) kb_ghost(BatchDev B, int post) {
  const Member& m = B.mem[B.mapGhost[blockIdx.x]];
  if (post ? halted_post(m.st) : halted_pre(m.st)) return;
  const int line = (int)(blockIdx.x - m.ghost0) * GHOST_LINES + threadIdx.x;
  if (line >= m.g.nyd + m.g.nxd) return;
  const int k = (int)blockIdx.y;
  if (m.mode == 0 && k % 3 != 0) return;   // a landslide's momenta ghosts hold 0
  T* U = (T*)(post ? m.U1 : m.U0);
  const T* prev = (const T*)(post ? m.U0 : m.U1);
  ghost_line_skip(m.g, m.mode, U, prev, line, k, post ? m.dry.z1 : m.dry.z0,
                  post ? m.dry.g1 : m.dry.g0, m.dry.nseg);
}

// Blocks stride over the batch's CFL boxes, CFL_GROUP at a time: one round of loads tests
// all of them (CFL_LANES lanes per box; members without dry-tile skipping are always
// computed) and writes the dry boxes' slots (the wave-speed floor), then the block computes
// the group's wet boxes one after another, each maximum into its box's slot.  Halted
// members' slots are left alone: their dt kernel does not read them.  The blocks after the
// first nbb take the ring faces in the same launch: each of their 4 waves is one ring block
// of the block map (one member's 32 ring faces), with its own slot.  nbb is B.cfl_blocks
// or fewer (a tuned parameter).
constexpr int CFL_GROUP = 64;
constexpr int CFL_LANES = 256 / CFL_GROUP;

// Waves per SIMD: the ring faces and the dry tests are chains of dependent loads, so
// occupancy sets their pace (at the compiler's choice, 3 in FP64 and 4 in FP32, a 256-member
// ensemble step was 4% slower in FP32, 2% in FP64).
template <typename T, bool UNI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CFL_WAVES<T>, CFL_WAVES<T>)))
    kb_cfl(BatchDev B, int nbb) {
  __shared__ int state[CFL_GROUP];   // 0: dry, 1: compute, 2: halted or past the end
  const int tid = threadIdx.y * 64 + threadIdx.x;
  if ((int)blockIdx.x >= nbb) {   // ring faces, one ring block per wave
    const int rb = ((int)blockIdx.x - nbb) * 4 + threadIdx.y;
    if (rb >= B.nRing) return;
    const Member& m = B.mem[B.mapRing[rb]];
    if (halted_pre(m.st)) return;
    const int t = (rb - m.ring0) * 64 + threadIdx.x;
    double v = cfl_ring_face<T, UNI>((const T*)m.U0, (const T*)m.MT, m.g, m.ph, t, m.dry);
    v = wave_cfl(v);
    if (threadIdx.x == 0) B.part[B.nBox + rb] = v;
    return;
  }
  for (int t0 = blockIdx.x; t0 < B.nBox; t0 += CFL_GROUP * nbb) {
    __syncthreads();   // every thread is done with the previous group's state
    {
      const int k = tid / CFL_LANES, q = tid % CFL_LANES, t = t0 + k * nbb;
      int wet = 0, st = 2;
      const Member* mp = nullptr;
      if (t < B.nBox) {
        mp = &B.mem[B.mapBox[t]];
        if (halted_pre(mp->st)) {
          st = 2;
        } else if (mp->dry.nseg == 0) {
          st = 1;
        } else {
          st = -1;   // from the test
          const int lb = t - mp->box0;
          const DryRegion d = box_region(mp->g, lb % mp->nbx, lb / mp->nbx);
          wet = dry_flags_wet3<7, CFL_LANES>(d, mp->dry.z0, mp->dry.nseg, q);
          for (int e = d.nf + q; e < d.n; e += CFL_LANES)
            wet |= dry_item_wet(d, (const T*)mp->U0, mp->g, mp->dry.z0, mp->dry.nseg, e);
        }
      }
      // the box's CFL_LANES lanes are adjacent lanes of one wave
      const unsigned long long bal = __ballot(wet != 0);
      const int l0 = (int)threadIdx.x & ~(CFL_LANES - 1);
      const bool any = ((bal >> l0) & ((1ull << CFL_LANES) - 1ull)) != 0ull;
      if (q == 0) {
        if (st < 0) {
          st = any ? 1 : 0;
          if (!any) B.part[t] = cfl_floor<T>(mp->g);
        }
        state[k] = st;
      }
    }
    __syncthreads();
    for (int k = 0; k < CFL_GROUP; k++) {
      if (state[k] != 1) continue;
      const int t = t0 + k * nbb;
      const Member& m = B.mem[B.mapBox[t]];
      const int lb = t - m.box0;
      block_cfl_store(cfl_box<T, UNI>((const T*)m.U0, (const T*)m.MT, m.g, m.ph, lb % m.nbx,
                                      lb / m.nbx),
                      B.part + t);
    }
  }
}

// One block per member: its CFL and corrector partials, then the dt control.
template <typename T>
__global__ void __launch_bounds__(256) kb_dt(BatchDev B) {
  const Member& m = B.mem[blockIdx.x];
  if (B.plan_on && blockIdx.x == 0 && threadIdx.x < PLAN_COUNTERS) B.plan_cnt[threadIdx.x] = 0;
  __shared__ int halted;
  if (threadIdx.x == 0) halted = dt_halted(m.st) ? 1 : 0;
  __syncthreads();
  if (halted) return;
  const double mb = block_cfl_of(B.part + m.box0, m.box_n);
  const double mr = block_cfl_of(B.part + B.nBox + m.ring0, m.ring_n);
  if (threadIdx.x == 0) dt_body(m.dp, m.st, m.tf, m.itime, cfl_join(mb, mr));
}

// One block per member after the corrector: its squared max flow speed into DevState.
template <typename T>
__global__ void __launch_bounds__(256) kb_vmax(BatchDev B) {
  const Member& m = B.mem[blockIdx.x];
  if (halted_post(m.st) || !(m.st->vstop2 > 0.0)) return;
  const double v2 = block_max_of(B.part + B.nBox + B.nRing + m.stage0, m.stage_n);
  if (threadIdx.x == 0) {
    m.st->v2 = v2;
    m.st->v2_nt = m.st->nt;
  }
}

template <typename T>
hipError_t launch_batch_steps(const BatchDev& B, int ty, int w, bool any_inflow, int nsteps,
                              hipStream_t s) {
  if (B.nmem == 0 || nsteps <= 0) return hipSuccess;
  for (int n = 0; n < nsteps; n++) {
    if (any_inflow) kb_inflow<T><<<B.nmem, 64, 0, s>>>(B);
    kb_ghost<T><<<dim3(B.nGhost, MN), GHOST_LINES, 0, s>>>(B, 0);
    {
      const int nbb = B.nBox < B.cfl_blocks ? B.nBox : B.cfl_blocks;
      const unsigned nb = nbb + (B.nRing + 3) / 4;
      bool done = false;
      if (B.uni) {
        kb_cfl<T, true><<<nb, dim3(64, 4), 0, s>>>(B, nbb);
        done = true;
      }
      if (!done) kb_cfl<T, false><<<nb, dim3(64, 4), 0, s>>>(B, nbb);
    }
    kb_dt<T><<<B.nmem, 256, 0, s>>>(B);
    if (!launch_batch_stage<T>(B, ty, w, 0, s)) return hipErrorInvalidValue;
    kb_ghost<T><<<dim3(B.nGhost, MN), GHOST_LINES, 0, s>>>(B, 1);
    if (!launch_batch_stage<T>(B, ty, w, 1, s)) return hipErrorInvalidValue;
    if (B.vmax) kb_vmax<T><<<B.nmem, 256, 0, s>>>(B);
  }
  return hipGetLastError();
}

template hipError_t launch_batch_steps<M2P_BATCH_T>(const BatchDev&, int, int, bool, int, hipStream_t);

}  // namespace m2p
