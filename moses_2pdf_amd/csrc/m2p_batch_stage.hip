// m2p_batch_stage.hip -- the batched stage kernels (plan and stage, every compiled tile
// shape) and their launcher, in a translation unit of their own so that the build can give
// them the max-memory-clause scheduling strategy, as the single-scenario stage kernels.
// Compiled once per scalar type: M2P_BATCH_T is double or float.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "m2p_batch.h"
#include "m2p_tiles.h"

#ifndef M2P_BATCH_T
#error "M2P_BATCH_T must be double or float"
#endif

namespace m2p {

// One wave per stage tile of the batch: its plan entry (see TilePlan).
template <int IO, typename T, int TY>
__global__ void __launch_bounds__(256) kb_plan(BatchDev B) {
  int t = (blockIdx.x * 4 + threadIdx.y) * (64 / PLAN_LANES) + threadIdx.x / PLAN_LANES;
  int tx = 0, ty = 0;
  const Member* m = nullptr;
  if (t < B.nStage) {
    m = &B.mem[B.mapStage[t]];
    if (halted_post(m->st) || !tile_of(m->g, TY, t - m->stage0, tx, ty)) t = -1;
  } else {
    t = -1;
  }
  const Grid g = m ? m->g : Grid{};
  const DryMap dm = m ? m->dry : DryMap{};
  const T* Uin = m ? (const T*)(IO == 0 ? m->U0 : m->U1) : nullptr;
  const int e = plan_tile<IO, T, TY>(g, Uin, dm, tx, ty, t, threadIdx.x % PLAN_LANES);
  const TilePlan& pl = B.plan[IO];
  const bool leader = t >= 0 && threadIdx.x % PLAN_LANES == 0;
  plan_append(pl, leader, e, t >= 0 ? plan_list_of(t) : 0);
  if (IO == 1 && leader && e < 0) B.part[B.nBox + B.nRing + t] = 0.0;
}

// One stage tile t of the batch (with zero: a dry predictor tile to clear).  Every thread
// of the block calls it (it holds barriers unless the tile is cleared or its member halted).
template <int IO, typename T, int TY, int W, bool PLANNED, bool UNI>
__device__ __forceinline__ void batch_tile(const BatchDev& B, T (*F)[TY + 1][64], int t, bool zero) {
  if (t >= B.nStage) return;
  const Member& m = B.mem[B.mapStage[t]];
  DevState* st = m.st;
  if (halted_post(st)) return;
  int tx, ty;
  if (!tile_of(m.g, TY, t - m.stage0, tx, ty)) return;
  const T* Uin = (const T*)(IO == 0 ? m.U0 : m.U1);
  T* Uout = (T*)(IO == 0 ? m.U1 : m.U0);
  if (zero) {
    stage_zero<T, TY>(m.g, Uout, (T*)m.ONE, m.dry, tx, ty);
    return;
  }
  constexpr bool LB = !(std::is_same<T, float>::value && W >= 7);
  const bool want_v2 = IO == 1 && st->vstop2 > 0.0;
  const double v2 = stage_tile<IO, T, TY, PLANNED, UNI, LB>(F, m.g, m.ph, Uin, (const T*)m.U0, Uout, (T*)m.ONE,
                                          (const T*)m.MT, m.LIM,
                                          IO == 0 ? st->dt_pred : st->dt_corr, tx, ty, m.dry,
                                          IO == 0 ? st->dtdx_pred : st->dtdx_corr,
                                          IO == 0 ? st->dtdy_pred : st->dtdy_corr, want_v2);
  if (want_v2) block_max_store(v2, B.part + B.nBox + B.nRing + t);
}

// PLANNED: dry-tile skipping, one block per plan-list slot of the first plan_cap slots of
// each list (halted members have no entries); a kernel of its own, so the dense kernel
// keeps its registers.  UNI: every member's metric has the uniform form (BatchDev::uni).
template <int IO, typename T, int TY, int W, bool PLANNED, bool UNI>
__global__ void __launch_bounds__(64 * (TY + 1)) __attribute__((amdgpu_waves_per_eu(W, W)))
    kb_stage(BatchDev B) {
  __shared__ T F[ST_NQ][TY + 1][64];
  const int b = blockIdx.x;
  int t = b;
  bool zero = false;
  if (PLANNED) {
    const int e = plan_entry(B.plan[IO]);
    if (e < 0) return;
    t = e & ~PLAN_ZERO;
    zero = IO == 0 && (e & PLAN_ZERO);
  } else if (B.remap) {
    const int per = (B.nStage + N_XCD - 1) / N_XCD;
    t = (b % N_XCD) * per + b / N_XCD;
  }
  batch_tile<IO, T, TY, W, PLANNED, UNI>(B, F, t, zero);
}

// Kept out of line: inlined into the loop below, the tile body spilled 24-29 VGPRs at 8
// waves per SIMD (6 this way).
template <int IO, typename T, int TY, int W, bool UNI>
__device__ __attribute__((noinline)) void batch_tile_call(const BatchDev& B, T (*F)[TY + 1][64], int t, bool zero) {
  batch_tile<IO, T, TY, W, true, UNI>(B, F, t, zero);
}

// The plan entries past the first plan_cap of each list: STAGE_MORE_BLOCKS blocks, block b
// taking entries plan_cap + b / 8, plan_cap + b / 8 + STAGE_MORE_BLOCKS / 8, ... of list
// b % 8.  Usually there are none and every block returns at once; a tile costs about twice
// what it does in kb_stage.
template <int IO, typename T, int TY, int W, bool UNI>
__global__ void __launch_bounds__(64 * (TY + 1)) __attribute__((amdgpu_waves_per_eu(W, W)))
    kb_stage_more(BatchDev B) {
  __shared__ T F[ST_NQ][TY + 1][64];
  const TilePlan& pl = B.plan[IO];
  const int x = blockIdx.x % 8, n = pl.cnt[x];
  for (int j = B.plan_cap + (int)blockIdx.x / 8; j < n; j += STAGE_MORE_BLOCKS / 8) {
    const int e = pl.list[x * pl.per + j];
    batch_tile_call<IO, T, TY, W, UNI>(B, F, e & ~PLAN_ZERO, IO == 0 && (e & PLAN_ZERO));
  }
}

template <typename T, int TY, int W, bool UNI>
static void stage_pair_t(const BatchDev& B, int io, hipStream_t s) {
  const dim3 blk(64, TY + 1);
  const unsigned nb = B.plan_on ? (unsigned)(8 * std::min(plan_per(B.nStage), B.plan_cap))
                   : B.remap ? (unsigned)(((B.nStage + N_XCD - 1) / N_XCD) * N_XCD) : (unsigned)B.nStage;
  const unsigned np = (unsigned)((B.nStage + 4 * (64 / PLAN_LANES) - 1) / (4 * (64 / PLAN_LANES)));
  if (B.plan_on) {
    const bool more = plan_per(B.nStage) > B.plan_cap;   // a list can outgrow the capped launch
    if (io == 0) {
      kb_plan<0, T, TY><<<np, dim3(64, 4), 0, s>>>(B);
      kb_stage<0, T, TY, W, true, UNI><<<nb, blk, 0, s>>>(B);
      if (more) kb_stage_more<0, T, TY, W, UNI><<<STAGE_MORE_BLOCKS, blk, 0, s>>>(B);
    } else {
      kb_plan<1, T, TY><<<np, dim3(64, 4), 0, s>>>(B);
      kb_stage<1, T, TY, W, true, UNI><<<nb, blk, 0, s>>>(B);
      if (more) kb_stage_more<1, T, TY, W, UNI><<<STAGE_MORE_BLOCKS, blk, 0, s>>>(B);
    }
  } else {
    if (io == 0)
      kb_stage<0, T, TY, W, false, UNI><<<nb, blk, 0, s>>>(B);
    else
      kb_stage<1, T, TY, W, false, UNI><<<nb, blk, 0, s>>>(B);
  }
}

template <typename T, int TY, int W>
static void stage_pair(const BatchDev& B, int io, hipStream_t s) {
  if (B.uni)
    stage_pair_t<T, TY, W, true>(B, io, s);
  else
    stage_pair_t<T, TY, W, false>(B, io, s);
}

// The compiled (tile height, waves per SIMD) variants; the list in stage_variants() must
// match.
template <typename T>
bool launch_batch_stage(const BatchDev& B, int ty, int w, int io, hipStream_t s) {
  if constexpr (std::is_same<T, double>::value) {
    if (ty == 7 && w == 4) stage_pair<double, 7, 4>(B, io, s);
    else if (ty == 3 && w == 4) stage_pair<double, 3, 4>(B, io, s);
    else if (ty == 7 && w == 3) stage_pair<double, 7, 3>(B, io, s);
    else return false;
  } else {
    if (ty == 7 && w == 6) stage_pair<float, 7, 6>(B, io, s);
    else if (ty == 7 && w == 4) stage_pair<float, 7, 4>(B, io, s);
    else if (ty == 3 && w == 6) stage_pair<float, 3, 6>(B, io, s);
    else if (ty == 15 && w == 4) stage_pair<float, 15, 4>(B, io, s);
    // the 8-wave variants run their corrector at 7 waves (at 8 it spills 4 VGPRs):
    // ensemble stepping 3% faster
    else if (ty == 3 && w == 8) { if (io == 0) stage_pair<float, 3, 8>(B, io, s); else stage_pair<float, 3, 7>(B, io, s); }
    else if (ty == 7 && w == 8) { if (io == 0) stage_pair<float, 7, 8>(B, io, s); else stage_pair<float, 7, 7>(B, io, s); }
    else return false;
  }
  return true;
}

template bool launch_batch_stage<M2P_BATCH_T>(const BatchDev&, int, int, int, hipStream_t);

}  // namespace m2p
