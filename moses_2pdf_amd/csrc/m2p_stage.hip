// m2p_stage.hip -- the single-scenario stage kernels (predictor and corrector) and their
// launcher.  They sit in a translation unit of their own, compiled once per scalar type
// (M2P_STAGE_T is double or float), so that the build can give each the LLVM scheduling
// strategy that suits it (a different one suits the CFL kernel).  The device body is
// stage_tile (m2p_tiles.h), shared with the batched kernels.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "m2p_kernels.h"
#include "m2p_numerics.h"
#include "m2p_tiles.h"

#ifndef M2P_STAGE_T
#error "M2P_STAGE_T must be double or float"
#endif

namespace m2p {

// PLANNED: dry-tile skipping, one block per plan-list slot (a kernel of its own, so the
// dense kernel keeps its registers).  UNI: the uniform-metric form (stage_tile).
template <int IO, typename T, int TY, int W, bool PLANNED, bool UNI>
__global__ void __launch_bounds__(64 * (TY + 1)) __attribute__((amdgpu_waves_per_eu(W, W)))
    k_stage(Grid g, Phys ph, const T* __restrict__ Uin, const T* Uzero, T* Uout,
            T* __restrict__ ONE, const T* __restrict__ M, unsigned char* __restrict__ LIM,
            DevState* __restrict__ st, double* __restrict__ part, DryMap dm, TilePlan pl) {
  __shared__ T F[ST_NQ][TY + 1][64];
  if (halted_post(st)) return;
  const int b = blockIdx.x;
  int t;
  if (PLANNED) {   // dry-tile skipping: this block's plan entry
    const int e = plan_entry(pl);
    if (e < 0) return;
    t = e & ~PLAN_ZERO;
    if (IO == 0 && (e & PLAN_ZERO)) {
      int tx, ty;
      tile_of(g, TY, t, tx, ty);
      stage_zero<T, TY>(g, Uout, ONE, dm, tx, ty);
      return;
    }
  } else {
    const int per = (stage_tiles(g, TY) + N_XCD - 1) / N_XCD;
    t = (b % N_XCD) * per + b / N_XCD;
  }
  int tx, ty;
  if (!tile_of(g, TY, t, tx, ty)) return;
  const bool want_v2 = IO == 1 && st->vstop2 > 0.0;
  // an FP32 corrector above 6 waves keeps the compiler's load order (it spills otherwise)
  constexpr bool LB = !(std::is_same<T, float>::value && IO == 1 && W > 6);
  const double v2 = stage_tile<IO, T, TY, PLANNED, UNI, LB>(F, g, ph, Uin, Uzero, Uout, ONE, M, LIM,
                                                   IO == 0 ? st->dt_pred : st->dt_corr, tx, ty, dm,
                                                   IO == 0 ? st->dtdx_pred : st->dtdx_corr,
                                                   IO == 0 ? st->dtdy_pred : st->dtdy_corr, want_v2);
  if (want_v2) block_max_store(v2, part + t);
}

template <typename T, bool UNI, int TY>
static void launch_stage_ty(int io, const Grid& g, const Phys& ph, const T* Uin, const T* Uzero,
                            T* Uout, T* ONE, const T* M, unsigned char* LIM, DevState* st,
                            double* part, const DryMap& dm, const TilePlan& pl, hipStream_t s) {
  constexpr int W = PlainStage<T>::W;
  constexpr int WP = UNI ? PlainStage<T>::W_PRED : W, WC = UNI ? PlainStage<T>::W_CORR : W;
  const dim3 b(64, TY + 1);
  const dim3 gr(pl.list ? plan_blocks(stage_tiles(g, TY))
                        : ((stage_tiles(g, TY) + N_XCD - 1) / N_XCD) * N_XCD);
  double* sp = part + stage_partials_offset(g);
  if (pl.list) {
    if (io == 0)
      k_stage<0, T, TY, W, true, UNI><<<gr, b, 0, s>>>(g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, sp, dm, pl);
    else
      k_stage<1, T, TY, W, true, UNI><<<gr, b, 0, s>>>(g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, sp, dm, pl);
  } else {
    if (io == 0)
      k_stage<0, T, TY, WP, false, UNI><<<gr, b, 0, s>>>(g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, sp, dm, pl);
    else
      k_stage<1, T, TY, WC, false, UNI><<<gr, b, 0, s>>>(g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, sp, dm, pl);
  }
}

// the tile height plain_ty picks for the grid, and 3-row tiles for planned launches (with
// dry-tile skipping the wet tiles are few: shorter tiles finish sooner; at 2048^2 the
// skipping step took 13% less time in FP64, 10% in FP32, than with 7-row tiles)
template <typename T, bool UNI>
static void launch_stage_t(int io, const Grid& g, const Phys& ph, const T* Uin, const T* Uzero,
                           T* Uout, T* ONE, const T* M, unsigned char* LIM, DevState* st,
                           double* part, const DryMap& dm, const TilePlan& pl, hipStream_t s) {
  if (pl.list || plain_ty(g, PlainStage<T>::TY) == PLAIN_TY_SMALL)
    launch_stage_ty<T, UNI, PLAIN_TY_SMALL>(io, g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, part, dm, pl, s);
  else
    launch_stage_ty<T, UNI, PlainStage<T>::TY>(io, g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, part, dm, pl, s);
}

template <typename T>
void launch_stage(int io, const Grid& g, const Phys& ph, const T* Uin, const T* Uzero, T* Uout,
                  T* ONE, const T* M, unsigned char* LIM, DevState* st, double* part,
                  const DryMap& dm, const TilePlan& pl, bool uni, hipStream_t s) {
  if (uni)
    launch_stage_t<T, true>(io, g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, part, dm, pl, s);
  else
    launch_stage_t<T, false>(io, g, ph, Uin, Uzero, Uout, ONE, M, LIM, st, part, dm, pl, s);
}

template void launch_stage<M2P_STAGE_T>(int, const Grid&, const Phys&, const M2P_STAGE_T*,
                                        const M2P_STAGE_T*, M2P_STAGE_T*, M2P_STAGE_T*,
                                        const M2P_STAGE_T*, unsigned char*, DevState*, double*,
                                        const DryMap&, const TilePlan&, bool, hipStream_t);

#ifdef M2P_STAGE_PROBE   // in the FP64 object only
// Confirms the library holds a code object for the current device (a missing gfx950 image
// otherwise surfaces only at the first launch).
hipError_t probe_kernels() {
  hipFuncAttributes a;
  return hipFuncGetAttributes(&a, reinterpret_cast<const void*>(
                                   &k_stage<0, double, PlainStage<double>::TY, PlainStage<double>::W, false, false>));
}

#endif

}  // namespace m2p
