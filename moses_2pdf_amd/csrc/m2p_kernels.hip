// m2p_kernels.hip -- HIP kernels of the adNOC_2Pb engine for gfx950.
//
// The reference launches ~85 kernels and synchronises after each of them per time step
// (MoSES_2PDF_landslide/adNOC_2Pb_runKernels.cu:646-1330).  Here one step is a handful of
// launches on one stream, with dt kept on the device:
//   ghost(U0) -> cfl(U0) -> dt -> stage<0>(U0 -> U1) -> ghost(U1) -> stage<1>(U1,U0 -> U0)
// (+ the inflow injection first in inflow mode).  The stage kernels are in m2p_stage.hip
// (compiled with their own scheduling option).  A stage kernel fuses the reference's
// TVD slopes, interface states, positivity, local speeds, central-upwind fluxes, pressure
// and viscous terms, sources, update, wet/dry clipping and the velocity-crossing limiter
// (L/K:1000-2937).  The per-step kernels are templated on the state's scalar type (double:
// the plain path; float: the FP32-bulk variant); the metric setup is always FP64.  Tile
// bodies are device functions so the batched-ensemble kernels reuse them.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>

#include "m2p_kernels.h"
#include "m2p_numerics.h"
#include "m2p_tiles.h"

namespace m2p {

// ======================================================================================
// terrain metric (once per run; L/R:395-642)
// ======================================================================================

// bfkt plane 0 / plane 1 after makeTopo1Kernel and makeTopo2Kernel (L/K:5-56).
__device__ __forceinline__ double coord(int c, int n, double d) {
  const double xmin = 0.0;
  if (c < MD) return (xmin + 0 * d) - (MD - c) * d;
  if (c >= n + MD) return (xmin + (n - 1) * d) + (c - (n + MD - 1)) * d;
  return xmin + (c - MD) * d;
}

// three-point non-uniform derivative of makeTopo4Kernel / JacobKernel (L/K:88-110, 674-696)
__device__ __forceinline__ double d3(double fm, double f0, double fp, double xm, double x0,
                                     double xp) {
  return fm * (x0 - xp) / ((xm - x0) * (xm - xp)) +
         f0 * ((xm - x0) - (x0 - xp)) / ((xm - x0) * (x0 - xp)) -
         fp * (xm - x0) / ((xm - xp) * (x0 - xp));
}

// makeTopo4..9Kernel (L/K:75-252) on [2, N-3]^2, fused per cell.
__global__ void k_metric_cell(Grid g, const double* __restrict__ topo, double* __restrict__ M) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int r = blockIdx.y * blockDim.y + threadIdx.y + 2;
  if (r > g.nyd - 3 || c > g.nxd - 3) return;
  const long long P = g.P, i = idx(g, r, c);
  const double xm = coord(c - 1, g.nx, g.dx), x0 = coord(c, g.nx, g.dx), xp = coord(c + 1, g.nx, g.dx);
  const double ym = coord(r - 1, g.ny, g.dy), y0 = coord(r, g.ny, g.dy), yp = coord(r + 1, g.ny, g.dy);
  const double d11 = d3(xm, x0, xp, xm, x0, xp);
  const double d12 = d3(x0, x0, x0, ym, y0, yp);
  const double d21 = d3(y0, y0, y0, xm, x0, xp);
  const double d22 = d3(ym, y0, yp, ym, y0, yp);
  const double bx = d3(topo[i - 1], topo[i], topo[i + 1], xm, x0, xp);
  const double by = d3(topo[i - g.nxd], topo[i], topo[i + g.nxd], ym, y0, yp);
  const double cv = 1.0 / sqrt(1.0 + bx * bx + by * by);              // makeTopo5
  const double s0 = cv * bx, s1 = cv * by;                            // makeTopo6
  const double Jb31 = (s0 * d11 + s1 * d21) / cv;                     // makeTopo7
  const double Jb32 = (s0 * d12 + s1 * d22) / cv;
  const double dettmp = d11 * d22 - d21 * d12;
  const double det = d11 * (d22 * cv - (-s1) * Jb32) -               // makeTopo8
                     d21 * (d12 * cv - Jb32 * (-s0)) + Jb31 * (d12 * (-s1) - d22 * (-s0));
  const double i11 = d22 / dettmp, i12 = -d12 / dettmp, i21 = -d21 / dettmp, i22 = d11 / dettmp;
  M[MC_C * P + i] = cv;                                               // makeTopo9
  M[MC_S0 * P + i] = s0;
  M[MC_S1 * P + i] = s1;
  M[MC_DET * P + i] = det;
  M[MC_I11 * P + i] = i11 - i11 * s0 * s0 - i12 * s1 * s0;
  M[MC_I12 * P + i] = i12 - i11 * s0 * s1 - i12 * s1 * s1;
  M[MC_I21 * P + i] = i21 - i21 * s0 * s0 - i22 * s1 * s0;
  M[MC_I22 * P + i] = i22 - i21 * s0 * s1 - i22 * s1 * s1;
  M[MC_I13 * P + i] = cv * (s0 * i11 + s1 * i12);
  M[MC_I23 * P + i] = cv * (s0 * i21 + s1 * i22);
  M[MC_X11 * P + i] = d11;
  M[MC_X12 * P + i] = d12;
  M[MC_X21 * P + i] = d21;
  M[MC_X22 * P + i] = d22;
}

// makeTopo10Kernel + makeTopo11Kernel (L/K:254-302), landslide: every cell, with cvalue as
// it is before the metric ghost copy (0 on the outer two rings).
__global__ void k_init_landslide(Grid g, const double* __restrict__ depth,
                                 const double* __restrict__ phicell, double phiS0,
                                 const double* __restrict__ M, double* __restrict__ U) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.P) return;
  const double cv = M[MC_C * g.P + i];
  const double phi = phicell ? phicell[i] : phiS0;
  const double u0 = depth[i] * phi * cv;
  const double u3 = depth[i] * (1.0 - phi) * cv;
  U[0 * g.P + i] = fmax(0.0, u0);
  U[1 * g.P + i] = 0.0;
  U[2 * g.P + i] = 0.0;
  U[3 * g.P + i] = fmax(0.0, u3);
  U[4 * g.P + i] = 0.0;
  U[5 * g.P + i] = 0.0;
  U[6 * g.P + i] = phi;
}

// Boundary1Kernel (L/K:304-479): x-ghost copy on rows [3, N-4] of planes [p0, p1).  The
// reference writes dxdxi21's column nx+MD+1 twice and never nx+MD+2 (L/K:407-409).
// (The inflow variant runs on every row, I/K:321; Boundary2 then overwrites the ghost
// rows, so the results are the same.)
__global__ void k_ghost_copy_x(Grid g, double* __restrict__ M, int p0, int p1) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x + 3;
  const int p = blockIdx.y + p0;
  if (r > g.nyd - 4 || p >= p1) return;
  double* A = M + (long long)p * g.P + (long long)r * g.nxd;
  const double a = A[MD], b = A[g.nx + MD - 1];
  A[0] = a;
  A[1] = a;
  A[2] = a;
  A[g.nx + MD + 0] = b;
  A[g.nx + MD + 1] = b;
  if (p != MC_X21) A[g.nx + MD + 2] = b;
}

// Boundary2Kernel (L/K:482-658): y-ghost copy on every column of planes [p0, p1).
__global__ void k_ghost_copy_y(Grid g, double* __restrict__ M, int p0, int p1) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y + p0;
  if (c >= g.nxd || p >= p1) return;
  double* A = M + (long long)p * g.P + c;
  const long long n = g.nxd;
  const double a = A[MD * n], b = A[(g.ny + MD - 1) * n];
  A[0 * n] = a;
  A[1 * n] = a;
  A[2 * n] = a;
  A[(g.ny + MD + 0) * n] = b;
  A[(g.ny + MD + 1) * n] = b;
  A[(g.ny + MD + 2) * n] = b;
}

// JacobKernel (L/K:660-701) on [3, N-4]^2, after the svec/cvalue ghost copies.
// MeanKernel (L/K:1121-1224): the face averages the flux code reads, once, on [1, N-3]^2,
// formed in FP64 from the cell planes of M and stored (rounded to TO) into the 18 face
// planes at F.
template <typename TO>
__global__ void k_face_metric(Grid g, const double* __restrict__ M, TO* __restrict__ F) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int r = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (r > g.nyd - 3 || c > g.nxd - 3) return;
  const long long P = g.P, i = idx(g, r, c), e = i + 1, n = i + g.nxd;
  const int cellq[FQ_COUNT] = {MC_I11, MC_I12, MC_I13, MC_I21, MC_I22, MC_I23, MC_DET, 0, 0};
#pragma unroll
  for (int q = 0; q < FQ_COUNT; q++) {
    const double* ax = M + (q == FQ_XA ? MC_X11 : q == FQ_XB ? MC_X21 : cellq[q]) * P;
    const double* ay = M + (q == FQ_XA ? MC_X12 : q == FQ_XB ? MC_X22 : cellq[q]) * P;
    F[q * P + i] = (TO)(0.5 * (ax[i] + ax[e]));
    F[(FQ_COUNT + q) * P + i] = (TO)(0.5 * (ay[i] + ay[n]));
  }
}

__global__ void k_jacob(Grid g, double* __restrict__ M) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 3;
  const int r = blockIdx.y * blockDim.y + threadIdx.y + 3;
  if (r > g.nyd - 4 || c > g.nxd - 4) return;
  const long long P = g.P, i = idx(g, r, c), w = i - 1, e = i + 1, s = i - g.nxd, n = i + g.nxd;
  const double* S0 = M + MC_S0 * P;
  const double* S1 = M + MC_S1 * P;
  const double* C = M + MC_C * P;
  const double xm = coord(c - 1, g.nx, g.dx), x0 = coord(c, g.nx, g.dx), xp = coord(c + 1, g.nx, g.dx);
  const double ym = coord(r - 1, g.ny, g.dy), y0 = coord(r, g.ny, g.dy), yp = coord(r + 1, g.ny, g.dy);
  M[MC_J13X * P + i] = d3(-S0[w], -S0[i], -S0[e], xm, x0, xp);
  M[MC_J23X * P + i] = d3(-S1[w], -S1[i], -S1[e], xm, x0, xp);
  M[MC_J33X * P + i] = d3(C[w], C[i], C[e], xm, x0, xp);
  M[MC_J13E * P + i] = d3(-S0[s], -S0[i], -S0[n], ym, y0, yp);
  M[MC_J23E * P + i] = d3(-S1[s], -S1[i], -S1[n], ym, y0, yp);
  M[MC_J33E * P + i] = d3(C[s], C[i], C[n], ym, y0, yp);
}

// ======================================================================================
// per step: single-scenario kernels
// ======================================================================================

template <typename T>
__global__ void k_inflow(Grid g, T* __restrict__ U, const double* __restrict__ M,
                         const double* __restrict__ table, const double* __restrict__ loc,
                         const int* __restrict__ dire, int nloc, int nsamp, double phiS0,
                         const DevState* __restrict__ st, unsigned char* __restrict__ LIM,
                         DryMap dm) {
  if (blockIdx.x != 0 || threadIdx.x != 0 || halted_pre(st)) return;
  inflow_body(g, U, M, table, loc, dire, nloc, nsamp, phiS0, st->inflowCount, LIM, dm);
}

__global__ void k_dry_fill(unsigned char* __restrict__ z, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    z[i] = 1;
}

template <typename T>
__global__ void k_ghost(Grid g, int mode, T* __restrict__ Tg, const T* __restrict__ prev,
                        const DevState* __restrict__ st, int post, const unsigned char* __restrict__ z,
                        unsigned char* __restrict__ gf, int nseg) {
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= g.nyd + g.nxd || (post ? halted_post(st) : halted_pre(st))) return;
  // a landslide fills planes 0, 3, 6 only: its momenta ghosts hold 0 (ghost_momenta_kept)
  const int k = mode == 0 ? 3 * (int)blockIdx.y : (int)blockIdx.y;
  ghost_line_skip(g, mode, Tg, prev, line, k, z, gf, nseg);
}

// The ghost ring of the momenta planes (1, 2, 4, 5) set to 0, as a landslide ghost fill
// would (after a host-side write of the state).
template <typename T>
__global__ void k_zero_ghost_momenta(Grid g, T* __restrict__ U) {
  const int line = blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= g.nyd + g.nxd) return;
  const int k = blockIdx.y < 2 ? 1 + (int)blockIdx.y : 2 + (int)blockIdx.y;
  ghost_line(g, 0, U, U, line, k);
}

// Blocks [0, nbb) stride over the 64x4 face boxes (at most CFL_BLOCKS, or CFL_BLOCKS_DRY, so the dt
// kernel reduces few partials); the blocks after them take the ring faces, 256 ring threads
// each (one wave per 32 faces), in the same launch.
constexpr int CFL_BLOCKS = 4096;
// with dry-tile skipping (most boxes dry): fewer box blocks, so the dt kernel reduces fewer
// slots (2048^2: skipping step FP64 -5%, FP32 -9%; the dense pass is fastest at 4096)
constexpr int CFL_BLOCKS_DRY = 1024;

// Waves per SIMD CFL_WAVES: with dry-tile skipping the box tests are chains of dependent
// loads (at the compiler's choice of 3 waves the FP64 2048^2 skipping step was 6% slower).
template <typename T, bool UNI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CFL_WAVES<T>, CFL_WAVES<T>)))
    k_cfl(Grid g, Phys ph, const T* __restrict__ U,
                                             const T* __restrict__ M, DevState* __restrict__ st,
                                             double* __restrict__ part, int nbx, int nbox,
                                             int nbb, DryMap dm) {
  if (halted_pre(st)) return;
  double m = 0.0;
  if ((int)blockIdx.x >= nbb) {   // ring faces
    const int t = ((int)blockIdx.x - nbb) * 256 + threadIdx.y * 64 + threadIdx.x;
    block_cfl_store(cfl_ring_face<T, UNI>(U, M, g, ph, t, dm), part + blockIdx.x);
    return;
  }
  if (dm.nseg == 0) {
    for (int t = blockIdx.x; t < nbox; t += nbb)
      m = cfl_join(m, cfl_box<T, UNI>(U, M, g, ph, t % nbx, t / nbx));
    block_cfl_store(m, part + blockIdx.x);
    return;
  }
  // Dry-tile skipping: the dry tests of up to 8 boxes of this block in one round of loads
  // (32 threads per box), then only the wet boxes are computed.
  __shared__ int wetk[8];
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const double fl = cfl_floor<T>(g);
  for (int t0 = blockIdx.x; t0 < nbox; t0 += 8 * nbb) {
    __syncthreads();   // every thread is done with the previous group's wetk
    if (tid < 8) wetk[tid] = 0;
    __syncthreads();
    {
      const int k = tid >> 5, t = t0 + k * nbb;
      if (t < nbox) {
        const DryRegion d = box_region(g, t % nbx, t / nbx);
        int wet = dry_flags_wet3<7, 32>(d, dm.z0, dm.nseg, tid & 31);
        for (int e = d.nf + (tid & 31); e < d.n; e += 32) wet |= dry_item_wet(d, U, g, dm.z0, dm.nseg, e);
        if (wet) atomicOr(&wetk[k], 1);
      }
    }
    __syncthreads();
    for (int k = 0; k < 8; k++) {
      const int t = t0 + k * nbb;
      if (t >= nbox) break;
      m = cfl_join(m, wetk[k] ? cfl_box<T, UNI>(U, M, g, ph, t % nbx, t / nbx) : fl);
    }
  }
  block_cfl_store(m, part + blockIdx.x);
}

// One block: the max over the CFL partials [0, ncfl), then the dt control in thread 0.
__global__ void __launch_bounds__(256) k_dt(DevParams dp, DevState* __restrict__ st,
                                            const double* __restrict__ tf,
                                            const double* __restrict__ itime,
                                            const double* __restrict__ part, int ncfl,
                                            int* __restrict__ plan_cnt) {
  __shared__ int halted;
  if (threadIdx.x == 0) halted = dt_halted(st) ? 1 : 0;
  __syncthreads();
  if (halted) return;
  if (plan_cnt && threadIdx.x < PLAN_COUNTERS) plan_cnt[threadIdx.x] = 0;
  const double mx = block_cfl_of(part, ncfl);
  if (threadIdx.x == 0) dt_body(dp, st, tf, itime, mx);
}

// One block, after the corrector of a step with a stop speed: the squared max flow speed
// over its tile partials, stamped with the step.
__global__ void __launch_bounds__(256) k_vmax(DevState* __restrict__ st,
                                              const double* __restrict__ part, int n) {
  if (halted_post(st) || !(st->vstop2 > 0.0)) return;
  const double v2 = block_max_of(part, n);
  if (threadIdx.x == 0) {
    st->v2 = v2;
    st->v2_nt = st->nt;
  }
}

// Two lanes per stage tile: its plan entry into its list (plan_list_of; tiles that need
// nothing have no entry; in the corrector their flow-speed partial is set to 0).
template <int IO, typename T, int TY>
__global__ void __launch_bounds__(256) k_plan(Grid g, const T* __restrict__ Uin, DryMap dm,
                                              TilePlan pl, double* __restrict__ spart,
                                              const DevState* __restrict__ st) {
  if (halted_post(st)) return;
  int t = (blockIdx.x * 4 + threadIdx.y) * (64 / PLAN_LANES) + threadIdx.x / PLAN_LANES;
  int tx = 0, ty = 0;
  if (!tile_of(g, TY, t, tx, ty)) t = -1;
  const int e = plan_tile<IO, T, TY>(g, Uin, dm, tx, ty, t, threadIdx.x % PLAN_LANES);
  const bool leader = t >= 0 && threadIdx.x % PLAN_LANES == 0;
  plan_append(pl, leader, e, t >= 0 ? plan_list_of(t) : 0);
  if (IO == 1 && leader && e < 0) spart[t] = 0.0;
}

// Sum of u0*Det and u3*Det over interior cells with double-double accumulation; one
// (hi, lo) pair per block and quantity, summed in block order on the host.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dd_add(double& hi, double& lo, double x) {
  double s, e;
  two_sum(hi, x, s, e);
  e += lo;
  two_sum(s, e, hi, lo);
}

template <typename T>
__global__ void k_mass(Grid g, const T* __restrict__ U, const double* __restrict__ det,
                       double* __restrict__ out) {
  const int tid = threadIdx.x;
  double sh = 0, sl = 0, fh = 0, fl = 0;
  const long long nint = (long long)g.nx * g.ny;
  for (long long q = (long long)blockIdx.x * blockDim.x + tid; q < nint;
       q += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(q / g.nx) + MD, c = (int)(q % g.nx) + MD;
    const long long i = idx(g, r, c);
    const double d = det[i];
    dd_add(sh, sl, (double)U[i] * d);
    dd_add(fh, fl, (double)U[3 * g.P + i] * d);
  }
  __shared__ double buf[4][256];
  buf[0][tid] = sh;
  buf[1][tid] = sl;
  buf[2][tid] = fh;
  buf[3][tid] = fl;
  __syncthreads();
  if (tid == 0) {
    double a = 0, al = 0, b = 0, bl = 0;
    for (int t = 0; t < (int)blockDim.x; t++) {
      dd_add(a, al, buf[0][t]);
      dd_add(a, al, buf[1][t]);
      dd_add(b, bl, buf[2][t]);
      dd_add(b, bl, buf[3][t]);
    }
    out[blockIdx.x * 4 + 0] = a;
    out[blockIdx.x * 4 + 1] = al;
    out[blockIdx.x * 4 + 2] = b;
    out[blockIdx.x * 4 + 3] = bl;
  }
}

// FP64 -> T copy of n values (the FP32 variant's metric and initial state).
template <typename T>
__global__ void k_convert(const double* __restrict__ a, T* __restrict__ b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (T)a[i];
}

// ======================================================================================
// launchers
// ======================================================================================

static dim3 grid2(int w, int h, dim3 b) { return dim3((w + b.x - 1) / b.x, (h + b.y - 1) / b.y); }

void launch_metric_setup(const Grid& g, int mode, const double* topo, const double* depth,
                         const double* phicell, double phiS0, double* M, double* U,
                         hipStream_t s) {
  const dim3 b(64, 4);
  k_metric_cell<<<grid2(g.nxd - 4, g.nyd - 4, b), b, 0, s>>>(g, topo, M);
  if (mode == 0) {
    k_init_landslide<<<(unsigned)((g.P + 255) / 256), 256, 0, s>>>(g, depth, phicell, phiS0, M, U);
  }
  k_ghost_copy_x<<<dim3((g.nyd + 63) / 64, MC_J13X), 64, 0, s>>>(g, M, 0, MC_J13X);
  k_ghost_copy_y<<<dim3((g.nxd + 63) / 64, MC_J13X), 64, 0, s>>>(g, M, 0, MC_J13X);
  k_jacob<<<grid2(g.nxd - 6, g.nyd - 6, b), b, 0, s>>>(g, M);
  k_ghost_copy_x<<<dim3((g.nyd + 63) / 64, MC_CELL_END - MC_J13X), 64, 0, s>>>(g, M, MC_J13X, MC_CELL_END);
  k_ghost_copy_y<<<dim3((g.nxd + 63) / 64, MC_CELL_END - MC_J13X), 64, 0, s>>>(g, M, MC_J13X, MC_CELL_END);
}

template <typename TO>
void launch_face_metric(const Grid& g, const double* M, TO* F, hipStream_t s) {
  const dim3 b(64, 4);
  k_face_metric<TO><<<grid2(g.nxd - 3, g.nyd - 3, b), b, 0, s>>>(g, M, F);
}
template void launch_face_metric<double>(const Grid&, const double*, double*, hipStream_t);
template void launch_face_metric<float>(const Grid&, const double*, float*, hipStream_t);

// One thread per padded cell (r, c): the x- and y-face averages at (r, c) when it is a face
// of [1, N-3]^2, and the cell's inverse Jacobian when (r, c) is in [3, N-4]^2, each formed
// as FM_UNIFORM forms them and compared bit for bit with what is stored.
__global__ void k_metric_uniform_check(Grid g, const double* __restrict__ M, int* __restrict__ ok) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= g.nyd || c >= g.nxd) return;
  const long long P = g.P, i = idx(g, r, c);
  auto same = [](double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
  };
  bool good = true;
  if (r >= 1 && r <= g.nyd - 3 && c >= 1 && c <= g.nxd - 3) {
    for (int dir = 0; dir < 2; dir++) {
      FaceMet<double> f;
      face_metric<double, FM_UNIFORM>(M, g, dir, r, c, f);
      const double* Fq = M + (dir == 0 ? MC_FX : MC_FY) * P;
      good = good && same(f.j11, Fq[FQ_I11 * P + i]) && same(f.j12, Fq[FQ_I12 * P + i]) &&
             same(f.j13, Fq[FQ_I13 * P + i]) && same(f.j21, Fq[FQ_I21 * P + i]) &&
             same(f.j22, Fq[FQ_I22 * P + i]) && same(f.j23, Fq[FQ_I23 * P + i]) &&
             same(f.xa, Fq[FQ_XA * P + i]) && same(f.xb, Fq[FQ_XB * P + i]);
    }
  }
  if (r >= MD && r <= g.nyd - 4 && c >= MD && c <= g.nxd - 4) {
    CellMet<double> a, b;
    cell_metric<double, true>(M, P, i, a);
    cell_metric<double, false>(M, P, i, b);
    good = good && same(a.i11, b.i11) && same(a.i12, b.i12) && same(a.i13, b.i13) &&
           same(a.i21, b.i21) && same(a.i22, b.i22) && same(a.i23, b.i23);
  }
  if (!good) atomicAnd(ok, 0);
}

void launch_metric_uniform_check(const Grid& g, const double* M, int* ok, hipStream_t s) {
  const dim3 b(64, 4);
  k_metric_uniform_check<<<grid2(g.nxd, g.nyd, b), b, 0, s>>>(g, M, ok);
}

__global__ void k_metric_uniform_cells_check(Grid g, const double* __restrict__ M,
                                             int* __restrict__ ok) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r < 1 || c < 1 || r > g.nyd - 2 || c > g.nxd - 2) return;
  const long long P = g.P, i = idx(g, r, c);
  auto same = [](double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
  };
  CellMet<double> a, b;
  cell_metric<double, true>(M, P, i, a);
  cell_metric<double, false>(M, P, i, b);
  const bool good = same(M[MC_X11 * P + i], 1.0) && same(M[MC_X22 * P + i], 1.0) &&
                    same(M[MC_X12 * P + i], 0.0) && same(M[MC_X21 * P + i], 0.0) &&
                    same(a.i11, b.i11) && same(a.i12, b.i12) && same(a.i13, b.i13) &&
                    same(a.i21, b.i21) && same(a.i22, b.i22) && same(a.i23, b.i23);
  if (!good) atomicAnd(ok, 0);
}

void launch_metric_uniform_cells_check(const Grid& g, const double* M, int* ok, hipStream_t s) {
  const dim3 b(64, 4);
  k_metric_uniform_cells_check<<<grid2(g.nxd, g.nyd, b), b, 0, s>>>(g, M, ok);
}

template <typename T>
void launch_convert(const double* a, T* b, long long n, hipStream_t s) {
  k_convert<T><<<1024, 256, 0, s>>>(a, b, n);
}
template void launch_convert<float>(const double*, float*, long long, hipStream_t);

template <typename T>
void launch_inflow(const Grid& g, T* U, const double* M, const double* table, const double* loc,
                   const int* dire, int nloc, int nsamp, double phiS0, const DevState* st,
                   unsigned char* LIM, const DryMap& dm, hipStream_t s) {
  k_inflow<T><<<1, 64, 0, s>>>(g, U, M, table, loc, dire, nloc, nsamp, phiS0, st, LIM, dm);
}


// One wave per row segment (63 cells, one per lane); blocks of 4 waves.
template <typename T>
__global__ void __launch_bounds__(256) k_dry_scan(Grid g, const T* __restrict__ U,
                                                  unsigned char* __restrict__ z, int nseg) {
  const long long seg = (long long)blockIdx.x * 4 + threadIdx.y;
  if (seg >= (long long)g.ny * nseg) return;   // the whole wave
  const int r = (int)(seg / nseg), c = (int)(seg % nseg) * 63 + (int)threadIdx.x;
  bool wet = false;
  if (threadIdx.x < 63 && c < g.nx) {
    const long long i = idx(g, r + MD, c + MD);
#pragma unroll
    for (int k = 0; k < 7; k++) wet = wet || nonzero_bits(U[k * g.P + i]);
  }
  const unsigned long long b = __ballot(wet);
  if (threadIdx.x == 0) z[seg] = b != 0ull;
}

template <typename T>
void launch_dry_scan(const Grid& g, const DryMap& dm, const T* U0, const T* U1, hipStream_t s) {
  if (dm.nseg <= 0) return;
  const long long n = (long long)g.ny * dm.nseg;
  const unsigned nb = (unsigned)((n + 3) / 4);
  k_dry_scan<T><<<nb, dim3(64, 4), 0, s>>>(g, U0, dm.z0, dm.nseg);
  k_dry_scan<T><<<nb, dim3(64, 4), 0, s>>>(g, U1, dm.z1, dm.nseg);
  launch_ghost_flags_fill(g, dm, s);   // the ghost cells may hold anything: fill every line
}
template void launch_dry_scan<double>(const Grid&, const DryMap&, const double*, const double*,
                                      hipStream_t);
template void launch_dry_scan<float>(const Grid&, const DryMap&, const float*, const float*,
                                     hipStream_t);

// COPY_BLOCKS_PER_ITEM blocks per item, each striding over the item.
__global__ void __launch_bounds__(256) k_copy_list(const CopyItem* items) {
  const CopyItem c = items[blockIdx.x / COPY_BLOCKS_PER_ITEM];
  const size_t lane = (size_t)(blockIdx.x % COPY_BLOCKS_PER_ITEM) * 256 + threadIdx.x;
  const size_t stride = (size_t)COPY_BLOCKS_PER_ITEM * 256;
  size_t done = 0;
  if ((((uintptr_t)c.src | (uintptr_t)c.dst) & 15) == 0) {
    const uint4* s = (const uint4*)c.src;
    uint4* d = (uint4*)c.dst;
    const size_t n16 = c.n / 16;
    for (size_t i = lane; i < n16; i += stride) d[i] = s[i];
    done = n16 * 16;
  }
  for (size_t i = done + lane; i < c.n; i += stride) c.dst[i] = c.src[i];
}

__global__ void __launch_bounds__(64) k_gather_states(const DevState* const* __restrict__ src,
                                                     DevState* __restrict__ dst, int n) {
  const int k = blockIdx.x;
  if (k >= n) return;
  static_assert(sizeof(DevState) % 4 == 0, "DevState is copied in 4-byte words");
  const unsigned* a = (const unsigned*)src[k];
  unsigned* d = (unsigned*)(dst + k);
  for (int i = threadIdx.x; i < (int)(sizeof(DevState) / 4); i += 64) d[i] = a[i];
}

void launch_gather_states(const DevState* const* src, DevState* dst, int n, hipStream_t s) {
  if (n > 0) k_gather_states<<<n, 64, 0, s>>>(src, dst, n);
}

void launch_copy_list(const CopyItem* items, int nitems, hipStream_t s) {
  if (nitems <= 0) return;
  k_copy_list<<<(unsigned)nitems * COPY_BLOCKS_PER_ITEM, 256, 0, s>>>(items);
}

void launch_dry_fill(const Grid& g, const DryMap& dm, hipStream_t s) {
  if (dm.nseg <= 0) return;
  const long long n = dry_flag_count(g);
  const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 1024);
  k_dry_fill<<<nb, 256, 0, s>>>(dm.z0, n);
  k_dry_fill<<<nb, 256, 0, s>>>(dm.z1, n);
  launch_ghost_flags_fill(g, dm, s);
}

void launch_ghost_flags_fill(const Grid& g, const DryMap& dm, hipStream_t s) {
  if (!dm.g0) return;
  const long long n = ghost_flag_count(g);
  const unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 1024);
  k_dry_fill<<<nb, 256, 0, s>>>(dm.g0, n);
  k_dry_fill<<<nb, 256, 0, s>>>(dm.g1, n);
}

template <typename T>
void launch_ghost(const Grid& g, int mode, T* Tg, const T* prev, const DevState* st, int post,
                  const unsigned char* z, unsigned char* gf, int nseg, hipStream_t s) {
  const int lines = g.nyd + g.nxd;
  k_ghost<T><<<dim3((lines + 127) / 128, mode == 0 ? 3 : MN), 128, 0, s>>>(g, mode, Tg, prev, st,
                                                                          post, z, gf, nseg);
}

template <typename T>
void launch_zero_ghost_momenta(const Grid& g, T* U, hipStream_t s) {
  k_zero_ghost_momenta<T><<<dim3((g.nyd + g.nxd + 127) / 128, 4), 128, 0, s>>>(g, U);
}

static int cfl_boxes(const Grid& g) { return ((g.nx + 63) / 64) * ((g.ny + 3) / 4); }
// The box blocks of a CFL launch (dry: with dry-tile skipping).  The partial array keeps
// room for the dense count, so the stage partials sit at the same offset either way.
static int cfl_box_blocks(const Grid& g, bool dry = false) {
  return std::min(cfl_boxes(g), dry ? CFL_BLOCKS_DRY : CFL_BLOCKS);
}
static int cfl_ring_blocks(const Grid& g) { return (ring_threads(g) + 255) / 256; }

// Tile height of the single-scenario stage: planned (dry-tile skipping) launches always take
// 3-row tiles, dense ones plain_ty's.
static int plain_tile_height(const Grid& g, int prec, bool planned) {
  return planned ? PLAIN_TY_SMALL : plain_ty(g, prec == 1 ? PlainStage<float>::TY : PlainStage<double>::TY);
}

size_t partial_count(const Grid& g, int prec) {
  return (size_t)cfl_box_blocks(g) + cfl_ring_blocks(g) + plain_stage_tiles(g, prec);
}

template <typename T>
void launch_cfl(const Grid& g, const Phys& ph, const T* U, const T* M, DevState* st, double* part,
                const DryMap& dm, bool uni, hipStream_t s) {
  const dim3 b(64, 4);
  const bool dry = dm.nseg > 0;
  const int nb = cfl_box_blocks(g, dry) + cfl_ring_blocks(g);
  if (uni) {
    k_cfl<T, true><<<nb, b, 0, s>>>(g, ph, U, M, st, part, (g.nx + 63) / 64, cfl_boxes(g),
                                    cfl_box_blocks(g, dry), dm);
    return;
  }
  k_cfl<T, false><<<nb, b, 0, s>>>(g, ph, U, M, st, part, (g.nx + 63) / 64, cfl_boxes(g),
                                   cfl_box_blocks(g, dry), dm);
}

void launch_dt(const DevParams& dp, DevState* st, const double* tf, const double* itime,
               const Grid& g, int prec, const double* part, int* plan_cnt, hipStream_t s) {
  (void)prec;
  // the plan counters are passed exactly when dry-tile skipping is on
  const int ncfl = cfl_box_blocks(g, plan_cnt != nullptr) + cfl_ring_blocks(g);
  k_dt<<<1, 256, 0, s>>>(dp, st, tf, itime, part, ncfl, plan_cnt);
}

int plain_stage_tiles(const Grid& g, int prec) {
  return std::max(stage_tiles(g, plain_tile_height(g, prec, true)),
                  stage_tiles(g, plain_tile_height(g, prec, false)));
}

template <typename T, int TY>
static void launch_plan_ty(int io, const Grid& g, const T* Uin, const DryMap& dm,
                           const TilePlan& pl, double* part, const DevState* st, hipStream_t s) {
  const unsigned nb = (unsigned)((stage_tiles(g, TY) + 4 * (64 / PLAN_LANES) - 1) / (4 * (64 / PLAN_LANES)));
  double* sp = part + cfl_box_blocks(g) + cfl_ring_blocks(g);
  if (io == 0)
    k_plan<0, T, TY><<<nb, dim3(64, 4), 0, s>>>(g, Uin, dm, pl, sp, st);
  else
    k_plan<1, T, TY><<<nb, dim3(64, 4), 0, s>>>(g, Uin, dm, pl, sp, st);
}

template <typename T>
void launch_plan(int io, const Grid& g, const T* Uin, const DryMap& dm, const TilePlan& pl,
                 double* part, const DevState* st, hipStream_t s) {
  launch_plan_ty<T, PLAIN_TY_SMALL>(io, g, Uin, dm, pl, part, st, s);
}

void launch_vmax(DevState* st, const Grid& g, int prec, bool planned, const double* part,
                 hipStream_t s) {
  k_vmax<<<1, 256, 0, s>>>(st, part + cfl_box_blocks(g) + cfl_ring_blocks(g),
                           stage_tiles(g, plain_tile_height(g, prec, planned)));
}

size_t stage_partials_offset(const Grid& g) {
  return (size_t)cfl_box_blocks(g) + cfl_ring_blocks(g);
}

int mass_blocks() { return 512; }

template <typename T>
void launch_mass(const Grid& g, const T* U, const double* det, double* out, hipStream_t s) {
  k_mass<T><<<mass_blocks(), 256, 0, s>>>(g, U, det, out);
}

// One thread: hi + lo accumulation of the partials in block order (a two-sum of the
// running high part and the partial, the low parts added, then renormalised).
__global__ void k_mass_sum(const double* __restrict__ part, int nb, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  auto add = [](double& hi, double& lo, double x) {
    double s = hi + x, bb = s - hi, e = (hi - (s - bb)) + (x - bb);
    e += lo;
    hi = s + e;
    lo = e - (hi - s);
  };
  double a = 0, al = 0, b = 0, bl = 0;
  for (int i = 0; i < nb; i++) {
    add(a, al, part[4 * i + 0]);
    add(a, al, part[4 * i + 1]);
    add(b, bl, part[4 * i + 2]);
    add(b, bl, part[4 * i + 3]);
  }
  out[0] = a + al;
  out[1] = b + bl;
}

void launch_mass_sum(const double* part, int nb, double* out, hipStream_t s) {
  k_mass_sum<<<1, 64, 0, s>>>(part, nb, out);
}

#define M2P_INSTANTIATE(T)                                                                    \
  template void launch_inflow<T>(const Grid&, T*, const double*, const double*, const double*, \
                                 const int*, int, int, double, const DevState*,                \
                                 unsigned char*, const DryMap&, hipStream_t);                  \
  template void launch_ghost<T>(const Grid&, int, T*, const T*, const DevState*, int,         \
                                const unsigned char*, unsigned char*, int, hipStream_t);      \
  template void launch_zero_ghost_momenta<T>(const Grid&, T*, hipStream_t);                     \
  template void launch_cfl<T>(const Grid&, const Phys&, const T*, const T*, DevState*, double*, \
                              const DryMap&, bool, hipStream_t);                               \
  template void launch_plan<T>(int, const Grid&, const T*, const DryMap&, const TilePlan&,     \
                               double*, const DevState*, hipStream_t);                         \
  template void launch_mass<T>(const Grid&, const T*, const double*, double*, hipStream_t);
M2P_INSTANTIATE(double)
M2P_INSTANTIATE(float)

}  // namespace m2p
