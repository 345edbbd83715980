// m2p_numerics.h -- per-face and per-cell numerics of the adNOC_2Pb scheme as device
// functions, shared by the CFL pass and the stage kernels, templated on the scalar type.
//
// T = double is the plain path: every expression keeps the operand order of the
// reference kernel it restates and the library is compiled with -ffp-contract=off, so
// results are bit-identical to a serial IEEE FP64 evaluation of
// MoSES_2PDF_landslide/adNOC_2Pb_kernels.cu (cited as L/K).
// T = float is the FP32-bulk variant: the same expressions in single precision, with the
// hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32) in place of the correctly
// rounded sequences.  Its metric comes from the FP64 setup; dt, the CFL maximum and the
// mass totals stay FP64.
#pragma once

#include <hip/hip_runtime.h>

namespace m2p {

constexpr int MD = 3;             // ghost width, L/KH:7
constexpr int MN = 7;             // state components, L/KH:8
constexpr double H0 = 1.0e-6;     // L/KH:11
constexpr double H00 = 0.0;       // L/KH:12
constexpr double U0 = 1.0e-6;     // L/KH:13

// Metric planes kept on the device (padded nxd*nyd each).  Cell-centred planes first:
// invJ31, invJ32 and invJ33 are exactly -svec0, -svec1 and cvalue (L/K:244-248), so they
// are not stored.  Then the face averages of MeanKernel (L/K:1121-1224) for x-faces
// (r,c)-(r,c+1) and y-faces (r,c)-(r+1,c), stored at (r,c) for faces in [1, N-3]^2, in
// the order of FaceQ (the x-face's XA/XB are dxdxi11/dxdxi21, the y-face's dxdxi12/22).
enum MetricPlane {
  MC_C = 0, MC_S0, MC_S1, MC_DET,
  MC_I11, MC_I12, MC_I13, MC_I21, MC_I22, MC_I23,
  MC_X11, MC_X12, MC_X21, MC_X22,
  MC_J13X, MC_J23X, MC_J33X, MC_J13E, MC_J23E, MC_J33E,
  MC_CELL_END,
  MC_FX = MC_CELL_END,           // 9 x-face planes
  MC_FY = MC_FX + 9,             // 9 y-face planes
  MC_COUNT = MC_FY + 9
};
enum FaceQ { FQ_I11 = 0, FQ_I12, FQ_I13, FQ_I21, FQ_I22, FQ_I23, FQ_DET, FQ_XA, FQ_XB, FQ_COUNT };

struct Grid {
  int nx, ny, nxd, nyd;
  long long P;          // nxd*nyd
  double dx, dy;
};

struct Phys {
  double rr;        // density ratio (1.42/2.60 in the reference)
  double Drtio1;    // 1 - rr
  double Cd, N_R, varTheta;
  double tande;     // tan(delta0*3.14159/180), one value for the whole plane (L/K:298)
};

__device__ __forceinline__ long long idx(const Grid& g, int r, int c) {
  return (long long)r * g.nxd + c;
}

// quotient and square root: IEEE for FP64, the hardware approximations for FP32
__device__ __forceinline__ double qdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double qsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float qsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }

// TVD1..5Kernel (L/K:1000-1119) for one component: the slope at a cell from its two
// neighbours, dA = dux[c-1], dB = dux[c].  The reference forms
//   t1 = (sA+sB)/2 * min(dA*sA, dB*sB),  t2 = (dA+dB)/2,  slope = (sA'+sB')/2 * min(t1*sA', t2*sB')
// with s = (x > 0 ? 1 : -1).  For finite dA, dB that equals minmod(dA, dB): when both are
// positive or both negative it is sign(dA)*min(|dA|,|dB|) exactly (|t2| >= |t1| holds after
// rounding because rounding is monotonic), otherwise a zero.  minmod(dA, dB) is in turn the
// median of dA, dB and 0 (both positive: the smaller; both negative: the one nearer 0; else
// 0), formed here as max(min(dA, dB), min(max(dA, dB), 0)): four min/max instructions, no
// product (dA*dB underflows for the tiny differences at a wet/dry front) and no sign tests.
// Only the sign of a zero slope can differ from the reference's, and a zero's sign never
// reaches a nonzero value in this scheme (no quantity is divided by a slope and no branch
// tests the sign of a zero).  A NaN difference (a state already non-finite) gives another
// result than the reference's products, which the run reports through its non-finite flag.
template <typename T>
__device__ __forceinline__ T tvd_slope(T um, T u0, T up) {
  const T dA = u0 - um;
  const T dB = up - u0;
  return fmax(fmin(dA, dB), fmin(fmax(dA, dB), T(0)));
}

// FP32: the same median as one v_med3_f32.
template <>
__device__ __forceinline__ float tvd_slope<float>(float um, float u0, float up) {
  return __builtin_amdgcn_fmed3f(u0 - um, up - u0, 0.0f);
}

// Interface states of InterfacesKernel (L/K:1226-1252) for one face, components 0..5
// (component 6 of a face state is always replaced by KeepPositivi2Kernel).
//   dir 0: x-face between (r,c) and (r,c+1): L = W state, R = E state.
//   dir 1: y-face between (r,c) and (r+1,c): L = S state, R = N state.
// A face's loads are split from its arithmetic: the four cells of each component are
// loaded first (load_stencil), and __builtin_amdgcn_sched_barrier(0) keeps the compiler's
// scheduler from moving the arithmetic above the loads.  Left to itself, the FP64 stage
// kernel's scheduler (it keeps the kernel at 128 VGPRs) waits for each component's loads
// before issuing the next component's, six dependent memory round trips per face instead of
// one.  Callers that load more for the same face (its metric, the cell's phi) issue those
// loads before the barrier too.
template <typename T>
struct Stencil {
  T v[6][4];   // components 0..5 at i - st, i, i + st, i + 2 st
};

template <typename T>
__device__ __forceinline__ void load_stencil(const T* __restrict__ U, const Grid& g, int dir, int r,
                                             int c, Stencil<T>& s) {
  const long long st = dir == 0 ? 1 : g.nxd;
  const long long i = idx(g, r, c);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const T* a = U + k * g.P;
    s.v[k][0] = a[i - st];
    s.v[k][1] = a[i];
    s.v[k][2] = a[i + st];
    s.v[k][3] = a[i + 2 * st];
  }
}

template <typename T>
__device__ __forceinline__ void stencil_states(const Stencil<T>& s, T L[6], T R[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const T um = s.v[k][0], u0 = s.v[k][1], u1 = s.v[k][2], u2 = s.v[k][3];
    const T s0 = tvd_slope(um, u0, u1);
    const T s1 = tvd_slope(u0, u1, u2);
    L[k] = u0 + T(0.5) * s0;
    R[k] = u1 - T(0.5) * s1;
  }
}

// GROUPED = false leaves the order of the loads to the compiler (the FP32 kernels held to
// 64 VGPRs spill when a face's loads are all in flight at once).
template <typename T, bool GROUPED = true>
__device__ __forceinline__ void base_states(const T* __restrict__ U, const Grid& g, int dir, int r,
                                            int c, T L[6], T R[6]) {
  Stencil<T> s;
  load_stencil(U, g, dir, r, c, s);
  if constexpr (GROUPED) __builtin_amdgcn_sched_barrier(0);
  stencil_states(s, L, R);
}

// Face states after Interfaces2Kernel (L/K:1254-1337), taken as: all x-ghost overrides
// (columns 0-2 from 3,4; columns nx+3.. from nx+2,nx+1) from the InterfacesKernel values,
// then all y-ghost overrides (rows 0-2 from 3,4; rows ny+3.. from ny+2,ny+1) from the
// x-overridden values.  Valid for 1 <= r,c <= N-3.
template <typename T>
__device__ __forceinline__ void x_overridden(const T* __restrict__ U, const Grid& g, int dir, int r,
                                             int c, T L[6], T R[6]) {
  if (c <= 2 || c >= g.nx + MD) {
    const int ca = (c <= 2) ? MD : g.nx + MD - 1;
    const int cb = (c <= 2) ? MD + 1 : g.nx + MD - 2;
    T La[6], Ra[6], Lb[6], Rb[6];
    base_states(U, g, dir, r, ca, La, Ra);
    base_states(U, g, dir, r, cb, Lb, Rb);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      L[k] = T(0.5) * (La[k] + Lb[k]);
      R[k] = T(0.5) * (Ra[k] + Rb[k]);
    }
  } else {
    base_states(U, g, dir, r, c, L, R);
  }
}

template <typename T>
__device__ __forceinline__ void final_states_edge(const T* __restrict__ U, const Grid& g, int dir,
                                                  int r, int c, T L[6], T R[6]) {
  if (r <= 2 || r >= g.ny + MD) {
    const int ra = (r <= 2) ? MD : g.ny + MD - 1;
    const int rb = (r <= 2) ? MD + 1 : g.ny + MD - 2;
    T La[6], Ra[6], Lb[6], Rb[6];
    x_overridden(U, g, dir, ra, c, La, Ra);
    x_overridden(U, g, dir, rb, c, Lb, Rb);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      L[k] = T(0.5) * (La[k] + Lb[k]);
      R[k] = T(0.5) * (Ra[k] + Rb[k]);
    }
  } else {
    x_overridden(U, g, dir, r, c, L, R);
  }
}

// KeepPositivi1Kernel + KeepPositivi2Kernel (L/K:1339-1424) on one face state.
template <typename T>
__device__ __forceinline__ T keep_positive(T a[6]) {
  if (a[0] <= T(H00)) {
    a[0] = T(0);
    a[1] = T(0);
    a[2] = T(0);
  }
  if (a[3] <= T(H00)) {
    a[3] = T(0);
    a[4] = T(0);
    a[5] = T(0);
  }
  return qdiv(a[0], a[0] + a[3] + T(H0));
}

// Face-averaged metric of MeanKernel (L/K:1121-1224, read from the stored face planes) plus
// the cell values Flux2Kernel reads on either side (L/K:1506-1536).
template <typename T>
struct FaceMet {
  T j11, j12, j13, j21, j22, j23, j31, j32, j33;  // invJ*_avg
  T det, cav;                                    // Detmin_avg, cval_avg
  T xa, xb;      // x-face: dxdxi11_avgEW, dxdxi21_avgEW; y-face: dxdxi12_avgSN, dxdxi22_avgSN
  T sR0, sR1, cR;   // svec, cvalue of the E (x) / N (y) cell
  T sL0, sL1, cL;   // of the W / S cell
};

// The inverse Jacobian of makeTopo9Kernel (L/K:210-252) at a cell of a grid whose
// coordinate derivatives are exactly the identity (dxdxi11 = dxdxi22 = 1, dxdxi12 =
// dxdxi21 = +0, as makeTopo4Kernel gives them on the reference's uniform grid): then
// i11 = i22 = 1 and i12 = i21 = -0 (L/K:197-203), and the reference's expressions
//   invJ11 = i11 - i11 s0 s0 - i12 s1 s0,   invJ12 = i12 - i11 s0 s1 - i12 s1 s1,
//   invJ21 = i21 - i21 s0 s0 - i22 s1 s0,   invJ22 = i22 - i21 s0 s1 - i22 s1 s1,
//   invJ13 = c (s0 i11 + s1 i12),           invJ23 = c (s0 i21 + s1 i22)
// reduce, operation for operation and including the sign of a zero result, to the ones
// below: a product with -0 is a zero that cannot change a nonzero sum or difference, and
// the zero results come out +0 both ways.  The library uses them only after a set-up
// check has found them bit-identical to the stored planes wherever the step reads them.
template <typename T>
struct InvJ {
  T j11, j12, j13, j21, j22, j23;
};

template <typename T>
__device__ __forceinline__ InvJ<T> invj_uniform(T s0, T s1, T c) {
  InvJ<T> v;
  v.j11 = T(1.0) - s0 * s0;
  v.j12 = T(0.0) - s0 * s1;
  v.j21 = v.j12;
  v.j22 = T(1.0) - s1 * s1;
  v.j13 = c * (s0 + s1 * T(-0.0));
  v.j23 = c * (s0 * T(-0.0) + s1);
  return v;
}

// Where face_metric takes the face averages of MeanKernel from: FM_CELLS forms them from
// the stored cell planes (the FP32 stage: fewer bytes, and its float averages differ from
// the rounded FP64 ones by an ulp); FM_STORED reads the stored face planes; FM_UNIFORM (a
// grid whose coordinate derivatives are exactly the identity, checked at set-up) forms
// the inverse-Jacobian averages from the cells' svec and cvalue, which it reads anyway,
// takes the coordinate-derivative averages as the constants 1 and +0, and reads only
// the Detmin average.  All three give the stored values bit for bit in FP64.
enum FaceMetSrc { FM_CELLS = 0, FM_STORED = 1, FM_UNIFORM = 2 };

// dir 0: face (r,c)-(r,c+1); dir 1: face (r,c)-(r+1,c).  Valid for faces in [1,N-3]^2,
// where MeanKernel writes the averages.  face_metric_load issues the loads (raw plane
// values), face_metric_form does the arithmetic, so that a caller can issue a face's metric
// loads together with its state loads.
template <typename T, int SRC>
struct FaceMetIn {
  T sL0, sL1, cL, sR0, sR1, cR;
  // FM_STORED: the 9 face planes (FaceQ order); FM_UNIFORM: the Detmin average;
  // FM_CELLS: the cell planes at (r,c) and the neighbour, in FaceQ order, interleaved
  T a[SRC == FM_CELLS ? 2 * FQ_COUNT : SRC == FM_STORED ? FQ_COUNT : 1];
};

template <typename T, int SRC>
__device__ __forceinline__ void face_metric_load(const T* __restrict__ M, const Grid& g, int dir,
                                                 int r, int c, FaceMetIn<T, SRC>& m) {
  const long long i = idx(g, r, c);
  const long long j = i + (dir == 0 ? 1 : g.nxd);
  const long long P = g.P;
  const T* Fq = M + (dir == 0 ? MC_FX : MC_FY) * P;   // the stored face averages
  m.sL0 = M[MC_S0 * P + i];
  m.sL1 = M[MC_S1 * P + i];
  m.cL = M[MC_C * P + i];
  m.sR0 = M[MC_S0 * P + j];
  m.sR1 = M[MC_S1 * P + j];
  m.cR = M[MC_C * P + j];
  if constexpr (SRC == FM_STORED) {
#pragma unroll
    for (int q = 0; q < FQ_COUNT; q++) m.a[q] = Fq[q * P + i];
  } else if constexpr (SRC == FM_UNIFORM) {
    m.a[0] = Fq[FQ_DET * P + i];
  } else {
    const int pl[FQ_COUNT] = {MC_I11, MC_I12, MC_I13, MC_I21, MC_I22, MC_I23, MC_DET,
                              dir == 0 ? MC_X11 : MC_X12, dir == 0 ? MC_X21 : MC_X22};
#pragma unroll
    for (int q = 0; q < FQ_COUNT; q++) {
      m.a[2 * q] = M[pl[q] * P + i];
      m.a[2 * q + 1] = M[pl[q] * P + j];
    }
  }
}

template <typename T, int SRC>
__device__ __forceinline__ void face_metric_form(const FaceMetIn<T, SRC>& m, int dir, FaceMet<T>& f) {
  f.sL0 = m.sL0;
  f.sL1 = m.sL1;
  f.cL = m.cL;
  f.sR0 = m.sR0;
  f.sR1 = m.sR1;
  f.cR = m.cR;
  if constexpr (SRC == FM_STORED) {
    f.j11 = m.a[FQ_I11];
    f.j12 = m.a[FQ_I12];
    f.j13 = m.a[FQ_I13];
    f.j21 = m.a[FQ_I21];
    f.j22 = m.a[FQ_I22];
    f.j23 = m.a[FQ_I23];
    f.det = m.a[FQ_DET];
    f.xa = m.a[FQ_XA];
    f.xb = m.a[FQ_XB];
  } else if constexpr (SRC == FM_UNIFORM) {
    const InvJ<T> a = invj_uniform(f.sL0, f.sL1, f.cL), b = invj_uniform(f.sR0, f.sR1, f.cR);
    f.j11 = T(0.5) * (a.j11 + b.j11);
    f.j12 = T(0.5) * (a.j12 + b.j12);
    f.j13 = T(0.5) * (a.j13 + b.j13);
    f.j21 = T(0.5) * (a.j21 + b.j21);
    f.j22 = T(0.5) * (a.j22 + b.j22);
    f.j23 = T(0.5) * (a.j23 + b.j23);
    f.det = m.a[0];
    f.xa = dir == 0 ? T(1.0) : T(0.0);   // dxdxi11 avg (x) / dxdxi12 avg (y)
    f.xb = dir == 0 ? T(0.0) : T(1.0);   // dxdxi21 avg (x) / dxdxi22 avg (y)
  } else {
    // MeanKernel's average 0.5 * (a[i] + a[j]) of each cell plane
    f.j11 = T(0.5) * (m.a[2 * FQ_I11] + m.a[2 * FQ_I11 + 1]);
    f.j12 = T(0.5) * (m.a[2 * FQ_I12] + m.a[2 * FQ_I12 + 1]);
    f.j13 = T(0.5) * (m.a[2 * FQ_I13] + m.a[2 * FQ_I13 + 1]);
    f.j21 = T(0.5) * (m.a[2 * FQ_I21] + m.a[2 * FQ_I21 + 1]);
    f.j22 = T(0.5) * (m.a[2 * FQ_I22] + m.a[2 * FQ_I22 + 1]);
    f.j23 = T(0.5) * (m.a[2 * FQ_I23] + m.a[2 * FQ_I23 + 1]);
    f.det = T(0.5) * (m.a[2 * FQ_DET] + m.a[2 * FQ_DET + 1]);
    f.xa = T(0.5) * (m.a[2 * FQ_XA] + m.a[2 * FQ_XA + 1]);
    f.xb = T(0.5) * (m.a[2 * FQ_XB] + m.a[2 * FQ_XB + 1]);
  }
  f.j31 = T(0.5) * ((-f.sL0) + (-f.sR0));
  f.j32 = T(0.5) * ((-f.sL1) + (-f.sR1));
  f.j33 = T(0.5) * (f.cL + f.cR);
  f.cav = f.j33;
}

template <typename T, int SRC = FM_STORED>
__device__ __forceinline__ void face_metric(const T* __restrict__ M, const Grid& g, int dir, int r,
                                            int c, FaceMet<T>& f) {
  FaceMetIn<T, SRC> m;
  face_metric_load(M, g, dir, r, c, m);
  face_metric_form(m, dir, f);
}

// Wave speeds of one face (Flux1..6Kernel, L/K:1426-1848): em for the solid (emS) and the
// fluid (emF) after the 1e-10 floor.  L/R are final states; they leave with positivity
// applied.  Also returns the quantities the flux needs.
template <typename T>
struct FaceSpeed {
  T phiL, phiR;
  T qL, qR, qFL, qFR;     // normal contravariant speeds (q_xi for x, q_et for y)
  T ap, apF;              // apEW/apSN, apFEW/apFSN
  T emS, emF;
  bool finite;            // all four one-sided speeds finite (the maxima above drop a NaN)
};

template <typename T>
__device__ __forceinline__ void face_speed(int dir, T L[6], T R[6], const FaceMet<T>& f,
                                           const Phys& ph, FaceSpeed<T>& o) {
  o.phiR = keep_positive(R);
  o.phiL = keep_positive(L);
  const T h0 = T(H0);
  // Flux1Kernel L/K:1448-1470
  const T vxR = qdiv(R[1], R[0] + h0), vyR = qdiv(R[2], R[0] + h0);
  const T vxL = qdiv(L[1], L[0] + h0), vyL = qdiv(L[2], L[0] + h0);
  const T vxFR = qdiv(R[4], R[3] + h0), vyFR = qdiv(R[5], R[3] + h0);
  const T vxFL = qdiv(L[4], L[3] + h0), vyFL = qdiv(L[5], L[3] + h0);
  // Flux2Kernel L/K:1506-1536
  const T wR = qdiv(f.sR0 * vxR + f.sR1 * vyR, f.cR);
  const T wL = qdiv(f.sL0 * vxL + f.sL1 * vyL, f.cL);
  const T wFR = qdiv(f.sR0 * vxFR + f.sR1 * vyFR, f.cR);
  const T wFL = qdiv(f.sL0 * vxFL + f.sL1 * vyFL, f.cL);
  // Flux3Kernel L/K:1593-1676 (only the normal component enters the flux)
  T a1, a2, a3, an;
  if (dir == 0) {
    a1 = f.j11; a2 = f.j12; a3 = f.j13; an = f.j11;
  } else {
    a1 = f.j21; a2 = f.j22; a3 = f.j23; an = f.j22;
  }
  o.qR = a1 * vxR + a2 * vyR + a3 * wR;
  o.qL = a1 * vxL + a2 * vyL + a3 * wL;
  o.qFR = a1 * vxFR + a2 * vyFR + a3 * wFR;
  o.qFL = a1 * vxFL + a2 * vyFL + a3 * wFL;
  const T NpR = (R[0] + R[3]) * f.cav;
  const T NpL = (L[0] + L[3]) * f.cav;
  const T Drtio1 = T(ph.Drtio1);
  // Flux4Kernel L/K:1715-1737
  const T apR = fabs(o.qR) + an * qsqrt(NpR * Drtio1);
  const T apL = fabs(o.qL) + an * qsqrt(NpL * Drtio1);
  const T apFR = fabs(o.qFR) + an * qsqrt(qdiv(NpR, T(1.0) - o.phiR + h0));
  const T apFL = fabs(o.qFL) + an * qsqrt(qdiv(NpL, T(1.0) - o.phiL + h0));
  o.finite = isfinite(apR) && isfinite(apL) && isfinite(apFR) && isfinite(apFL);
  // Flux5Kernel L/K:1762-1766 (max(E, W) / max(N, S))
  o.ap = fmax(apR, apL);
  o.apF = fmax(apFR, apFL);
  // Flux6Kernel L/K:1803-1807
  o.emS = fmax(T(1.0e-10), o.ap);
  o.emF = fmax(T(1.0e-10), o.apF);
}

// Numerical flux of one face: Hp (6), Pp (components 1,2,4,5) and PD (4,5).
template <typename T>
struct FaceFlux {
  T H[6];
  T P1, P2, P4, P5;
  T D4, D5;
};

// Cell quotients the viscous term needs on both sides of the face (Flux8Kernel
// L/K:2182-2205), each over the fluid depth u3 + H0: the fluid momenta u4, u5 and their
// limited slopes along the face (uy for x-faces, ux for y-faces).  They depend on the cell
// only, so a cell shared by two faces can hand the same values to both.
template <typename T>
struct PDSide {
  T q4, q5;   // u4 / (u3 + H0), u5 / (u3 + H0)
  T s4, s5;   // slope(u4) / (u3 + H0), slope(u5) / (u3 + H0)
};

template <typename T>
__device__ __forceinline__ void face_flux(int dir, const T L[6], const T R[6], const FaceMet<T>& f,
                                          const FaceSpeed<T>& s, const PDSide<T>& pl,
                                          const PDSide<T>& pr, const Grid& g, const Phys& ph,
                                          FaceFlux<T>& o) {
  const T M = (dir == 0) ? f.j11 * f.j11 + f.j12 * f.j12 + f.j13 * f.j13
                         : f.j21 * f.j21 + f.j22 * f.j22 + f.j23 * f.j23;
  const T Drtio1 = T(ph.Drtio1);
  // Flux6Kernel L/K:1809-1843
  const T epsilon0 = T(1.0);
  const T kk = T(0.5) * epsilon0 * f.cav * f.det * M;
  const T czwR = R[0] * (R[0] + R[3]) * (kk)*Drtio1;
  const T czwL = L[0] * (L[0] + L[3]) * (kk)*Drtio1;
  const T czwFR = (R[0] + R[3]) * (R[0] + R[3]) * (kk);
  const T czwFL = (L[0] + L[3]) * (L[0] + L[3]) * (kk);
  // Flux7Kernel L/K:1883-1932
  T FR[6], FL[6];
  FR[0] = R[0] * s.qR * f.det;
  FR[1] = R[1] * s.qR * f.det + czwR * f.xa;
  FR[2] = R[2] * s.qR * f.det + czwR * f.xb;
  FR[3] = R[3] * s.qFR * f.det;
  FR[4] = R[4] * s.qFR * f.det + czwFR * f.xa;
  FR[5] = R[5] * s.qFR * f.det + czwFR * f.xb;
  FL[0] = L[0] * s.qL * f.det;
  FL[1] = L[1] * s.qL * f.det + czwL * f.xa;
  FL[2] = L[2] * s.qL * f.det + czwL * f.xb;
  FL[3] = L[3] * s.qFL * f.det;
  FL[4] = L[4] * s.qFL * f.det + czwFL * f.xa;
  FL[5] = L[5] * s.qFL * f.det + czwFL * f.xb;
  // Flux8Kernel L/K:2081-2128: Hp = 0.5*(F_E + F_W - a*Det*(u_E - u_W))
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const T a = (k < 3) ? s.ap : s.apF;
    o.H[k] = T(0.5) * (FR[k] + FL[k] - a * f.det * (R[k] - L[k]));
  }
  // Flux8Kernel L/K:2131-2179: non-conservative pressure terms
  const T cc = T(0.5) * epsilon0 * f.cav * f.det;
  const T hh = T(0.5) * ((R[0] + R[3]) + (L[0] + L[3]));
  const T pa = (dir == 0) ? f.j11 : f.j21;
  const T pb = (dir == 0) ? f.j12 : f.j22;
  const T rr = T(ph.rr);
  o.P1 = -rr * (cc)*hh * hh * pa;
  o.P2 = -rr * (cc)*hh * hh * pb;
  o.P4 = (cc)*hh * hh * pa;
  o.P5 = (cc)*hh * hh * pb;
  // Flux8Kernel L/K:2182-2205: fluid velocity gradients at the face
  const T dx = T(g.dx), dy = T(g.dy);
  T dudx, dvdx, dudy, dvdy;
  if (dir == 0) {
    dudx = qdiv(pr.q4 - pl.q4, dx);
    dvdx = qdiv(pr.q5 - pl.q5, dx);
    dudy = qdiv(T(0.5) * (pr.s4 + pl.s4), dy);
    dvdy = qdiv(T(0.5) * (pr.s5 + pl.s5), dy);
  } else {
    dudx = qdiv(T(0.5) * (pr.s4 + pl.s4), dx);
    dvdx = qdiv(T(0.5) * (pr.s5 + pl.s5), dx);
    dudy = qdiv(pr.q4 - pl.q4, dy);
    dvdy = qdiv(pr.q5 - pl.q5, dy);
  }
  // Flux9Kernel L/K:2242-2273
  const T duxidxi = f.j11 * dudx + f.j12 * dvdx - qdiv(f.j13 * (f.j31 * dudx + f.j32 * dvdx), f.j33);
  const T dvetdxi = f.j21 * dudx + f.j22 * dvdx - qdiv(f.j23 * (f.j31 * dudx + f.j32 * dvdx), f.j33);
  const T duxidet = f.j11 * dudy + f.j12 * dvdy - qdiv(f.j13 * (f.j31 * dudy + f.j32 * dvdy), f.j33);
  const T dvetdet = f.j21 * dudy + f.j22 * dvdy - qdiv(f.j23 * (f.j31 * dudy + f.j32 * dvdy), f.j33);
  // Flux10Kernel L/K:2316-2343
  if (dir == 0) {
    o.D4 = T(2.0) * f.det * (hh) * (f.j11 * duxidxi + f.j21 * duxidet);
    o.D5 = f.det * (hh) * (f.j12 * duxidxi + f.j22 * duxidet + f.j11 * dvetdxi + f.j21 * dvetdet);
  } else {
    o.D4 = f.det * (hh) * (f.j12 * duxidxi + f.j22 * duxidet + f.j11 * dvetdxi + f.j21 * dvetdet);
    o.D5 = T(2.0) * f.det * (hh) * (f.j12 * dvetdxi + f.j22 * dvetdet);
  }
}

// Cell metric for the source terms.
template <typename T>
struct CellMet {
  T c, s0, s1, det;
  T i11, i12, i13, i21, i22, i23;
  T j13x, j23x, j33x, j13e, j23e, j33e;
};

// UNI: the inverse Jacobian from svec and cvalue (invj_uniform) instead of its planes.
// cell_metric_load issues the loads, cell_metric_form the arithmetic (as for faces).
template <typename T, bool UNI = false>
__device__ __forceinline__ void cell_metric_load(const T* __restrict__ M, long long P, long long i,
                                                 CellMet<T>& m) {
  m.c = M[MC_C * P + i];
  m.s0 = M[MC_S0 * P + i];
  m.s1 = M[MC_S1 * P + i];
  m.det = M[MC_DET * P + i];
  if constexpr (!UNI) {
    m.i11 = M[MC_I11 * P + i];
    m.i12 = M[MC_I12 * P + i];
    m.i13 = M[MC_I13 * P + i];
    m.i21 = M[MC_I21 * P + i];
    m.i22 = M[MC_I22 * P + i];
    m.i23 = M[MC_I23 * P + i];
  }
  m.j13x = M[MC_J13X * P + i];
  m.j23x = M[MC_J23X * P + i];
  m.j33x = M[MC_J33X * P + i];
  m.j13e = M[MC_J13E * P + i];
  m.j23e = M[MC_J23E * P + i];
  m.j33e = M[MC_J33E * P + i];
}

template <typename T, bool UNI = false>
__device__ __forceinline__ void cell_metric_form(CellMet<T>& m) {
  if constexpr (UNI) {
    const InvJ<T> v = invj_uniform(m.s0, m.s1, m.c);
    m.i11 = v.j11;
    m.i12 = v.j12;
    m.i13 = v.j13;
    m.i21 = v.j21;
    m.i22 = v.j22;
    m.i23 = v.j23;
  }
}

template <typename T, bool UNI = false>
__device__ __forceinline__ void cell_metric(const T* __restrict__ M, long long P, long long i,
                                            CellMet<T>& m) {
  cell_metric_load<T, UNI>(M, P, i, m);
  cell_metric_form<T, UNI>(m);
}

// Flux11..17Kernel (L/K:2349-2621): the source vector s (components 1,2,4,5; s0 = s3 = 0).
template <typename T>
__device__ __forceinline__ void cell_source(const T u[7], const CellMet<T>& m, const Phys& ph,
                                            T s[6]) {
  const T h = u[0], hf = u[3];
  const T h0 = T(H0), u0 = T(U0);
  T vex, vey, vexF, veyF;
  if (h <= T(H00)) {  // Flux11
    vex = T(0);
    vey = T(0);
  } else {
    vex = qdiv(u[1], h + h0);
    vey = qdiv(u[2], h + h0);
  }
  if (hf <= T(H00)) {
    vexF = T(0);
    veyF = T(0);
  } else {
    vexF = qdiv(u[4], hf + h0);
    veyF = qdiv(u[5], hf + h0);
  }
  const T w = qdiv(m.s0 * vex + m.s1 * vey, m.c);  // Flux12
  const T wF = qdiv(m.s0 * vexF + m.s1 * veyF, m.c);
  const T usw = vex * vex + vey * vey + w * w;  // Flux13
  T vel = qsqrt(vex * vex + vey * vey + w * w);
  T vexw, veyw;
  if (usw <= u0 * u0) {  // Flux14
    vexw = qdiv(vex, u0);
    veyw = qdiv(vey, u0);
    vel = u0;
  } else {
    vexw = qdiv(vex, vel);
    veyw = qdiv(vey, vel);
  }
  const T q_xi = m.i11 * vex + m.i12 * vey + m.i13 * w;
  const T q_et = m.i21 * vex + m.i22 * vey + m.i23 * w;
  const T q_xiF = m.i11 * vexF + m.i12 * veyF + m.i13 * wF;
  const T q_etF = m.i21 * vexF + m.i22 * veyF + m.i23 * wF;
  // Flux15
  const T Ac = vex * m.j13x * q_xi + vey * m.j23x * q_xi + w * m.j33x * q_xi +
               vex * m.j13e * q_et + vey * m.j23e * q_et + w * m.j33e * q_et;
  const T AcF = vexF * m.j13x * q_xiF + veyF * m.j23x * q_xiF + wF * m.j33x * q_xiF +
                vexF * m.j13e * q_etF + veyF * m.j23e * q_etF + wF * m.j33e * q_etF;
  // Flux16
  const T D = m.det;
  const T Drtio1 = T(ph.Drtio1), rr = T(ph.rr), Cd = T(ph.Cd), tande = T(ph.tande);
  const T Npress1 = fmax(T(0), (D * h * (Drtio1 * m.c - Ac + AcF * rr)));
  const T Npress2 = fmax(T(0), (D * h * (m.c - Ac)));
  const T NpressF = fmax(T(0), (D * hf * (m.c - AcF)));
  // Flux17
  const T phi = u[6];
  const T epsilon0 = T(1.0);
  const T Dratio = rr;
  const T varTheta = T(ph.varTheta), N_R = T(ph.N_R);
  s[0] = T(0);
  s[1] = -Npress2 * m.s0 - Npress1 * tande * vexw +
         Dratio * D * Cd * phi * (T(1.0) - phi) * (h + hf) * (vexF - vex);
  s[2] = -Npress2 * m.s1 - Npress1 * tande * veyw +
         Dratio * D * Cd * phi * (T(1.0) - phi) * (h + hf) * (veyF - vey);
  s[3] = T(0);
  s[4] = -NpressF * m.s0 - D * Cd * phi * (T(1.0) - phi) * (h + hf) * (vexF - vex) +
         qdiv(D * qdiv(-(T(1.0) - phi) * (h + hf) * varTheta * vexF, epsilon0), N_R);
  s[5] = -NpressF * m.s1 - D * Cd * phi * (T(1.0) - phi) * (h + hf) * (veyF - vey) +
         qdiv(D * qdiv(-(T(1.0) - phi) * (h + hf) * varTheta * veyF, epsilon0), N_R);
}

}  // namespace m2p
