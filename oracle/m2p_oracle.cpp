// m2p_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
//
// Serial FP64 CPU restatement of MoSES_2PDF's adNOC_2Pb time-step loop, landslide and
// inflow modes.  Every function below follows one CUDA kernel of the reference, over
// exactly that kernel's index range, in the launch order of RunKernels::kernelStep.
// Paths are relative to the MoSES_2PDF source tree:
//   L/K = MoSES_2PDF_landslide/adNOC_2Pb_kernels.cu     L/R = .../adNOC_2Pb_runKernels.cu
//   I/K = MoSES_2PDF_inflow/adNOC_2Pb_kernels.cu        I/R = .../adNOC_2Pb_runKernels.cu
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
// library.  It is built with -O2 -ffp-contract=off so that every expression is evaluated
// with the reference's operand order and one rounding per operation.
//
// Deterministic resolutions of the reference's undefined behaviour (cudaMalloc memory is
// never cleared there, and three kernels race):
//   * every device array starts at 0.0 (uninitialised reads: inflow u, the outer two
//     rings of cvalue in makeTopo10, the face-averaged metric outside [1,N-3], Val's last
//     row and column);
//   * Interfaces2Kernel (L/K:1255-1337): all x-ghost overrides first, reading the values
//     InterfacesKernel wrote, then all y-ghost overrides, reading the x-overridden values;
//   * the inflow Boundary3Kernel (I/K:704-815): x copies first, then y copies;
//   * CFL4Kernel (L/K:2023-2036): TotalTime advances by dt once per step.
//
// How it is pinned (tests/test_oracle_pins.py): closed forms of the metric on planes and
// valleys, textbook minmod, exact interface states of linear data, the lake at rest, mass
// conservation, mirror symmetry, the closed-form dt, the sliding-block acceleration
// (gravity, Coulomb friction, buoyancy), the midpoint rule of a fluid layer's
// gravity-resistance ODE, momentum exchange by the phase drag, the shallow-water fluxes
// and wave speeds of both phases, the central-upwind combination, the fluid shear
// viscosity on flat and tilted ground, the viscous fluxes as the depth-integrated
// Newtonian film stress on planes of any tilt (vector geometry: the factor 2 of the normal
// stresses and every term of Flux9's transform), the centripetal term of the basal loads
// as the second fundamental form of the bed on quadratic beds (Flux15-16), and the DF004
// injection strips.  Parity unpinned: only what the CUDA reference on NVIDIA hardware
// could show (the racy orders above, uninitialised memory, device tan).

#include <algorithm>
#include <cmath>
#include <omp.h>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

constexpr int MD = 3;           // L/KH:7
constexpr int MN = 7;           // L/KH:8
constexpr double H0 = 1.0e-6;   // L/KH:11
constexpr double H00 = 0.0;     // L/KH:12
constexpr double U0 = 1.0e-6;   // L/KH:13
constexpr double PI_REF = 3.14159;  // L/KH:17 (used in tan(delta0*PI/180))

using Vec = std::vector<double>;

struct Orc {
  // ---- configuration ----
  int mode = 0;  // 0 landslide, 1 inflow
  int nx = 0, ny = 0, nxd = 0, nyd = 0;
  size_t P = 0;
  double dx = 0, dy = 0;  // scaled (x10)
  double delta0 = 0, Cd = 0, N_R = 0, varTheta = 0, phiS0 = 0;
  double cfl = 0.1, rr = 1.420 / 2.60, dtlo = 0.001, dthi = 2.0;

  // ---- inflow tables (I/R:297-331, 394-428) ----
  int nloc = 0, nsamp = 0;
  Vec inflow, loc, inflowTime;
  std::vector<int> dire;
  int inflowCount = 0;

  // ---- host loop state (L/R:382-393) ----
  Vec tf;          // tf[0..outputStep]
  int iter = 1;
  int schreiben = 0;
  int nstop = 0;
  int nt = 0;
  double TotalTime = 0, dt = 0, dtval = 0, dmax = 0, stepStartTime = 0;
  int lastWritten = 0;  // iter of the output produced by the last step (0: none)

  // ---- inputs ----
  Vec topo, depth, phicell;

  // ---- terrain metric ----
  Vec bfkt, posx, posy, dxdxi11, dxdxi12, dxdxi21, dxdxi22, dbdx, dbdy, cvalue, svec, Jacb31,
      Jacb32, dettmp, Detmin, i11, i12, i21, i22, invJ11, invJ12, invJ13, invJ21, invJ22, invJ23,
      invJ31, invJ32, invJ33, tande, J13dxi, J23dxi, J33dxi, J13det, J23det, J33det;
  Vec dxdxi11_avgEW, dxdxi21_avgEW, dxdxi12_avgSN, dxdxi22_avgSN;
  Vec invJ11_avgEW, invJ12_avgEW, invJ13_avgEW, invJ21_avgEW, invJ22_avgEW, invJ23_avgEW,
      invJ31_avgEW, invJ32_avgEW, invJ33_avgEW;
  Vec invJ11_avgSN, invJ12_avgSN, invJ13_avgSN, invJ21_avgSN, invJ22_avgSN, invJ23_avgSN,
      invJ31_avgSN, invJ32_avgSN, invJ33_avgSN;
  Vec Detmin_avgEW, Detmin_avgSN, cval_avgEW, cval_avgSN;

  // ---- state (7 planes) ----
  Vec u, uzero, uone, utwo, v;
  // ---- stage temporaries ----
  Vec dux, duy, sgnAx, sgnBx, sgnAy, sgnBy, t1x, t2x, t1y, t2y, ux, uy;
  Vec uE, uW, uN, uS;
  Vec Hpx, Hpy, Ppx, Ppy, PDx, PDy, FpE, FpW, GpN, GpS, s;
  Vec vexE, veyE, vexW, veyW, vexFE, veyFE, vexFW, veyFW;
  Vec vexN, veyN, vexS, veyS, vexFN, veyFN, vexFS, veyFS;
  Vec wE, wW, wFE, wFW, wN, wS, wFN, wFS;
  Vec q_xiE, q_etE, q_xiW, q_etW, q_xiFE, q_etFE, q_xiFW, q_etFW;
  Vec q_xiN, q_etN, q_xiS, q_etS, q_xiFN, q_etFN, q_xiFS, q_etFS;
  Vec NpressFE, NpressFW, M11EW, NpressFN, NpressFS, M22SN;
  Vec apE, apW, apFE, apFW, apN, apS, apFN, apFS, apEW, apFEW, apSN, apFSN;
  Vec em_x, em_y, em_Fx, em_Fy, czw1x, czw2x, czwF1x, czwF2x, czw1y, czw2y, czwF1y, czwF2y;
  Vec em_valS, em_valF, Val;
  Vec dudxE, dvdxE, dudyE, dvdyE, dudxN, dvdxN, dudyN, dvdyN;
  Vec duxidxix, dvetdxix, duxidetx, dvetdetx, duxidxiy, dvetdxiy, duxidety, dvetdety;
  Vec vex, vey, vexF, veyF, w_wert, w_wertF, usw, vel, vexw, veyw, q_xi, q_et, q_xiF, q_etF;
  Vec Ac, AcF, Npress1, Npress2, NpressF;
  Vec usxnew, ufxnew, usxold, ufxold, usynew, ufynew, usyold, ufyold;

  size_t I(int r, int c) const { return (size_t)r * nxd + c; }
  size_t K(int k) const { return (size_t)k * P; }

  void alloc() {
    P = (size_t)nxd * nyd;
    Vec* planes1[] = {&posx, &posy, &dxdxi11, &dxdxi12, &dxdxi21, &dxdxi22, &dbdx, &dbdy,
                      &cvalue, &Jacb31, &Jacb32, &dettmp, &Detmin, &i11, &i12, &i21, &i22,
                      &invJ11, &invJ12, &invJ13, &invJ21, &invJ22, &invJ23, &invJ31, &invJ32,
                      &invJ33, &tande, &J13dxi, &J23dxi, &J33dxi, &J13det, &J23det, &J33det,
                      &dxdxi11_avgEW, &dxdxi21_avgEW, &dxdxi12_avgSN, &dxdxi22_avgSN,
                      &invJ11_avgEW, &invJ12_avgEW, &invJ13_avgEW, &invJ21_avgEW, &invJ22_avgEW,
                      &invJ23_avgEW, &invJ31_avgEW, &invJ32_avgEW, &invJ33_avgEW,
                      &invJ11_avgSN, &invJ12_avgSN, &invJ13_avgSN, &invJ21_avgSN, &invJ22_avgSN,
                      &invJ23_avgSN, &invJ31_avgSN, &invJ32_avgSN, &invJ33_avgSN,
                      &Detmin_avgEW, &Detmin_avgSN, &cval_avgEW, &cval_avgSN,
                      &vexE, &veyE, &vexW, &veyW, &vexFE, &veyFE, &vexFW, &veyFW,
                      &vexN, &veyN, &vexS, &veyS, &vexFN, &veyFN, &vexFS, &veyFS,
                      &wE, &wW, &wFE, &wFW, &wN, &wS, &wFN, &wFS,
                      &q_xiE, &q_etE, &q_xiW, &q_etW, &q_xiFE, &q_etFE, &q_xiFW, &q_etFW,
                      &q_xiN, &q_etN, &q_xiS, &q_etS, &q_xiFN, &q_etFN, &q_xiFS, &q_etFS,
                      &NpressFE, &NpressFW, &M11EW, &NpressFN, &NpressFS, &M22SN,
                      &apE, &apW, &apFE, &apFW, &apN, &apS, &apFN, &apFS, &apEW, &apFEW,
                      &apSN, &apFSN, &em_x, &em_y, &em_Fx, &em_Fy, &czw1x, &czw2x, &czwF1x,
                      &czwF2x, &czw1y, &czw2y, &czwF1y, &czwF2y, &em_valS, &em_valF, &Val,
                      &dudxE, &dvdxE, &dudyE, &dvdyE, &dudxN, &dvdxN, &dudyN, &dvdyN,
                      &duxidxix, &dvetdxix, &duxidetx, &dvetdetx, &duxidxiy, &dvetdxiy,
                      &duxidety, &dvetdety, &vex, &vey, &vexF, &veyF, &w_wert, &w_wertF, &usw,
                      &vel, &vexw, &veyw, &q_xi, &q_et, &q_xiF, &q_etF, &Ac, &AcF, &Npress1,
                      &Npress2, &NpressF, &usxnew, &ufxnew, &usxold, &ufxold, &usynew, &ufynew,
                      &usyold, &ufyold};
    for (Vec* p : planes1) p->assign(P, 0.0);
    svec.assign(2 * P, 0.0);
    bfkt.assign(3 * P, 0.0);
    Vec* planes7[] = {&u,   &uzero, &uone, &utwo, &v,   &dux, &duy, &sgnAx, &sgnBx, &sgnAy,
                      &sgnBy, &t1x, &t2x,  &t1y,  &t2y, &ux,  &uy,  &uE,    &uW,    &uN,
                      &uS,  &Hpx, &Hpy,  &Ppx,  &Ppy, &PDx, &PDy, &FpE,   &FpW,   &GpN,
                      &GpS, &s};
    for (Vec* p : planes7) p->assign(7 * P, 0.0);
  }

  // =====================================================================================
  // Terrain metric, once per run (L/R:395-642)
  // =====================================================================================

  // makeTopo1Kernel L/K:5-23.  Coordinates; writes spill into the next row / plane and
  // are overwritten by makeTopo2.
  void makeTopo1() {
    const double xmin = 0.0, ymin = 0.0;  // MINX, MINY (L/R:184-185)
    for (int row = 0; row < nyd; row++)
      for (int col = 0; col < nxd; col++) {
        bfkt[0 * P + (size_t)row * nxd + (col + MD)] = xmin + col * dx;
        bfkt[1 * P + (size_t)(row + MD) * nxd + col] = ymin + row * dy;
      }
  }
  // makeTopo2Kernel L/K:25-56
  void makeTopo2() {
    for (int row = 0; row < nyd; row++)
      for (int col = 0; col < nxd; col++) {
        double* b0 = &bfkt[0];
        double* b1 = &bfkt[P];
        b0[I(row, 0)] = b0[I(row, MD)] - (MD - 0) * dx;
        b0[I(row, 1)] = b0[I(row, MD)] - (MD - 1) * dx;
        b0[I(row, 2)] = b0[I(row, MD)] - (MD - 2) * dx;
        b1[I(0, col)] = b1[I(MD, col)] - (MD - 0) * dy;
        b1[I(1, col)] = b1[I(MD, col)] - (MD - 1) * dy;
        b1[I(2, col)] = b1[I(MD, col)] - (MD - 2) * dy;
        b0[I(row, nx + MD + 0)] = b0[I(row, nx + MD - 1)] + (1 + 0) * dx;
        b0[I(row, nx + MD + 1)] = b0[I(row, nx + MD - 1)] + (1 + 1) * dx;
        b0[I(row, nx + MD + 2)] = b0[I(row, nx + MD - 1)] + (1 + 2) * dx;
        b1[I(ny + MD + 0, col)] = b1[I(ny + MD - 1, col)] + (1 + 0) * dy;
        b1[I(ny + MD + 1, col)] = b1[I(ny + MD - 1, col)] + (1 + 1) * dy;
        b1[I(ny + MD + 2, col)] = b1[I(ny + MD - 1, col)] + (1 + 2) * dy;
        bfkt[2 * P + I(row, col)] = topo[I(row, col)];
      }
  }
  // makeTopo3Kernel L/K:58-73
  void makeTopo3() {
    for (size_t i = 0; i < P; i++) {
      posx[i] = bfkt[i];
      posy[i] = bfkt[P + i];
    }
  }
  // Three-point non-uniform derivative used by makeTopo4Kernel (L/K:88-110) and
  // JacobKernel (L/K:674-696): f at (m, 0, p), abscissae (xm, x0, xp).
  static double d3(double fm, double f0, double fp, double xm, double x0, double xp) {
    return fm * (x0 - xp) / ((xm - x0) * (xm - xp)) +
           f0 * ((xm - x0) - (x0 - xp)) / ((xm - x0) * (x0 - xp)) -
           fp * (xm - x0) / ((xm - xp) * (x0 - xp));
  }
  // makeTopo4Kernel L/K:75-114, range [2, N-3]^2
  void makeTopo4() {
    const double* b0 = &bfkt[0];
    const double* b1 = &bfkt[P];
    const double* b2 = &bfkt[2 * P];
    for (int row = 2; row < nyd - 2; row++)
      for (int col = 2; col < nxd - 2; col++) {
        size_t i = I(row, col), w = I(row, col - 1), e = I(row, col + 1), sN = I(row - 1, col),
               n = I(row + 1, col);
        dxdxi11[i] = d3(b0[w], b0[i], b0[e], posx[w], posx[i], posx[e]);
        dxdxi12[i] = d3(b0[sN], b0[i], b0[n], posy[sN], posy[i], posy[n]);
        dxdxi21[i] = d3(b1[w], b1[i], b1[e], posx[w], posx[i], posx[e]);
        dxdxi22[i] = d3(b1[sN], b1[i], b1[n], posy[sN], posy[i], posy[n]);
        dbdx[i] = d3(b2[w], b2[i], b2[e], b0[w], b0[i], b0[e]);
        dbdy[i] = d3(b2[sN], b2[i], b2[n], b1[sN], b1[i], b1[n]);
      }
  }
  // makeTopo5..9Kernel L/K:116-252, range [2, N-3]^2
  void makeTopo5to9() {
    for (int row = 2; row < nyd - 2; row++)
      for (int col = 2; col < nxd - 2; col++) {
        size_t i = I(row, col);
        cvalue[i] = 1.0 / std::sqrt(1.0 + dbdx[i] * dbdx[i] + dbdy[i] * dbdy[i]);  // topo5
      }
    for (int row = 2; row < nyd - 2; row++)
      for (int col = 2; col < nxd - 2; col++) {
        size_t i = I(row, col);
        svec[i] = cvalue[i] * dbdx[i];  // topo6
        svec[P + i] = cvalue[i] * dbdy[i];
      }
    for (int row = 2; row < nyd - 2; row++)
      for (int col = 2; col < nxd - 2; col++) {
        size_t i = I(row, col);
        const double s0 = svec[i], s1 = svec[P + i];
        Jacb31[i] = (s0 * dxdxi11[i] + s1 * dxdxi21[i]) / cvalue[i];  // topo7
        Jacb32[i] = (s0 * dxdxi12[i] + s1 * dxdxi22[i]) / cvalue[i];
        dettmp[i] = dxdxi11[i] * dxdxi22[i] - dxdxi21[i] * dxdxi12[i];
      }
    for (int row = 2; row < nyd - 2; row++)
      for (int col = 2; col < nxd - 2; col++) {
        size_t i = I(row, col);
        const double s0 = svec[i], s1 = svec[P + i], c = cvalue[i];
        Detmin[i] = dxdxi11[i] * (dxdxi22[i] * c - (-s1) * Jacb32[i]) -   // topo8
                    dxdxi21[i] * (dxdxi12[i] * c - Jacb32[i] * (-s0)) +
                    Jacb31[i] * (dxdxi12[i] * (-s1) - dxdxi22[i] * (-s0));
        i11[i] = dxdxi22[i] / dettmp[i];
        i12[i] = -dxdxi12[i] / dettmp[i];
        i21[i] = -dxdxi21[i] / dettmp[i];
        i22[i] = dxdxi11[i] / dettmp[i];
      }
    for (int row = 2; row < nyd - 2; row++)
      for (int col = 2; col < nxd - 2; col++) {
        size_t i = I(row, col);
        const double s0 = svec[i], s1 = svec[P + i], c = cvalue[i];
        invJ11[i] = i11[i] - i11[i] * s0 * s0 - i12[i] * s1 * s0;  // topo9
        invJ12[i] = i12[i] - i11[i] * s0 * s1 - i12[i] * s1 * s1;
        invJ21[i] = i21[i] - i21[i] * s0 * s0 - i22[i] * s1 * s0;
        invJ22[i] = i22[i] - i21[i] * s0 * s1 - i22[i] * s1 * s1;
        invJ13[i] = c * (s0 * i11[i] + s1 * i12[i]);
        invJ23[i] = c * (s0 * i21[i] + s1 * i22[i]);
        invJ31[i] = -s0;
        invJ32[i] = -s1;
        invJ33[i] = c;
      }
  }
  // makeTopo10Kernel L/K:254-280 (landslide only; L/R:479).  Uses cvalue before its
  // ghost copy, so the outer two rings read the initial 0.
  void makeTopo10() {
    for (size_t i = 0; i < P; i++) {
      const double phi = phicell.empty() ? phiS0 : phicell[i];
      u[K(0) + i] = depth[i] * phi * cvalue[i];
      u[K(1) + i] = 0.0;
      u[K(2) + i] = 0.0;
      u[K(3) + i] = depth[i] * (1.0 - phi) * cvalue[i];
      u[K(4) + i] = 0.0;
      u[K(5) + i] = 0.0;
      u[K(6) + i] = phi;
    }
  }
  // makeTopo11Kernel L/K:282-302 (inflow I/K: no clamp).  tan() of the reference's
  // 3.14159 is evaluated once on the host.
  void makeTopo11() {
    const double t = std::tan(delta0 * PI_REF / 180.);
    for (size_t i = 0; i < P; i++) {
      if (mode == 0) {
        u[K(0) + i] = std::fmax(0.0, u[K(0) + i]);
        u[K(3) + i] = std::fmax(0.0, u[K(3) + i]);
      }
      tande[i] = t;
    }
  }
  // Boundary1Kernel L/K:304-479 (inflow: all rows, I/K:321) and Boundary2Kernel
  // L/K:482-658: ghost copies of 17 metric planes.  dxdxi21 keeps the reference's
  // double write of column nx+MD+1 and never writes nx+MD+2 (L/K:407-409).
  void boundary12() {
    std::vector<double*> pl = {&dxdxi11[0], &dxdxi12[0], &dxdxi21[0], &dxdxi22[0], &cvalue[0],
                               &Detmin[0],  &svec[0],    &svec[P],    &Jacb31[0],  &Jacb32[0],
                               &invJ11[0],  &invJ12[0],  &invJ13[0],  &invJ21[0],  &invJ22[0],
                               &invJ23[0],  &invJ31[0],  &invJ32[0],  &invJ33[0]};
    const int r0 = (mode == 0) ? 3 : 0, r1 = (mode == 0) ? nyd - MD : nyd;
    for (size_t a = 0; a < pl.size(); a++) {
      double* A = pl[a];
      const bool quirk = (A == &dxdxi21[0]);
      for (int row = r0; row < r1; row++) {
        A[I(row, 0)] = A[I(row, MD)];
        A[I(row, 1)] = A[I(row, MD)];
        A[I(row, 2)] = A[I(row, MD)];
        A[I(row, nx + MD + 0)] = A[I(row, nx + MD - 1)];
        A[I(row, nx + MD + 1)] = A[I(row, nx + MD - 1)];
        if (!quirk) A[I(row, nx + MD + 2)] = A[I(row, nx + MD - 1)];
      }
    }
    for (size_t a = 0; a < pl.size(); a++) {
      double* A = pl[a];
      for (int col = 0; col < nxd; col++) {
        A[I(0, col)] = A[I(MD, col)];
        A[I(1, col)] = A[I(MD, col)];
        A[I(2, col)] = A[I(MD, col)];
        A[I(ny + MD + 0, col)] = A[I(ny + MD - 1, col)];
        A[I(ny + MD + 1, col)] = A[I(ny + MD - 1, col)];
        A[I(ny + MD + 2, col)] = A[I(ny + MD - 1, col)];
      }
    }
  }
  // JacobKernel L/K:660-701, range [3, N-4]^2
  void jacob() {
    for (int row = 3; row < nyd - MD; row++)
      for (int col = 3; col < nxd - MD; col++) {
        size_t i = I(row, col), w = I(row, col - 1), e = I(row, col + 1), sN = I(row - 1, col),
               n = I(row + 1, col);
        const double* s0 = &svec[0];
        const double* s1 = &svec[P];
        J13dxi[i] = d3(-s0[w], -s0[i], -s0[e], posx[w], posx[i], posx[e]);
        J23dxi[i] = d3(-s1[w], -s1[i], -s1[e], posx[w], posx[i], posx[e]);
        J33dxi[i] = d3(cvalue[w], cvalue[i], cvalue[e], posx[w], posx[i], posx[e]);
        J13det[i] = d3(-s0[sN], -s0[i], -s0[n], posy[sN], posy[i], posy[n]);
        J23det[i] = d3(-s1[sN], -s1[i], -s1[n], posy[sN], posy[i], posy[n]);
        J33det[i] = d3(cvalue[sN], cvalue[i], cvalue[n], posy[sN], posy[i], posy[n]);
      }
  }
  // Boundary3Kernel L/K:703-765 (rows [3,N-4]) + Boundary4Kernel L/K:767-829; the inflow
  // merged Boundary3Kernel (I/K:704-815) gives the same values when x copies run first.
  void boundary34() {
    Vec* pl[] = {&J13dxi, &J23dxi, &J33dxi, &J13det, &J23det, &J33det};
    const int r0 = (mode == 0) ? 3 : 0, r1 = (mode == 0) ? nyd - MD : nyd;
    for (Vec* p : pl) {
      double* A = &(*p)[0];
      for (int row = r0; row < r1; row++)
        for (int k = 0; k < 3; k++) {
          A[I(row, k)] = A[I(row, MD)];
          A[I(row, nx + MD + k)] = A[I(row, nx + MD - 1)];
        }
      for (int col = 0; col < nxd; col++)
        for (int k = 0; k < 3; k++) {
          A[I(k, col)] = A[I(MD, col)];
          A[I(ny + MD + k, col)] = A[I(ny + MD - 1, col)];
        }
    }
  }
  // MeanKernel L/K:1121-1224, range [1, N-3]^2.  The J*_avg and svec_avg outputs are
  // never read by the time loop and are not kept.
  void mean() {
    for (int row = 1; row < nyd - 2; row++)
      for (int col = 1; col < nxd - 2; col++) {
        size_t i = I(row, col), e = I(row, col + 1), n = I(row + 1, col);
        dxdxi11_avgEW[i] = 0.5 * (dxdxi11[i] + dxdxi11[e]);
        dxdxi21_avgEW[i] = 0.5 * (dxdxi21[i] + dxdxi21[e]);
        dxdxi12_avgSN[i] = 0.5 * (dxdxi12[i] + dxdxi12[n]);
        dxdxi22_avgSN[i] = 0.5 * (dxdxi22[i] + dxdxi22[n]);
        invJ11_avgEW[i] = 0.5 * (invJ11[i] + invJ11[e]);
        invJ12_avgEW[i] = 0.5 * (invJ12[i] + invJ12[e]);
        invJ13_avgEW[i] = 0.5 * (invJ13[i] + invJ13[e]);
        invJ21_avgEW[i] = 0.5 * (invJ21[i] + invJ21[e]);
        invJ22_avgEW[i] = 0.5 * (invJ22[i] + invJ22[e]);
        invJ23_avgEW[i] = 0.5 * (invJ23[i] + invJ23[e]);
        invJ31_avgEW[i] = 0.5 * (invJ31[i] + invJ31[e]);
        invJ32_avgEW[i] = 0.5 * (invJ32[i] + invJ32[e]);
        invJ33_avgEW[i] = 0.5 * (invJ33[i] + invJ33[e]);
        invJ11_avgSN[i] = 0.5 * (invJ11[i] + invJ11[n]);
        invJ12_avgSN[i] = 0.5 * (invJ12[i] + invJ12[n]);
        invJ13_avgSN[i] = 0.5 * (invJ13[i] + invJ13[n]);
        invJ21_avgSN[i] = 0.5 * (invJ21[i] + invJ21[n]);
        invJ22_avgSN[i] = 0.5 * (invJ22[i] + invJ22[n]);
        invJ23_avgSN[i] = 0.5 * (invJ23[i] + invJ23[n]);
        invJ31_avgSN[i] = 0.5 * (invJ31[i] + invJ31[n]);
        invJ32_avgSN[i] = 0.5 * (invJ32[i] + invJ32[n]);
        invJ33_avgSN[i] = 0.5 * (invJ33[i] + invJ33[n]);
        Detmin_avgEW[i] = 0.5 * (Detmin[i] + Detmin[e]);
        Detmin_avgSN[i] = 0.5 * (Detmin[i] + Detmin[n]);
        cval_avgEW[i] = 0.5 * (cvalue[i] + cvalue[e]);
        cval_avgSN[i] = 0.5 * (cvalue[i] + cvalue[n]);
      }
  }

  void setup() {
    makeTopo1();
    makeTopo2();
    makeTopo3();
    makeTopo4();
    makeTopo5to9();
    if (mode == 0) makeTopo10();
    makeTopo11();
    boundary12();
    jacob();
    boundary34();
    mean();
  }

  // =====================================================================================
  // Time loop kernels (L/R:646-1330)
  // =====================================================================================

  // Inflow1Kernel I/K:818-850 and Inflow2Kernel I/K:853-887
  void inflowInject() {
    for (int i = 0; i < nloc; i++) {
      const int direX = dire[i * 3 + 1], direY = dire[i * 3 + 2];
      const int xloc = (int)loc[i * 3 + 1], yloc = (int)loc[i * 3 + 2];
      const double Hin = inflow[0 * (size_t)nsamp * nloc + (size_t)inflowCount * nloc + i];
      for (int mm = 0; mm < direX + MD + 3; mm++)
        for (int nn = 0; nn < direY + MD + 3; nn++) {
          size_t c = I(yloc + nn, xloc + mm);
          u[K(0) + c] = Hin * (1 - phiS0) / phiS0;
          u[K(3) + c] = Hin * (1.0 - phiS0) * cvalue[c];
        }
    }
    for (int i = 0; i < nloc; i++) {
      const int direX = dire[i * 3 + 1], direY = dire[i * 3 + 2];
      const int xloc = (int)loc[i * 3 + 1], yloc = (int)loc[i * 3 + 2];
      const double Uin = inflow[1 * (size_t)nsamp * nloc + (size_t)inflowCount * nloc + i];
      const double Vin = inflow[2 * (size_t)nsamp * nloc + (size_t)inflowCount * nloc + i];
      for (int mm = 0; mm < direX + MD + 3; mm++)
        for (int nn = 0; nn < direY + MD + 3; nn++) {
          size_t c = I(yloc + nn, xloc + mm);
          u[K(1) + c] = Uin * (u[K(0) + c] + H0);
          u[K(2) + c] = Vin * (u[K(0) + c] + H0);
          u[K(4) + c] = Uin * (u[K(3) + c] + H0);
          u[K(5) + c] = Vin * (u[K(3) + c] + H0);
          u[K(6) + c] = u[K(0) + c] / (u[K(0) + c] + u[K(3) + c] + H0);
        }
    }
  }

  // UzeroKernel L/K:832-851
  void uzeroCopy() {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 3; row < nyd - MD; row++)
        for (int col = 3; col < nxd - MD; col++) uzero[K(k) + I(row, col)] = u[K(k) + I(row, col)];
  }
  // Boundary5Kernel L/K:853-876: x-ghost extrapolation, every row, sequential per row.
  void boundary5() {
    for (int row = 0; row < nyd; row++)
      for (int k = 0; k < MN; k++) {
        double* a = &u[K(k) + I(row, 0)];
        a[MD - 0 - 1] = 0.5 * (a[MD - 0] + a[MD - 0 + 1]);
        a[MD - 1 - 1] = 0.5 * (a[MD - 1] + a[MD - 1 + 1]);
        a[MD - 2 - 1] = 0.5 * (a[MD - 2] + a[MD - 2 + 1]);
        a[nx + MD + 0] = 0.5 * (a[nx + MD + 0 - 1] + a[nx + MD + 0 - 2]);
        a[nx + MD + 1] = 0.5 * (a[nx + MD + 1 - 1] + a[nx + MD + 1 - 2]);
        a[nx + MD + 2] = 0.5 * (a[nx + MD + 2 - 1] + a[nx + MD + 2 - 2]);
      }
  }
  // Boundary6Kernel L/K:878-902 (columns [3,N-4]); inflow all columns (I/K:943).
  void boundary6() {
    const int c0 = (mode == 0) ? 3 : 0, c1 = (mode == 0) ? nxd - MD : nxd;
    for (int col = c0; col < c1; col++)
      for (int k = 0; k < MN; k++) {
        double* a = &u[K(k) + col];
        auto R = [&](int r) -> double& { return a[(size_t)r * nxd]; };
        R(MD - 0 - 1) = 0.5 * (R(MD - 0) + R(MD - 0 + 1));
        R(MD - 1 - 1) = 0.5 * (R(MD - 1) + R(MD - 1 + 1));
        R(MD - 2 - 1) = 0.5 * (R(MD - 2) + R(MD - 2 + 1));
        R(ny + MD + 0) = 0.5 * (R(ny + MD + 0 - 1) + R(ny + MD + 0 - 2));
        R(ny + MD + 1) = 0.5 * (R(ny + MD + 1 - 1) + R(ny + MD + 1 - 2));
        R(ny + MD + 2) = 0.5 * (R(ny + MD + 2 - 1) + R(ny + MD + 2 - 2));
      }
  }
  // Boundary7Kernel L/K:904-929 and Boundary8Kernel L/K:931-956 (landslide; columns
  // [3,N-4]); inflow merged Boundary7Kernel I/K:962-997 (all rows and columns).
  void boundary78() {
    const int mom[4] = {1, 2, 4, 5};
    for (int row = 0; row < nyd; row++)
      for (int i = 0; i < MD; i++)
        for (int k : mom) {
          u[K(k) + I(row, MD - i - 1)] = 0.0;
          u[K(k) + I(row, nx + MD + i)] = 0.0;
        }
    const int c0 = (mode == 0) ? 3 : 0, c1 = (mode == 0) ? nxd - MD : nxd;
    for (int col = c0; col < c1; col++)
      for (int i = 0; i < MD; i++)
        for (int k : mom) {
          u[K(k) + I(MD - i - 1, col)] = 0.0;
          u[K(k) + I(ny + MD + i, col)] = 0.0;
        }
  }
  // Boundary9Kernel L/K:959-997
  void boundary9() {
    Vec* p7[] = {&Hpx, &Hpy, &Ppx, &Ppy, &PDx, &PDy, &ux, &uy};
    for (Vec* p : p7) std::fill(p->begin(), p->end(), 0.0);
    std::fill(apEW.begin(), apEW.end(), 0.0);
    std::fill(apSN.begin(), apSN.end(), 0.0);
    std::fill(apFEW.begin(), apFEW.end(), 0.0);
    std::fill(apFSN.begin(), apFSN.end(), 0.0);
  }
  // TVD1..5Kernel L/K:1000-1119: minmod(minmod(a,b), central) slopes.
  void tvd() {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 0; row < nyd - 1; row++)
        for (int col = 0; col < nxd - 1; col++) {
          size_t i = K(k) + I(row, col);
          dux[i] = u[K(k) + I(row, col + 1)] - u[i];  // TVD1
          duy[i] = u[K(k) + I(row + 1, col)] - u[i];
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 1; row < nyd - 1; row++)
        for (int col = 1; col < nxd - 1; col++) {
          size_t i = K(k) + I(row, col), w = K(k) + I(row, col - 1), sN = K(k) + I(row - 1, col);
          sgnAx[i] = (dux[w] > 0.0) ? 1.0 : -1.0;  // TVD2
          sgnBx[i] = (dux[i] > 0.0) ? 1.0 : -1.0;
          sgnAy[i] = (duy[sN] > 0.0) ? 1.0 : -1.0;
          sgnBy[i] = (duy[i] > 0.0) ? 1.0 : -1.0;
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 1; row < nyd - 1; row++)
        for (int col = 1; col < nxd - 1; col++) {
          size_t i = K(k) + I(row, col), w = K(k) + I(row, col - 1), sN = K(k) + I(row - 1, col);
          t1x[i] = 0.5 * (sgnAx[i] + sgnBx[i]) * std::fmin(dux[w] * sgnAx[i], dux[i] * sgnBx[i]);
          t2x[i] = 0.5 * (dux[w] + dux[i]);  // TVD3
          t1y[i] = 0.5 * (sgnAy[i] + sgnBy[i]) * std::fmin(duy[sN] * sgnAy[i], duy[i] * sgnBy[i]);
          t2y[i] = 0.5 * (duy[sN] + duy[i]);
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 1; row < nyd - 1; row++)
        for (int col = 1; col < nxd - 1; col++) {
          size_t i = K(k) + I(row, col);
          sgnAx[i] = (t1x[i] > 0.0) ? 1.0 : -1.0;  // TVD4
          sgnBx[i] = (t2x[i] > 0.0) ? 1.0 : -1.0;
          sgnAy[i] = (t1y[i] > 0.0) ? 1.0 : -1.0;
          sgnBy[i] = (t2y[i] > 0.0) ? 1.0 : -1.0;
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 1; row < nyd - 1; row++)
        for (int col = 1; col < nxd - 1; col++) {
          size_t i = K(k) + I(row, col);
          ux[i] = 0.5 * (sgnAx[i] + sgnBx[i]) * std::fmin(t1x[i] * sgnAx[i], t2x[i] * sgnBx[i]);
          uy[i] = 0.5 * (sgnAy[i] + sgnBy[i]) * std::fmin(t1y[i] * sgnAy[i], t2y[i] * sgnBy[i]);
        }
  }
  // InterfacesKernel L/K:1226-1252, range [0, N-2]^2
  void interfaces() {
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 0; row < nyd - 1; row++)
        for (int col = 0; col < nxd - 1; col++) {
          size_t i = K(k) + I(row, col), e = K(k) + I(row, col + 1), n = K(k) + I(row + 1, col);
          uE[i] = u[e] - 0.5 * ux[e];
          uW[i] = u[i] + 0.5 * ux[i];
          uN[i] = u[n] - 0.5 * uy[n];
          uS[i] = u[i] + 0.5 * uy[i];
        }
  }
  // Interfaces2Kernel L/K:1254-1337, racy in the reference; resolved as all x-ghost
  // writes (from the InterfacesKernel values) and then all y-ghost writes.
  void interfaces2() {
    Vec* F[4] = {&uE, &uW, &uN, &uS};
    for (Vec* f : F)
      for (int k = 0; k < MN; k++)
        for (int row = 0; row < nyd; row++) {
          double* a = &(*f)[K(k) + I(row, 0)];
          const double left = 0.5 * (a[MD] + a[MD + 1]);
          const double right = 0.5 * (a[nx + MD - 1] + a[nx + MD - 2]);
          a[0] = left;
          a[1] = left;
          a[2] = left;
          a[nx + MD + 0] = right;
          a[nx + MD + 1] = right;
          a[nx + MD + 2] = right;
        }
    for (Vec* f : F)
      for (int k = 0; k < MN; k++)
        for (int col = 0; col < nxd; col++) {
          double* a = &(*f)[K(k) + col];
          auto R = [&](int r) -> double& { return a[(size_t)r * nxd]; };
          const double top = 0.5 * (R(MD) + R(MD + 1));
          const double bot = 0.5 * (R(ny + MD - 1) + R(ny + MD - 2));
          R(0) = top;
          R(1) = top;
          R(2) = top;
          R(ny + MD + 0) = bot;
          R(ny + MD + 1) = bot;
          R(ny + MD + 2) = bot;
        }
  }
  // KeepPositivi1Kernel L/K:1339-1403 and KeepPositivi2Kernel L/K:1405-1424
  void keepPositivity() {
    Vec* F[4] = {&uE, &uW, &uN, &uS};
    for (Vec* f : F)
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < P; i++) {
        Vec& a = *f;
        if (a[K(0) + i] <= H00) {
          a[K(0) + i] = 0.0;
          a[K(1) + i] = 0.0;
          a[K(2) + i] = 0.0;
        }
        if (a[K(3) + i] <= H00) {
          a[K(3) + i] = 0.0;
          a[K(4) + i] = 0.0;
          a[K(5) + i] = 0.0;
        }
      }
    Vec* G[4] = {&uE, &uW, &uS, &uN};
    for (Vec* f : G)
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < P; i++) {
        Vec& a = *f;
        a[K(6) + i] = a[K(0) + i] / (a[K(0) + i] + a[K(3) + i] + H0);
      }
  }
  // Flux1..7Kernel L/K:1426-1936 and CFL1/CFL2Kernel L/K:1938-1974, range [0, N-2]^2
  void faceFluxes() {
    const double Drtio1 = 1.0 - rr;
    const double epsilon0 = 1.0;
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nyd - 1; row++)
      for (int col = 0; col < nxd - 1; col++) {
        const size_t i = I(row, col);
        auto E = [&](int k) { return uE[K(k) + i]; };
        auto W = [&](int k) { return uW[K(k) + i]; };
        auto N = [&](int k) { return uN[K(k) + i]; };
        auto S = [&](int k) { return uS[K(k) + i]; };
        // Flux1
        vexE[i] = E(1) / (E(0) + H0);
        veyE[i] = E(2) / (E(0) + H0);
        vexW[i] = W(1) / (W(0) + H0);
        veyW[i] = W(2) / (W(0) + H0);
        vexFE[i] = E(4) / (E(3) + H0);
        veyFE[i] = E(5) / (E(3) + H0);
        vexFW[i] = W(4) / (W(3) + H0);
        veyFW[i] = W(5) / (W(3) + H0);
        vexN[i] = N(1) / (N(0) + H0);
        veyN[i] = N(2) / (N(0) + H0);
        vexS[i] = S(1) / (S(0) + H0);
        veyS[i] = S(2) / (S(0) + H0);
        vexFN[i] = N(4) / (N(3) + H0);
        veyFN[i] = N(5) / (N(3) + H0);
        vexFS[i] = S(4) / (S(3) + H0);
        veyFS[i] = S(5) / (S(3) + H0);
        // Flux2
        const size_t e = I(row, col + 1), n = I(row + 1, col);
        const double* s0 = &svec[0];
        const double* s1 = &svec[P];
        wE[i] = (s0[e] * vexE[i] + s1[e] * veyE[i]) / cvalue[e];
        wW[i] = (s0[i] * vexW[i] + s1[i] * veyW[i]) / cvalue[i];
        wFE[i] = (s0[e] * vexFE[i] + s1[e] * veyFE[i]) / cvalue[e];
        wFW[i] = (s0[i] * vexFW[i] + s1[i] * veyFW[i]) / cvalue[i];
        wN[i] = (s0[n] * vexN[i] + s1[n] * veyN[i]) / cvalue[n];
        wS[i] = (s0[i] * vexS[i] + s1[i] * veyS[i]) / cvalue[i];
        wFN[i] = (s0[n] * vexFN[i] + s1[n] * veyFN[i]) / cvalue[n];
        wFS[i] = (s0[i] * vexFS[i] + s1[i] * veyFS[i]) / cvalue[i];
        // Flux3
        const double a11 = invJ11_avgEW[i], a12 = invJ12_avgEW[i], a13 = invJ13_avgEW[i];
        const double a21 = invJ21_avgEW[i], a22 = invJ22_avgEW[i], a23 = invJ23_avgEW[i];
        const double b11 = invJ11_avgSN[i], b12 = invJ12_avgSN[i], b13 = invJ13_avgSN[i];
        const double b21 = invJ21_avgSN[i], b22 = invJ22_avgSN[i], b23 = invJ23_avgSN[i];
        q_xiE[i] = a11 * vexE[i] + a12 * veyE[i] + a13 * wE[i];
        q_etE[i] = a21 * vexE[i] + a22 * veyE[i] + a23 * wE[i];
        q_xiW[i] = a11 * vexW[i] + a12 * veyW[i] + a13 * wW[i];
        q_etW[i] = a21 * vexW[i] + a22 * veyW[i] + a23 * wW[i];
        q_xiFE[i] = a11 * vexFE[i] + a12 * veyFE[i] + a13 * wFE[i];
        q_etFE[i] = a21 * vexFE[i] + a22 * veyFE[i] + a23 * wFE[i];
        q_xiFW[i] = a11 * vexFW[i] + a12 * veyFW[i] + a13 * wFW[i];
        q_etFW[i] = a21 * vexFW[i] + a22 * veyFW[i] + a23 * wFW[i];
        NpressFE[i] = (E(0) + E(3)) * cval_avgEW[i];
        NpressFW[i] = (W(0) + W(3)) * cval_avgEW[i];
        M11EW[i] = a11 * a11 + a12 * a12 + a13 * a13;
        q_xiN[i] = b11 * vexN[i] + b12 * veyN[i] + b13 * wN[i];
        q_etN[i] = b21 * vexN[i] + b22 * veyN[i] + b23 * wN[i];
        q_xiS[i] = b11 * vexS[i] + b12 * veyS[i] + b13 * wS[i];
        q_etS[i] = b21 * vexS[i] + b22 * veyS[i] + b23 * wS[i];
        q_xiFN[i] = b11 * vexFN[i] + b12 * veyFN[i] + b13 * wFN[i];
        q_etFN[i] = b21 * vexFN[i] + b22 * veyFN[i] + b23 * wFN[i];
        q_xiFS[i] = b11 * vexFS[i] + b12 * veyFS[i] + b13 * wFS[i];
        q_etFS[i] = b21 * vexFS[i] + b22 * veyFS[i] + b23 * wFS[i];
        NpressFN[i] = (N(0) + N(3)) * cval_avgSN[i];
        NpressFS[i] = (S(0) + S(3)) * cval_avgSN[i];
        M22SN[i] = b21 * b21 + b22 * b22 + b23 * b23;
        // Flux4
        apE[i] = std::fabs(q_xiE[i]) + a11 * std::sqrt(NpressFE[i] * Drtio1);
        apW[i] = std::fabs(q_xiW[i]) + a11 * std::sqrt(NpressFW[i] * Drtio1);
        apFE[i] = std::fabs(q_xiFE[i]) + a11 * std::sqrt(NpressFE[i] / (1.0 - E(6) + H0));
        apFW[i] = std::fabs(q_xiFW[i]) + a11 * std::sqrt(NpressFW[i] / (1.0 - W(6) + H0));
        apN[i] = std::fabs(q_etN[i]) + b22 * std::sqrt(NpressFN[i] * Drtio1);
        apS[i] = std::fabs(q_etS[i]) + b22 * std::sqrt(NpressFS[i] * Drtio1);
        apFN[i] = std::fabs(q_etFN[i]) + b22 * std::sqrt(NpressFN[i] / (1.0 - N(6) + H0));
        apFS[i] = std::fabs(q_etFS[i]) + b22 * std::sqrt(NpressFS[i] / (1.0 - S(6) + H0));
        // Flux5
        apEW[i] = std::fmax(apE[i], apW[i]);
        apFEW[i] = std::fmax(apFE[i], apFW[i]);
        apSN[i] = std::fmax(apN[i], apS[i]);
        apFSN[i] = std::fmax(apFN[i], apFS[i]);
        // Flux6
        em_x[i] = std::fmax(1.0e-10, apEW[i]);
        em_y[i] = std::fmax(1.0e-10, apSN[i]);
        em_Fx[i] = std::fmax(1.0e-10, apFEW[i]);
        em_Fy[i] = std::fmax(1.0e-10, apFSN[i]);
        const double kx = 0.5 * epsilon0 * cval_avgEW[i] * Detmin_avgEW[i] * M11EW[i];
        const double ky = 0.5 * epsilon0 * cval_avgSN[i] * Detmin_avgSN[i] * M22SN[i];
        czw1x[i] = E(0) * (E(0) + E(3)) * (kx)*Drtio1;
        czw2x[i] = W(0) * (W(0) + W(3)) * (kx)*Drtio1;
        czwF1x[i] = (E(0) + E(3)) * (E(0) + E(3)) * (kx);
        czwF2x[i] = (W(0) + W(3)) * (W(0) + W(3)) * (kx);
        czw1y[i] = N(0) * (N(0) + N(3)) * (ky)*Drtio1;
        czw2y[i] = S(0) * (S(0) + S(3)) * (ky)*Drtio1;
        czwF1y[i] = (N(0) + N(3)) * (N(0) + N(3)) * (ky);
        czwF2y[i] = (S(0) + S(3)) * (S(0) + S(3)) * (ky);
        // Flux7
        const double dEW = Detmin_avgEW[i], dSN = Detmin_avgSN[i];
        const double x11 = dxdxi11_avgEW[i], x21 = dxdxi21_avgEW[i];
        const double x12 = dxdxi12_avgSN[i], x22 = dxdxi22_avgSN[i];
        FpE[K(0) + i] = E(0) * q_xiE[i] * dEW;
        FpE[K(1) + i] = E(1) * q_xiE[i] * dEW + czw1x[i] * x11;
        FpE[K(2) + i] = E(2) * q_xiE[i] * dEW + czw1x[i] * x21;
        FpE[K(3) + i] = E(3) * q_xiFE[i] * dEW;
        FpE[K(4) + i] = E(4) * q_xiFE[i] * dEW + czwF1x[i] * x11;
        FpE[K(5) + i] = E(5) * q_xiFE[i] * dEW + czwF1x[i] * x21;
        FpW[K(0) + i] = W(0) * q_xiW[i] * dEW;
        FpW[K(1) + i] = W(1) * q_xiW[i] * dEW + czw2x[i] * x11;
        FpW[K(2) + i] = W(2) * q_xiW[i] * dEW + czw2x[i] * x21;
        FpW[K(3) + i] = W(3) * q_xiFW[i] * dEW;
        FpW[K(4) + i] = W(4) * q_xiFW[i] * dEW + czwF2x[i] * x11;
        FpW[K(5) + i] = W(5) * q_xiFW[i] * dEW + czwF2x[i] * x21;
        GpN[K(0) + i] = N(0) * q_etN[i] * dSN;
        GpN[K(1) + i] = N(1) * q_etN[i] * dSN + czw1y[i] * x12;
        GpN[K(2) + i] = N(2) * q_etN[i] * dSN + czw1y[i] * x22;
        GpN[K(3) + i] = N(3) * q_etFN[i] * dSN;
        GpN[K(4) + i] = N(4) * q_etFN[i] * dSN + czwF1y[i] * x12;
        GpN[K(5) + i] = N(5) * q_etFN[i] * dSN + czwF1y[i] * x22;
        GpS[K(0) + i] = S(0) * q_etS[i] * dSN;
        GpS[K(1) + i] = S(1) * q_etS[i] * dSN + czw2y[i] * x12;
        GpS[K(2) + i] = S(2) * q_etS[i] * dSN + czw2y[i] * x22;
        GpS[K(3) + i] = S(3) * q_etFS[i] * dSN;
        GpS[K(4) + i] = S(4) * q_etFS[i] * dSN + czwF2y[i] * x12;
        GpS[K(5) + i] = S(5) * q_etFS[i] * dSN + czwF2y[i] * x22;
        // CFL1, CFL2
        em_valS[i] = std::fmax(em_x[i] / dx, em_y[i] / dy);
        em_valF[i] = std::fmax(em_Fx[i] / dx, em_Fy[i] / dy);
        Val[i] = std::fmax(em_valS[i], em_valF[i]);
      }
  }
  // reduceKernel L/K:1977-2004 (fmaxf: the max is rounded to float) and CFL3Kernel
  // L/K:2006-2020 / I/K:2050-2063
  void cflReduce() {
    // fmaxf over all entries; per-thread maxima combined with fmaxf give the same value in
    // any order (NaN is ignored unless every entry is NaN)
    float m = 0.0f;
#pragma omp parallel
    {
      float t = 0.0f;
#pragma omp for schedule(static) nowait
      for (size_t i = 0; i < P; i++) t = std::fmax(t, (float)Val[i]);
#pragma omp critical
      m = std::fmax(m, t);
    }
    dmax = (double)m;
    dtval = cfl / dmax;
    dtval = std::fmin(dthi, dtval);
    dtval = std::fmax(dtlo, dtval);
  }
  // Flux8..10Kernel L/K:2038-2347, range [1, N-2]^2
  void numFluxes() {
    const double epsilon0 = 1.0;
    const double Dratio = rr;
#pragma omp parallel for schedule(static)
    for (int row = 1; row < nyd - 1; row++)
      for (int col = 1; col < nxd - 1; col++) {
        const size_t i = I(row, col);
        auto E = [&](int k) { return uE[K(k) + i]; };
        auto W = [&](int k) { return uW[K(k) + i]; };
        auto N = [&](int k) { return uN[K(k) + i]; };
        auto S = [&](int k) { return uS[K(k) + i]; };
        const double dEW = Detmin_avgEW[i], dSN = Detmin_avgSN[i];
        for (int k = 0; k < 6; k++) {
          const double ax = (k < 3) ? apEW[i] : apFEW[i];
          const double ay = (k < 3) ? apSN[i] : apFSN[i];
          Hpx[K(k) + i] = 0.5 * (FpE[K(k) + i] + FpW[K(k) + i] - ax * dEW * (E(k) - W(k)));
          Hpy[K(k) + i] = 0.5 * (GpN[K(k) + i] + GpS[K(k) + i] - ay * dSN * (N(k) - S(k)));
        }
        const double cx = 0.5 * epsilon0 * cval_avgEW[i] * dEW;
        const double hx = 0.5 * ((E(0) + E(3)) + (W(0) + W(3)));
        Ppx[K(0) + i] = 0.0;
        Ppx[K(1) + i] = -Dratio * (cx)*hx * hx * invJ11_avgEW[i];
        Ppx[K(2) + i] = -Dratio * (cx)*hx * hx * invJ12_avgEW[i];
        Ppx[K(3) + i] = 0.0;
        Ppx[K(4) + i] = (cx)*hx * hx * invJ11_avgEW[i];
        Ppx[K(5) + i] = (cx)*hx * hx * invJ12_avgEW[i];
        const double cy = 0.5 * epsilon0 * cval_avgSN[i] * dSN;
        const double hy = 0.5 * ((N(0) + N(3)) + (S(0) + S(3)));
        Ppy[K(0) + i] = 0.0;
        Ppy[K(1) + i] = -Dratio * (cy)*hy * hy * invJ21_avgSN[i];
        Ppy[K(2) + i] = -Dratio * (cy)*hy * hy * invJ22_avgSN[i];
        Ppy[K(3) + i] = 0.0;
        Ppy[K(4) + i] = (cy)*hy * hy * invJ21_avgSN[i];
        Ppy[K(5) + i] = (cy)*hy * hy * invJ22_avgSN[i];
        const size_t e = I(row, col + 1), n = I(row + 1, col);
        const double* u3 = &u[K(3)];
        const double* u4 = &u[K(4)];
        const double* u5 = &u[K(5)];
        dudxE[i] = (u4[e] / (u3[e] + H0) - u4[i] / (u3[i] + H0)) / dx;
        dvdxE[i] = (u5[e] / (u3[e] + H0) - u5[i] / (u3[i] + H0)) / dx;
        dudyE[i] = 0.5 * (uy[K(4) + e] / (u3[e] + H0) + uy[K(4) + i] / (u3[i] + H0)) / dy;
        dvdyE[i] = 0.5 * (uy[K(5) + e] / (u3[e] + H0) + uy[K(5) + i] / (u3[i] + H0)) / dy;
        dudxN[i] = 0.5 * (ux[K(4) + n] / (u3[n] + H0) + ux[K(4) + i] / (u3[i] + H0)) / dx;
        dvdxN[i] = 0.5 * (ux[K(5) + n] / (u3[n] + H0) + ux[K(5) + i] / (u3[i] + H0)) / dx;
        dudyN[i] = (u4[n] / (u3[n] + H0) - u4[i] / (u3[i] + H0)) / dy;
        dvdyN[i] = (u5[n] / (u3[n] + H0) - u5[i] / (u3[i] + H0)) / dy;
      }
#pragma omp parallel for schedule(static)
    for (int row = 1; row < nyd - 1; row++)
      for (int col = 1; col < nxd - 1; col++) {
        const size_t i = I(row, col);
        // Flux9
        {
          const double a11 = invJ11_avgEW[i], a12 = invJ12_avgEW[i], a13 = invJ13_avgEW[i];
          const double a21 = invJ21_avgEW[i], a22 = invJ22_avgEW[i], a23 = invJ23_avgEW[i];
          const double a31 = invJ31_avgEW[i], a32 = invJ32_avgEW[i], a33 = invJ33_avgEW[i];
          duxidxix[i] = a11 * dudxE[i] + a12 * dvdxE[i] - a13 * (a31 * dudxE[i] + a32 * dvdxE[i]) / (a33);
          dvetdxix[i] = a21 * dudxE[i] + a22 * dvdxE[i] - a23 * (a31 * dudxE[i] + a32 * dvdxE[i]) / a33;
          duxidetx[i] = a11 * dudyE[i] + a12 * dvdyE[i] - a13 * (a31 * dudyE[i] + a32 * dvdyE[i]) / a33;
          dvetdetx[i] = a21 * dudyE[i] + a22 * dvdyE[i] - a23 * (a31 * dudyE[i] + a32 * dvdyE[i]) / a33;
        }
        {
          const double b11 = invJ11_avgSN[i], b12 = invJ12_avgSN[i], b13 = invJ13_avgSN[i];
          const double b21 = invJ21_avgSN[i], b22 = invJ22_avgSN[i], b23 = invJ23_avgSN[i];
          const double b31 = invJ31_avgSN[i], b32 = invJ32_avgSN[i], b33 = invJ33_avgSN[i];
          duxidxiy[i] = b11 * dudxN[i] + b12 * dvdxN[i] - b13 * (b31 * dudxN[i] + b32 * dvdxN[i]) / b33;
          dvetdxiy[i] = b21 * dudxN[i] + b22 * dvdxN[i] - b23 * (b31 * dudxN[i] + b32 * dvdxN[i]) / b33;
          duxidety[i] = b11 * dudyN[i] + b12 * dvdyN[i] - b13 * (b31 * dudyN[i] + b32 * dvdyN[i]) / b33;
          dvetdety[i] = b21 * dudyN[i] + b22 * dvdyN[i] - b23 * (b31 * dudyN[i] + b32 * dvdyN[i]) / b33;
        }
      }
#pragma omp parallel for schedule(static)
    for (int row = 1; row < nyd - 1; row++)
      for (int col = 1; col < nxd - 1; col++) {
        const size_t i = I(row, col);
        const double hx = 0.5 * ((uE[K(0) + i] + uE[K(3) + i]) + (uW[K(0) + i] + uW[K(3) + i]));
        const double hy = 0.5 * ((uN[K(0) + i] + uN[K(3) + i]) + (uS[K(0) + i] + uS[K(3) + i]));
        const double dEW = Detmin_avgEW[i], dSN = Detmin_avgSN[i];
        for (int k = 0; k < 4; k++) {  // Flux10
          PDx[K(k) + i] = 0.0;
          PDy[K(k) + i] = 0.0;
        }
        PDx[K(4) + i] = 2.0 * dEW * (hx) * (invJ11_avgEW[i] * duxidxix[i] + invJ21_avgEW[i] * duxidetx[i]);
        PDx[K(5) + i] = dEW * (hx) *
                        (invJ12_avgEW[i] * duxidxix[i] + invJ22_avgEW[i] * duxidetx[i] +
                         invJ11_avgEW[i] * dvetdxix[i] + invJ21_avgEW[i] * dvetdetx[i]);
        PDy[K(4) + i] = dSN * (hy) *
                        (invJ12_avgSN[i] * duxidxiy[i] + invJ22_avgSN[i] * duxidety[i] +
                         invJ11_avgSN[i] * dvetdxiy[i] + invJ21_avgSN[i] * dvetdety[i]);
        PDy[K(5) + i] = 2.0 * dSN * (hy) * (invJ12_avgSN[i] * dvetdxiy[i] + invJ22_avgSN[i] * dvetdety[i]);
      }
  }
  // Flux11..17Kernel L/K:2349-2621 (landslide [1,N-2]^2; inflow [0,N-2]^2, I/K:2404)
  void sources() {
    const double Drtio1 = 1.0 - rr;
    const double Dratio = rr;
    const double epsilon0 = 1.0;
    const int lo = (mode == 0) ? 1 : 0;
#pragma omp parallel for schedule(static)
    for (int row = lo; row < nyd - 1; row++)
      for (int col = lo; col < nxd - 1; col++) {
        const size_t i = I(row, col);
        const double h = u[K(0) + i], hf = u[K(3) + i];
        if (h <= H00) {  // Flux11
          vex[i] = 0.0;
          vey[i] = 0.0;
        } else {
          vex[i] = u[K(1) + i] / (h + H0);
          vey[i] = u[K(2) + i] / (h + H0);
        }
        if (hf <= H00) {
          vexF[i] = 0.0;
          veyF[i] = 0.0;
        } else {
          vexF[i] = u[K(4) + i] / (hf + H0);
          veyF[i] = u[K(5) + i] / (hf + H0);
        }
        const double s0 = svec[i], s1 = svec[P + i], c = cvalue[i];
        w_wert[i] = (s0 * vex[i] + s1 * vey[i]) / c;  // Flux12
        w_wertF[i] = (s0 * vexF[i] + s1 * veyF[i]) / c;
        usw[i] = vex[i] * vex[i] + vey[i] * vey[i] + w_wert[i] * w_wert[i];  // Flux13
        vel[i] = std::sqrt(vex[i] * vex[i] + vey[i] * vey[i] + w_wert[i] * w_wert[i]);
        if (usw[i] <= U0 * U0) {  // Flux14
          vexw[i] = vex[i] / U0;
          veyw[i] = vey[i] / U0;
          vel[i] = U0;
        } else {
          vexw[i] = vex[i] / vel[i];
          veyw[i] = vey[i] / vel[i];
        }
        q_xi[i] = invJ11[i] * vex[i] + invJ12[i] * vey[i] + invJ13[i] * w_wert[i];
        q_et[i] = invJ21[i] * vex[i] + invJ22[i] * vey[i] + invJ23[i] * w_wert[i];
        q_xiF[i] = invJ11[i] * vexF[i] + invJ12[i] * veyF[i] + invJ13[i] * w_wertF[i];
        q_etF[i] = invJ21[i] * vexF[i] + invJ22[i] * veyF[i] + invJ23[i] * w_wertF[i];
        // Flux15
        Ac[i] = vex[i] * J13dxi[i] * q_xi[i] + vey[i] * J23dxi[i] * q_xi[i] +
                w_wert[i] * J33dxi[i] * q_xi[i] + vex[i] * J13det[i] * q_et[i] +
                vey[i] * J23det[i] * q_et[i] + w_wert[i] * J33det[i] * q_et[i];
        AcF[i] = vexF[i] * J13dxi[i] * q_xiF[i] + veyF[i] * J23dxi[i] * q_xiF[i] +
                 w_wertF[i] * J33dxi[i] * q_xiF[i] + vexF[i] * J13det[i] * q_etF[i] +
                 veyF[i] * J23det[i] * q_etF[i] + w_wertF[i] * J33det[i] * q_etF[i];
        // Flux16
        const double D = Detmin[i];
        Npress1[i] = std::fmax(0.0, (D * h * (Drtio1 * c - Ac[i] + AcF[i] * Dratio)));
        Npress2[i] = std::fmax(0.0, (D * h * (c - Ac[i])));
        NpressF[i] = std::fmax(0.0, (D * hf * (c - AcF[i])));
        // Flux17
        const double phi = u[K(6) + i];
        const double td = tande[i];
        s[K(0) + i] = 0.0;
        s[K(1) + i] = -Npress2[i] * s0 - Npress1[i] * td * vexw[i] +
                      Dratio * D * Cd * phi * (1.0 - phi) * (h + hf) * (vexF[i] - vex[i]);
        s[K(2) + i] = -Npress2[i] * s1 - Npress1[i] * td * veyw[i] +
                      Dratio * D * Cd * phi * (1.0 - phi) * (h + hf) * (veyF[i] - vey[i]);
        s[K(3) + i] = 0.0;
        s[K(4) + i] = -NpressF[i] * s0 - D * Cd * phi * (1.0 - phi) * (h + hf) * (vexF[i] - vex[i]) +
                      D * (-(1.0 - phi) * (h + hf) * varTheta * vexF[i] / epsilon0) / N_R;
        s[K(5) + i] = -NpressF[i] * s1 - D * Cd * phi * (1.0 - phi) * (h + hf) * (veyF[i] - vey[i]) +
                      D * (-(1.0 - phi) * (h + hf) * varTheta * veyF[i] / epsilon0) / N_R;
      }
  }
  // Flux18..21Kernel L/K:2623-2726, range [3, N-4]^2
  void update() {
    const double epsilon0 = 1.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN - 1; k++)
      for (int row = 3; row < nyd - MD; row++)
        for (int col = 3; col < nxd - MD; col++) {
          const size_t i = I(row, col), w = I(row, col - 1), sN = I(row - 1, col), kk = K(k);
          const double phi = u[K(6) + i];
          v[kk + i] = uzero[kk + i] * Detmin[i] - (dt / dx) * (Hpx[kk + i] - Hpx[kk + w]) -
                      (dt / dy) * (Hpy[kk + i] - Hpy[kk + sN]) +
                      (dt / dx) * (Ppx[kk + i] - Ppx[kk + w]) * (epsilon0 * phi) +
                      (dt / dy) * (Ppy[kk + i] - Ppy[kk + sN]) * (epsilon0 * phi) +
                      (dt / dx) * (PDx[kk + i] - PDx[kk + w]) * (epsilon0 * (1.0 - phi) / N_R) +
                      (dt / dy) * (PDy[kk + i] - PDy[kk + sN]) * (epsilon0 * (1.0 - phi) / N_R) +
                      (dt)*s[kk + i];
        }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN - 1; k++)  // Flux19
      for (int row = 3; row < nyd - MD; row++)
        for (int col = 3; col < nxd - MD; col++) {
          const size_t i = I(row, col);
          u[K(k) + i] = v[K(k) + i] / Detmin[i];
        }
#pragma omp parallel for schedule(static)
    for (int row = 3; row < nyd - MD; row++)  // Flux20
      for (int col = 3; col < nxd - MD; col++) {
        const size_t i = I(row, col);
        if (u[K(0) + i] <= H00) {
          u[K(0) + i] = 0.0;
          u[K(1) + i] = 0.0;
          u[K(2) + i] = 0.0;
        }
        if (u[K(3) + i] <= H00) {
          u[K(3) + i] = 0.0;
          u[K(4) + i] = 0.0;
          u[K(5) + i] = 0.0;
        }
      }
#pragma omp parallel for schedule(static)
    for (int row = 3; row < nyd - MD; row++)  // Flux21
      for (int col = 3; col < nxd - MD; col++) {
        const size_t i = I(row, col);
        u[K(6) + i] = u[K(0) + i] / (u[K(0) + i] + u[K(3) + i] + H0);
      }
  }
  // Velocity-crossing limiter: Flux22..24Kernel (predictor, into uone) L/K:2730-2822 and
  // Flux25..28Kernel (corrector, into u via utwo) L/K:2826-2937.
  void crossing(int io) {
    Vec& dst = (io == 0) ? uone : utwo;
    const Vec& old = (io == 0) ? uzero : uone;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < MN; k++)
      for (int row = 3; row < nyd - MD; row++)
        for (int col = 3; col < nxd - MD; col++) dst[K(k) + I(row, col)] = u[K(k) + I(row, col)];
#pragma omp parallel for schedule(static)
    for (int row = 3; row < nyd - MD; row++)
      for (int col = 3; col < nxd - MD; col++) {
        const size_t i = I(row, col);
        usxnew[i] = u[K(1) + i] / (u[K(0) + i] + H0);
        ufxnew[i] = u[K(4) + i] / (u[K(3) + i] + H0);
        usxold[i] = old[K(1) + i] / (old[K(0) + i] + H0);
        ufxold[i] = old[K(4) + i] / (old[K(3) + i] + H0);
        usynew[i] = u[K(2) + i] / (u[K(0) + i] + H0);
        ufynew[i] = u[K(5) + i] / (u[K(3) + i] + H0);
        usyold[i] = old[K(2) + i] / (old[K(0) + i] + H0);
        ufyold[i] = old[K(5) + i] / (old[K(3) + i] + H0);
      }
#pragma omp parallel for schedule(static)
    for (int row = 3; row < nyd - MD; row++)
      for (int col = 3; col < nxd - MD; col++) {
        const size_t i = I(row, col);
        if ((usxnew[i] - ufxnew[i]) * (usxold[i] - ufxold[i]) >= 0.0)
          dst[K(4) + i] = u[K(4) + i];
        else
          dst[K(4) + i] = usxnew[i] * u[K(3) + i];
        if ((usynew[i] - ufynew[i]) * (usyold[i] - ufyold[i]) >= 0.0)
          dst[K(5) + i] = u[K(5) + i];
        else
          dst[K(5) + i] = usynew[i] * u[K(3) + i];
        dst[K(6) + i] = u[K(6) + i];
      }
    if (io == 1)
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 0; k < MN; k++)
        for (int row = 3; row < nyd - MD; row++)
          for (int col = 3; col < nxd - MD; col++) u[K(k) + I(row, col)] = utwo[K(k) + I(row, col)];
  }

  // One pass of the `for (nt...)` body of kernelStep (L/R:646-1330; inflow I/R:721-1424).
  void step() {
    lastWritten = 0;
    if (mode == 1) inflowInject();
    for (int io = 0; io < 2; io++) {
      if (io == 0) uzeroCopy();
      boundary5();
      boundary6();
      boundary78();
      boundary9();
      tvd();
      interfaces();
      interfaces2();
      keepPositivity();
      faceFluxes();
      if (io == 0) {
        cflReduce();
        // host dt control L/R:978-991; inflow sample advance I/R:1078-1081
        double dt_h = dtval;
        const double T_h = TotalTime;
        stepStartTime = T_h;
        if (iter < (int)tf.size() && (T_h + dt_h) >= tf[iter]) {
          dt_h = tf[iter] - T_h;
          schreiben = 1;
          dtval = dt_h;
          iter++;
        }
        if (mode == 1 && T_h >= inflowTime[inflowCount] && inflowCount < (nsamp - 1)) inflowCount++;
      }
      // CFL4Kernel L/K:2022-2036
      dt = dtval / (2.0 - io);
      if (io == 1) TotalTime = dt + TotalTime;
      numFluxes();
      sources();
      update();
      crossing(io);
    }
    if (schreiben == 1) {  // output bookkeeping L/R:1260-1324
      lastWritten = iter;
      if (iter == (int)tf.size()) nstop = 1;
      schreiben = 0;
    }
    nt++;
  }
};

}  // namespace

// =========================================================================================
// C interface (ctypes)
// =========================================================================================
extern "C" {

// Builds the padded, x10-scaled input planes exactly as RunKernels::dataInitialization
// does (L/R:178-301; inflow I/R: topography only): file line j -> padded row j+MD, value i
// -> padded column i+MD, negative values clamped to 0, edge values copied into the ghosts.
// raw[j*NX+i] holds the parsed file value.  out has (NX+6)*(NY+6) entries.
void orc_pad_input(int NX, int NY, const double* raw, double* out) {
  const int nxd = NX + 2 * MD, nyd = NY + 2 * MD;
  std::vector<double> m((size_t)nxd * nyd, 0.0);  // m[i][j] stored as m[i*nyd+j]
  auto at = [&](int i, int j) -> double& { return m[(size_t)i * nyd + j]; };
  if (raw)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        at(i + MD, j + MD) = raw[(size_t)j * NX + i];
        at(i + MD, j + MD) = 10.0 * at(i + MD, j + MD);
        if (at(i + MD, j + MD) < 0) at(i + MD, j + MD) = 0;
      }
  for (int i = 0; i < MD; i++)
    for (int j = MD; j < NY + MD; j++) {
      at(i, j) = at(MD, j);
      at(NX + MD + i, j) = at(NX + MD - 1, j);
    }
  for (int j = 0; j < MD; j++)
    for (int i = 0; i < NX + 2 * MD; i++) {
      at(i, j) = at(i, MD);
      at(i, NY + MD + j) = at(i, NY + MD - 1);
    }
  for (int i = 0; i < nxd; i++)
    for (int j = 0; j < nyd; j++) out[(size_t)j * nxd + i] = at(i, j);
}

// par[0..8] = delta0, Cd, N_R, varTheta, phiS0, cfl, rho_ratio, dt_min, dt_max.
// topo/depth: padded planes from orc_pad_input (depth ignored in inflow mode).
// phi: optional padded per-cell solid fraction (landslide extension), else NULL.
// tf: n_tf output times (scaled), tf[0] = 0; n_tf = 0 disables outputs.
void* orc_new(int mode, int nx, int ny, double dx_scaled, const double* topo, const double* depth,
              const double* phi, const double* par, int n_tf, const double* tf, int nloc,
              int nsamp, const double* inflow, const double* loc, const int* dire,
              const double* inflowTime) {
  Orc* o = new Orc();
  o->mode = mode;
  o->nx = nx;
  o->ny = ny;
  o->nxd = nx + 2 * MD;
  o->nyd = ny + 2 * MD;
  o->dx = dx_scaled;
  o->dy = dx_scaled;
  o->delta0 = par[0];
  o->Cd = par[1];
  o->N_R = par[2];
  o->varTheta = par[3];
  o->phiS0 = par[4];
  o->cfl = par[5];
  o->rr = par[6];
  o->dtlo = par[7];
  o->dthi = par[8];
  o->alloc();
  o->topo.assign(topo, topo + o->P);
  o->depth.assign(o->P, 0.0);
  if (mode == 0 && depth) o->depth.assign(depth, depth + o->P);
  if (phi) o->phicell.assign(phi, phi + o->P);
  if (n_tf > 0) o->tf.assign(tf, tf + n_tf);
  if (mode == 1) {
    o->nloc = nloc;
    o->nsamp = nsamp;
    o->inflow.assign(inflow, inflow + 3 * (size_t)nloc * nsamp);
    o->loc.assign(loc, loc + 3 * (size_t)nloc);
    o->dire.assign(dire, dire + 3 * (size_t)nloc);
    o->inflowTime.assign(inflowTime, inflowTime + nsamp);
  }
  o->setup();
  return o;
}

void orc_free(void* h) { delete (Orc*)h; }

// Runs n steps, or fewer if the run's last output time is reached; returns steps done.
int orc_step(void* h, int n) {
  Orc* o = (Orc*)h;
  int done = 0;
  for (; done < n && !o->nstop; done++) o->step();
  return done;
}

// Runs the first half of a predictor (ghost fill, slopes, interface states, positivity,
// local speeds and face fluxes, L/R:649-961) without updating u, so ux, uy, uE.., Val hold
// that stage's values.  Test hook.
void orc_faces(void* h) {
  Orc* o = (Orc*)h;
  o->uzeroCopy();
  o->boundary5();
  o->boundary6();
  o->boundary78();
  o->boundary9();
  o->tvd();
  o->interfaces();
  o->interfaces2();
  o->keepPositivity();
  o->faceFluxes();
  o->cflReduce();
}

// orc_faces, then the numerical and viscous fluxes (Flux8-10, L/K:2039-2347) and the cell
// sources (Flux11-17, L/K:2349-2621) of that predictor, still without updating u, so
// PDx, PDy, the transformed gradients, Ac, AcF and the basal loads hold that stage's
// values.  Test hook.
void orc_stage_terms(void* h) {
  Orc* o = (Orc*)h;
  orc_faces(h);
  o->numFluxes();
  o->sources();
}

// Applies only the inflow injection of the next step (I/K:818-887).  Test hook.
void orc_inject(void* h) {
  Orc* o = (Orc*)h;
  if (o->mode == 1) o->inflowInject();
}

int orc_nxd(void* h) { return ((Orc*)h)->nxd; }
int orc_nyd(void* h) { return ((Orc*)h)->nyd; }

// Copies a named array (planes of P doubles) into out; returns its plane count or -1.
int orc_get(void* h, const char* name, double* out) {
  Orc* o = (Orc*)h;
  struct Entry {
    const char* n;
    Vec* v;
  };
  Entry tab[] = {{"u", &o->u},
                 {"uone", &o->uone},
                 {"uzero", &o->uzero},
                 {"bfkt", &o->bfkt},
                 {"posx", &o->posx},
                 {"posy", &o->posy},
                 {"dxdxi11", &o->dxdxi11},
                 {"dxdxi12", &o->dxdxi12},
                 {"dxdxi21", &o->dxdxi21},
                 {"dxdxi22", &o->dxdxi22},
                 {"dbdx", &o->dbdx},
                 {"dbdy", &o->dbdy},
                 {"cvalue", &o->cvalue},
                 {"svec", &o->svec},
                 {"Detmin", &o->Detmin},
                 {"invJ11", &o->invJ11},
                 {"invJ12", &o->invJ12},
                 {"invJ13", &o->invJ13},
                 {"invJ21", &o->invJ21},
                 {"invJ22", &o->invJ22},
                 {"invJ23", &o->invJ23},
                 {"invJ31", &o->invJ31},
                 {"invJ32", &o->invJ32},
                 {"invJ33", &o->invJ33},
                 {"J13dxi", &o->J13dxi},
                 {"J23dxi", &o->J23dxi},
                 {"J33dxi", &o->J33dxi},
                 {"J13det", &o->J13det},
                 {"J23det", &o->J23det},
                 {"J33det", &o->J33det},
                 {"tande", &o->tande},
                 {"Jacb31", &o->Jacb31},
                 {"Jacb32", &o->Jacb32},
                 {"ux", &o->ux},
                 {"uy", &o->uy},
                 {"uE", &o->uE},
                 {"uW", &o->uW},
                 {"uN", &o->uN},
                 {"uS", &o->uS},
                 {"Hpx", &o->Hpx},
                 {"Hpy", &o->Hpy},
                 {"Ppx", &o->Ppx},
                 {"Ppy", &o->Ppy},
                 {"PDx", &o->PDx},
                 {"PDy", &o->PDy},
                 {"s", &o->s},
                 {"Val", &o->Val},
                 {"dbdx", &o->dbdx},
                 {"dbdy", &o->dbdy},
                 {"v", &o->v},
                 {"apEW", &o->apEW},
                 {"apSN", &o->apSN},
                 {"invJ11_avgEW", &o->invJ11_avgEW},
                 {"invJ12_avgEW", &o->invJ12_avgEW},
                 {"invJ13_avgEW", &o->invJ13_avgEW},
                 {"invJ21_avgEW", &o->invJ21_avgEW},
                 {"invJ22_avgEW", &o->invJ22_avgEW},
                 {"invJ23_avgEW", &o->invJ23_avgEW},
                 {"Detmin_avgEW", &o->Detmin_avgEW},
                 {"dxdxi11_avgEW", &o->dxdxi11_avgEW},
                 {"dxdxi21_avgEW", &o->dxdxi21_avgEW},
                 {"invJ11_avgSN", &o->invJ11_avgSN},
                 {"invJ12_avgSN", &o->invJ12_avgSN},
                 {"invJ13_avgSN", &o->invJ13_avgSN},
                 {"invJ21_avgSN", &o->invJ21_avgSN},
                 {"invJ22_avgSN", &o->invJ22_avgSN},
                 {"invJ23_avgSN", &o->invJ23_avgSN},
                 {"Detmin_avgSN", &o->Detmin_avgSN},
                 {"dxdxi12_avgSN", &o->dxdxi12_avgSN},
                 {"dxdxi22_avgSN", &o->dxdxi22_avgSN},
                 {"dudxE", &o->dudxE},
                 {"dvdxE", &o->dvdxE},
                 {"dudyE", &o->dudyE},
                 {"dvdyE", &o->dvdyE},
                 {"dudxN", &o->dudxN},
                 {"dvdxN", &o->dvdxN},
                 {"dudyN", &o->dudyN},
                 {"dvdyN", &o->dvdyN},
                 {"Ac", &o->Ac},
                 {"AcF", &o->AcF},
                 {"Npress1", &o->Npress1},
                 {"Npress2", &o->Npress2},
                 {"NpressF", &o->NpressF}};
  for (auto& t : tab)
    if (std::strcmp(t.n, name) == 0) {
      std::memcpy(out, t.v->data(), t.v->size() * sizeof(double));
      return (int)(t.v->size() / o->P);
    }
  return -1;
}

// Overwrites the state u (7 planes of P) -- used to start both implementations from
// the same mid-run state.
void orc_set_u(void* h, const double* in) {
  Orc* o = (Orc*)h;
  std::memcpy(o->u.data(), in, o->u.size() * sizeof(double));
}

// Threads of the OpenMP loops (0: leave the runtime's default).  Results do not depend on
// it: every parallel loop writes only its own cells and the only reduction is a max.
int orc_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

double orc_scalar(void* h, const char* name) {
  Orc* o = (Orc*)h;
  std::string n(name);
  if (n == "time") return o->TotalTime;
  if (n == "dt") return o->dt;
  if (n == "dtval") return o->dtval;
  if (n == "dmax") return o->dmax;
  if (n == "iter") return o->iter;
  if (n == "nt") return o->nt;
  if (n == "nstop") return o->nstop;
  if (n == "inflowCount") return o->inflowCount;
  if (n == "stepStartTime") return o->stepStartTime;
  if (n == "lastWritten") return o->lastWritten;
  return NAN;
}

}  // extern "C"
