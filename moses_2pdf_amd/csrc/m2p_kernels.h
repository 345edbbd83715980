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
  unsigned long long maxbits;  // running max (IEEE bits of a positive double)
  int iter;           // next output index (L/R:390)
  int schreiben;      // output due after this step
  int nstop;
  int inflowCount;
  int nt;             // steps done
  int pad;
};

struct DevParams {
  double cfl, dt_min, dt_max;
  int ntf;     // entries of tf (0: no output clipping)
  int nsamp;   // hydrograph samples (0: landslide)
};

void launch_metric_setup(const Grid& g, int mode, const double* topo, const double* depth,
                         const double* phicell, double phiS0, double* M, double* U,
                         hipStream_t s);
void launch_inflow(const Grid& g, double* U, const double* M, const double* table,
                   const double* loc, const int* dire, int nloc, int nsamp, double phiS0,
                   const DevState* st, hipStream_t s);
void launch_ghost(const Grid& g, int mode, double* T, const double* prev, hipStream_t s);
void launch_cfl(const Grid& g, const Phys& ph, const double* U, const double* M, DevState* st,
                hipStream_t s);
void launch_dt(const DevParams& dp, DevState* st, const double* tf, const double* itime,
               hipStream_t s);
void launch_stage(int io, const Grid& g, const Phys& ph, const double* Uin, const double* Uzero,
                  double* Uout, double* ONE, const double* M, const DevState* st, hipStream_t s);
int mass_blocks();
hipError_t probe_kernels();
void launch_mass(const Grid& g, const double* U, const double* M, double* out, hipStream_t s);

}  // namespace m2p
