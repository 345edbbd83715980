// m2p_ctx.h -- the scenario context behind the C ABI's m2p_ctx handle (internal to the
// library: the engine and the batch driver both reach into it).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/moses2pdf.h"
#include "m2p_kernels.h"

// the stream of the batch that steps a scenario (m2p_batch.hip)
hipStream_t m2p_batch_stream(m2p_batch* b);
// the batch must re-read its members' settings before its next step
void m2p_batch_mark_dirty(m2p_batch* b);

namespace m2p {
// The message of the calling thread's last failed call that takes no context
// (m2p_mass_noalloc), or "".
const char* context_free_error();
}  // namespace m2p

using m2p::DevParams;
using m2p::DevState;

struct m2p_ctx {
  int mode = 0;
  int NX = 0, NY = 0, nxd = 0, nyd = 0;
  long long P = 0;
  double dx = 0, dy = 0, xll = 0, yll = 0;
  m2p_params par{};
  std::vector<double> topo, depth, phi;  // padded, scaled (x10)
  std::vector<double> tf;                // output times (scaled), tf[0] = 0
  std::vector<double> outtime;           // Time.dat entries written so far (m2p_run)
  int nsamp = 0, nloc = 0;
  std::vector<double> table, loc, itime;
  std::vector<int> dire;
  hipStream_t s = nullptr;
  int prec = M2P_FP64;     // scalar type of the state and of the per-step metric copy
  size_t es = 8;           // its size in bytes
  // every device buffer below lies in one of two allocations: dArena (held until the
  // context closes) and dTmp (the set-up temporaries dTopo, dDepth, dPhi, freed after it)
  void *dArena = nullptr, *dTmp = nullptr;
  double *dTopo = nullptr, *dDepth = nullptr, *dPhi = nullptr, *dM = nullptr;
  // state planes u^n (U0), u^{n+1/2} (U1), the limited fluid momenta (ONE) and the metric
  // the step kernels read (MT: dM itself in FP64, a float copy in FP32), in type `prec`
  void *dU0 = nullptr, *dU1 = nullptr, *dONE = nullptr, *dMT = nullptr;
  unsigned char* dLIM = nullptr;  // FP32 only: limiter memory flags, one byte per cell
  unsigned char* dZ = nullptr;    // dry-tile flags of U0 and U1 (m2p::DryMap)
  unsigned char* dGF = nullptr;   // ghost-fill flags of U0 and U1 (m2p::DryMap)
  m2p::DryMap dry{};              // nseg 0 while dry-tile skipping is off
  int* dPlan = nullptr;           // tile plans of the two stages and their 16 counters
  m2p::TilePlan plan[2]{};
  double *dTf = nullptr, *dMass = nullptr;
  double* dPart = nullptr;  // per-block partial maxima of a step (m2p::partial_count)
  double *dTable = nullptr, *dLoc = nullptr, *dItime = nullptr;
  int* dDire = nullptr;
  DevState* dSt = nullptr;
  m2p::Grid g{};
  m2p::Phys ph{};
  std::string err;
  bool prof = false;
  const char* prof_only = nullptr;   // profile only this kernel family (nullptr: every one)
  std::map<std::string, std::pair<double, int>> pstats;
  std::vector<hipEvent_t> evpool;
  size_t evnext = 0;
  std::vector<std::pair<const char*, std::pair<hipEvent_t, hipEvent_t>>> pending;
  m2p_batch* batch = nullptr;  // the batch that steps this scenario, if any
  bool vstop_on = false;       // a stop speed is set (the step reduces corrector speeds)
  // the metric passed the uniform-metric check, so the stage and CFL kernels form the
  // inverse Jacobian from svec and cvalue instead of loading it (FP64: the same values bit
  // for bit; FP32: in float from the rounded planes)
  bool uni_metric = false;
  // a hash of what a checkpoint's state is only valid with: the padded topography,
  // the run parameters and the hydrograph tables (set when the context is opened)
  uint64_t scen_hash = 0;
};
