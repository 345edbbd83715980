// m2p_engine.hip -- host side of the adNOC_2Pb engine: the m2p_ctx context, input
// preparation, the time loop driver and the C ABI declared in include/moses2pdf.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/moses2pdf.h"
#include "m2p_ctx.h"
#include "m2p_io.h"
#include "m2p_kernels.h"

namespace {

constexpr int MD = m2p::MD;

int fail(m2p_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// The stream a scenario's work is ordered on: its own, or its batch's while in a batch.
hipStream_t strm(m2p_ctx* c) { return c->batch ? m2p_batch_stream(c->batch) : c->s; }

// Scenarios take their stream from a small per-device pool, round robin, instead of each
// creating one: creating a HIP stream costs about a millisecond (a quarter of opening a
// 128^2 scenario), and the device serves all streams from a few hardware queues (4 by
// default).  Every operation of a scenario stays on its one stream, so its work keeps its
// order; scenarios sharing a stream also wait for each other's queued work.  The pool's
// streams live until the process ends.
constexpr int STREAM_POOL = 8;

hipError_t pool_stream(hipStream_t* out) {
  static std::mutex mu;
  static std::map<int, std::vector<hipStream_t>> pools;
  static std::map<int, unsigned> next;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  std::vector<hipStream_t>& v = pools[dev];
  if (v.size() < (size_t)STREAM_POOL) {
    hipStream_t s = nullptr;
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return e;
    v.push_back(s);
  }
  *out = v[next[dev]++ % v.size()];
  return hipSuccess;
}

#define IN_BATCH(c)                                                                        \
  if ((c)->batch) return fail((c), M2P_ESTATE, "scenario is stepped by a batch")

#define HIPCHK(ctx, call)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail((ctx), M2P_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

// dataInitialization (L/R:216-301): x10, clamp < 0 to 0, edge copies into the ghosts,
// transposed to [row*nxd+col] (L/R:334-341).
std::vector<double> pad_plane(int NX, int NY, const double* raw, bool scale = true) {
  const int nxd = NX + 2 * MD, nyd = NY + 2 * MD;
  std::vector<double> a((size_t)nxd * nyd, 0.0);
  auto at = [&](int i, int j) -> double& { return a[(size_t)j * nxd + i]; };
  if (raw)
    for (int j = 0; j < NY; j++)
      for (int i = 0; i < NX; i++) {
        double v = raw[(size_t)j * NX + i];
        if (scale) v = 10.0 * v;
        if (v < 0) v = 0;
        at(i + MD, j + MD) = v;
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
  return a;
}

// bfkt planes 0 and 1 (makeTopo1Kernel/makeTopo2Kernel, L/K:5-56), for DEM.dat.
double coord_host(int c, int n, double d) {
  const double xmin = 0.0;
  if (c < MD) return (xmin + 0 * d) - (MD - c) * d;
  if (c >= n + MD) return (xmin + (n - 1) * d) + (c - (n + MD - 1)) * d;
  return xmin + (c - MD) * d;
}

int upload(m2p_ctx* c, double* d, const std::vector<double>& h) {
  if (h.empty()) return 0;
  HIPCHK(c, hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->s));
  return 0;
}

void release(m2p_ctx* c) {
  if (c->s) (void)hipStreamSynchronize(c->s);   // this scenario's queued work is done
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  c->evpool.clear();
  c->pending.clear();
  if (c->dArena) (void)hipFree(c->dArena);
  if (c->dTmp) (void)hipFree(c->dTmp);
  c->dArena = c->dTmp = nullptr;
  c->dTopo = c->dDepth = c->dPhi = c->dM = c->dTf = c->dMass = c->dTable = c->dLoc = c->dItime = nullptr;
  c->dU0 = c->dU1 = c->dONE = c->dMT = nullptr;
  c->dLIM = nullptr;
  c->dPart = nullptr;
  c->dZ = nullptr;
  c->dGF = nullptr;
  c->dPlan = nullptr;
  c->dry = m2p::DryMap{};
  c->dDire = nullptr;
  c->dSt = nullptr;
  c->s = nullptr;   // a pool stream: not destroyed
}

// A landslide's momenta ghost cells in both state buffers set to 0 after a host-side write
// of the state, which the ghost fills rely on (launch_zero_ghost_momenta).
void zero_ghost_momenta(m2p_ctx* c, hipStream_t s) {
  if (c->mode != 0) return;
  if (c->prec == M2P_FP64) {
    m2p::launch_zero_ghost_momenta<double>(c->g, (double*)c->dU0, s);
    m2p::launch_zero_ghost_momenta<double>(c->g, (double*)c->dU1, s);
  } else {
    m2p::launch_zero_ghost_momenta<float>(c->g, (float*)c->dU0, s);
    m2p::launch_zero_ghost_momenta<float>(c->g, (float*)c->dU1, s);
  }
}

// Both dry-flag maps (dm.nseg segments per row) from the state buffers, after they were
// replaced (a new state, a restore, skipping switched on).
void rescan_dry(m2p_ctx* c, const m2p::DryMap& dm, hipStream_t s) {
  if (c->prec == M2P_FP64)
    m2p::launch_dry_scan<double>(c->g, dm, (const double*)c->dU0, (const double*)c->dU1, s);
  else
    m2p::launch_dry_scan<float>(c->g, dm, (const float*)c->dU0, (const float*)c->dU1, s);
}

// Device buffers carved out of one allocation (one hipMalloc and one clear instead of one
// per buffer: they dominated the set-up of small grids).  Each starts on a 256-byte boundary.
struct Arena {
  struct Item {
    void** p;
    size_t off;
  };
  std::vector<Item> items;
  size_t total = 0;
  template <typename T>
  void add(T** p, size_t n) {
    items.push_back({(void**)p, total});
    total += (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~size_t(255);
  }
  // Allocates, clears on the context's stream and sets every pointer; *base owns it.
  int alloc(m2p_ctx* c, void** base) {
    hipError_t e = hipMalloc(base, total);
    if (e != hipSuccess) {
      *base = nullptr;
      return fail(c, M2P_ENOMEM, "hipMalloc of " + std::to_string(total) + " bytes failed");
    }
    e = hipMemsetAsync(*base, 0, total, c->s);
    if (e != hipSuccess) return fail(c, M2P_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    for (const Item& it : items) *it.p = (char*)*base + it.off;
    return 0;
  }
};

// Device allocation, uploads, metric setup (L/R:395-642) and the initial state.
int setup(m2p_ctx* c) {
  HIPCHK(c, pool_stream(&c->s));
  HIPCHK(c, m2p::probe_kernels());
  const size_t P = (size_t)c->P;
  c->g.nx = c->NX;
  c->g.ny = c->NY;
  c->g.nxd = c->nxd;
  c->g.nyd = c->nyd;
  c->g.P = c->P;
  c->g.dx = c->dx;
  c->g.dy = c->dy;
  int rc;
  double* dU64 = nullptr;  // FP64 initial state (the metric setup writes it)
  const int per = m2p::plan_per(m2p::plain_stage_tiles(c->g, c->prec));
  const long long nz = m2p::dry_flag_count(c->g);
  Arena keep, tmp;   // the context's buffers; the set-up temporaries, freed at its end
  tmp.add(&c->dTopo, P);
  tmp.add(&c->dDepth, P);
  if (!c->phi.empty()) tmp.add(&c->dPhi, P);
  // FP32 keeps only the FP64 cell planes; its face planes go straight into the float copy
  keep.add(&c->dM, (size_t)(c->prec == M2P_FP64 ? m2p::MC_COUNT : m2p::MC_CELL_END) * P);
  if (c->prec == M2P_FP64) {
    keep.add(&dU64, 7 * P);
    keep.add((double**)&c->dU1, 7 * P);
    keep.add((double**)&c->dONE, 2 * P);
  } else {
    tmp.add(&dU64, 7 * P);
    keep.add((float**)&c->dU0, 7 * P);
    keep.add((float**)&c->dU1, 7 * P);
    keep.add((float**)&c->dONE, 2 * P);
    keep.add((float**)&c->dMT, (size_t)m2p::MC_COUNT * P);
    keep.add(&c->dLIM, P);
  }
  keep.add(&c->dTf, c->tf.size());
  keep.add(&c->dMass, 4 * (size_t)m2p::mass_blocks());
  keep.add(&c->dSt, 1);
  if (c->mode == 1) {
    keep.add(&c->dTable, c->table.size());
    keep.add(&c->dLoc, c->loc.size());
    keep.add(&c->dItime, c->itime.size());
    keep.add(&c->dDire, c->dire.size());
  }
  keep.add(&c->dPart, m2p::partial_count(c->g, c->prec));
  keep.add(&c->dZ, 2 * (size_t)nz);
  keep.add(&c->dGF, 2 * (size_t)m2p::ghost_flag_count(c->g));
  keep.add(&c->dPlan, m2p::PLAN_COUNTERS + 16 * (size_t)per);
  if ((rc = keep.alloc(c, &c->dArena)) || (rc = tmp.alloc(c, &c->dTmp))) return rc;
  if (c->prec == M2P_FP64) {
    c->dU0 = dU64;
    c->dMT = c->dM;
  }
  if (c->mode == 1) {
    if ((rc = upload(c, c->dTable, c->table)) || (rc = upload(c, c->dLoc, c->loc)) ||
        (rc = upload(c, c->dItime, c->itime)))
      return rc;
    HIPCHK(c, hipMemcpyAsync(c->dDire, c->dire.data(), c->dire.size() * sizeof(int),
                             hipMemcpyHostToDevice, c->s));
  }
  if ((rc = upload(c, c->dTopo, c->topo)) || (rc = upload(c, c->dDepth, c->depth)) ||
      (rc = upload(c, c->dTf, c->tf)))
    return rc;
  if (c->dPhi && (rc = upload(c, c->dPhi, c->phi))) return rc;

  c->ph.rr = c->par.rho_ratio;
  c->ph.Drtio1 = 1.0 - c->par.rho_ratio;
  c->ph.Cd = c->par.Cd;
  c->ph.N_R = c->par.N_R;
  c->ph.varTheta = c->par.varTheta;
  c->ph.tande = std::tan(c->par.delta0 * 3.14159 / 180.);  // makeTopo11Kernel, L/K:298
  c->dry.z0 = c->dZ;
  c->dry.z1 = c->dZ + nz;
  c->dry.g0 = c->dGF;
  c->dry.g1 = c->dGF + m2p::ghost_flag_count(c->g);
  c->dry.nseg = c->par.dry_skip && m2p::dry_capable(c->g) ? m2p::dry_segments(c->g) : 0;
  m2p::launch_dry_fill(c->g, c->dry, c->s);
  for (int io = 0; io < 2; io++) {
    c->plan[io].cnt = c->dPlan + 8 * io;
    c->plan[io].list = c->dPlan + m2p::PLAN_COUNTERS + 8 * (size_t)per * io;
    c->plan[io].per = per;
  }

  DevState st{};
  st.iter = 1;
  st.t_end = -1.0;
  st.halt_at = INT_MAX;
  st.v2_nt = -1;
  st.call_t_end = -1.0;
  st.call_halt = INT_MAX;
  st.nonfinite_nt = -1;
  HIPCHK(c, hipMemcpyAsync(c->dSt, &st, sizeof st, hipMemcpyHostToDevice, c->s));
  m2p::launch_metric_setup(c->g, c->mode, c->dTopo, c->dDepth, c->dPhi, c->par.phiS0, c->dM,
                           dU64, c->s);
  if (c->prec == M2P_FP32) {
    m2p::launch_convert(dU64, (float*)c->dU0, (long long)(7 * P), c->s);
    m2p::launch_convert(c->dM, (float*)c->dMT, (long long)(m2p::MC_CELL_END * P), c->s);
    m2p::launch_face_metric(c->g, c->dM, (float*)c->dMT + m2p::MC_FX * P, c->s);
  } else {
    m2p::launch_face_metric(c->g, c->dM, c->dM + m2p::MC_FX * P, c->s);
  }
  HIPCHK(c, hipGetLastError());
  // may the step form the inverse Jacobian instead of loading it?  (the mass buffer is
  // free during set-up and holds the flag)
  int uni = c->par.metric_form == 0 ? 1 : 0;
  if (uni) {
    int* ok = (int*)c->dMass;
    HIPCHK(c, hipMemcpyAsync(ok, &uni, sizeof uni, hipMemcpyHostToDevice, c->s));
    if (c->prec == M2P_FP64)
      m2p::launch_metric_uniform_check(c->g, c->dM, ok, c->s);
    else
      m2p::launch_metric_uniform_cells_check(c->g, c->dM, ok, c->s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&uni, ok, sizeof uni, hipMemcpyDeviceToHost, c->s));
  }
  // the first predictor's corner ghosts come from the initial state's own ghost rows
  HIPCHK(c, hipMemcpyAsync(c->dU1, c->dU0, 7 * P * c->es, hipMemcpyDeviceToDevice, c->s));
  rescan_dry(c, c->dry, c->s);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->s));
  c->uni_metric = uni != 0;
  // the grids (and the FP32 variant's FP64 initial state) are only needed by the setup
  (void)hipFree(c->dTmp);
  c->dTmp = nullptr;
  c->dTopo = c->dDepth = c->dPhi = nullptr;
  return 0;
}

// State transfers in the reference's FP64 layout, whatever the context's precision.
int download_state(m2p_ctx* c, void* dsrc, size_t planes, double* u) {
  const size_t n = planes * (size_t)c->P;
  hipStream_t s = strm(c);
  if (c->prec == M2P_FP64) {
    HIPCHK(c, hipMemcpyAsync(u, dsrc, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
  }
  std::vector<float> f(n);
  HIPCHK(c, hipMemcpyAsync(f.data(), dsrc, n * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  for (size_t i = 0; i < n; i++) u[i] = (double)f[i];
  return 0;
}

int upload_state(m2p_ctx* c, void* ddst, const double* u) {
  const size_t n = 7 * (size_t)c->P;
  if (c->prec == M2P_FP64) {
    HIPCHK(c, hipMemcpyAsync(ddst, u, n * sizeof(double), hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return 0;
  }
  std::vector<float> f(n);
  for (size_t i = 0; i < n; i++) f[i] = (float)u[i];
  HIPCHK(c, hipMemcpyAsync(ddst, f.data(), n * sizeof(float), hipMemcpyHostToDevice, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

// Per-launch HIP events on the context's stream, resolved only when read (no host
// synchronisation between launches).
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};

hipEvent_t pool_event(m2p_ctx* c);

struct Timer {
  m2p_ctx* c;
  const char* name;
  hipEvent_t a = nullptr;
  Timer(m2p_ctx* c_, const char* n) : c(c_), name(n) {
    if (c->prof && (!c->prof_only || std::strcmp(c->prof_only, n) == 0)) {
      a = pool_event(c);
      (void)hipEventRecord(a, c->s);
    }
  }
  ~Timer();
};

hipEvent_t pool_event(m2p_ctx* c) {
  if (c->evnext == c->evpool.size()) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    c->evpool.push_back(e);
  }
  return c->evpool[c->evnext++];
}

Timer::~Timer() {
  if (a) {
    hipEvent_t b = pool_event(c);
    (void)hipEventRecord(b, c->s);
    c->pending.push_back({name, {a, b}});
  }
}

// Folds recorded event pairs into pstats (waits for the last one).
void resolve_profile(m2p_ctx* c) {
  if (c->pending.empty()) return;
  (void)hipEventSynchronize(c->pending.back().second.second);
  for (auto& p : c->pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
    auto& e = c->pstats[p.first];
    e.first += ms;
    e.second += 1;
  }
  c->pending.clear();
  c->evnext = 0;
}

// One pass of the kernelStep loop body (L/R:646-1330; inflow I/R:721-1424).
template <typename T>
void enqueue_step_t(m2p_ctx* c, int ntf) {
  T* U0 = (T*)c->dU0;
  T* U1 = (T*)c->dU1;
  T* ONE = (T*)c->dONE;
  const T* MT = (const T*)c->dMT;
  DevParams dp{};
  dp.cfl = c->par.cfl;
  dp.dt_min = c->par.dt_min;
  dp.dt_max = c->par.dt_max;
  dp.dx = c->dx;
  dp.dy = c->dy;
  dp.ntf = ntf;
  dp.nsamp = c->mode == 1 ? c->nsamp : 0;
  if (c->mode == 1) {
    Timer t(c, "inflow");
    m2p::launch_inflow(c->g, U0, c->dM, c->dTable, c->dLoc, c->dDire, c->nloc, c->nsamp,
                       c->par.phiS0, c->dSt, c->dLIM, c->dry, c->s);
  }
  {
    Timer t(c, "ghost");
    m2p::launch_ghost(c->g, c->mode, U0, (const T*)U1, c->dSt, 0, c->dry.z0, c->dry.g0,
                      c->dry.nseg, c->s);
  }
  {
    Timer t(c, "cfl");
    m2p::launch_cfl(c->g, c->ph, (const T*)U0, MT, c->dSt, c->dPart, c->dry, c->uni_metric, c->s);
  }
  {
    Timer t(c, "dt");
    m2p::launch_dt(dp, c->dSt, c->dTf, c->dItime, c->g, c->prec, c->dPart,
                   c->dry.nseg > 0 ? c->dPlan : nullptr, c->s);
  }
  {
    Timer t(c, "stage0");
    if (c->dry.nseg > 0)
      m2p::launch_plan(0, c->g, (const T*)U0, c->dry, c->plan[0], c->dPart, c->dSt, c->s);
    m2p::launch_stage(0, c->g, c->ph, (const T*)U0, (const T*)U0, U1, ONE, MT, c->dLIM, c->dSt,
                      c->dPart, c->dry, c->dry.nseg > 0 ? c->plan[0] : m2p::TilePlan{},
                      c->uni_metric, c->s);
  }
  {
    Timer t(c, "ghost");
    m2p::launch_ghost(c->g, c->mode, U1, (const T*)U0, c->dSt, 1, c->dry.z1, c->dry.g1,
                      c->dry.nseg, c->s);
  }
  {
    Timer t(c, "stage1");
    if (c->dry.nseg > 0)
      m2p::launch_plan(1, c->g, (const T*)U1, c->dry, c->plan[1], c->dPart, c->dSt, c->s);
    m2p::launch_stage(1, c->g, c->ph, (const T*)U1, (const T*)U0, U0, ONE, MT, c->dLIM, c->dSt,
                      c->dPart, c->dry, c->dry.nseg > 0 ? c->plan[1] : m2p::TilePlan{},
                      c->uni_metric, c->s);
  }
  if (c->vstop_on) m2p::launch_vmax(c->dSt, c->g, c->prec, c->dry.nseg > 0, c->dPart, c->s);
}

void enqueue_step(m2p_ctx* c, int ntf) {
  if (c->prec == M2P_FP32)
    enqueue_step_t<float>(c, ntf);
  else
    enqueue_step_t<double>(c, ntf);
}

int parse_overlay(const char* ov, m2p_params& p) {
  if (!ov) return 0;
  std::string s(ov);
  size_t pos = 0;
  while (pos < s.size()) {
    size_t end = s.find_first_of(";,", pos);
    if (end == std::string::npos) end = s.size();
    std::string kv = s.substr(pos, end - pos);
    pos = end + 1;
    if (kv.empty()) continue;
    size_t eq = kv.find('=');
    if (eq == std::string::npos) return -1;
    std::string k = kv.substr(0, eq), val = kv.substr(eq + 1);
    if (k == "dry_skip" || k == "metric_form") {
      if (val != "0" && val != "1") return -1;
      (k == "dry_skip" ? p.dry_skip : p.metric_form) = val[0] - '0';
      continue;
    }
    if (k == "precision") {
      if (val == "fp64") p.precision = M2P_FP64;
      else if (val == "fp32") p.precision = M2P_FP32;
      else return -1;
      continue;
    }
    char* endp = nullptr;
    const double v = std::strtod(val.c_str(), &endp);
    if (val.empty() || endp != val.c_str() + val.size() || !std::isfinite(v)) return -1;
    if (k == "cfl") p.cfl = v;
    else if (k == "rho_ratio") p.rho_ratio = v;
    else if (k == "dt_min") p.dt_min = v;
    else if (k == "dt_max") p.dt_max = v;
    else return -1;
  }
  return 0;
}

// The reason a parameter set cannot run, or "" (the reference reads par_List with sscanf
// and checks nothing, adNOC_2Pb_runKernels.cu:11-48).
std::string check_params(int mode, const m2p_params& p) {
  auto fin = [](double x) { return std::isfinite(x); };
  if (!fin(p.total_sim) || p.total_sim < 0) return "total_sim must be finite and >= 0";
  if (!fin(p.each_out) || !(p.each_out > 0)) return "each_out must be finite and > 0";
  if (!fin(p.delta0) || p.delta0 < 0 || !(p.delta0 < 90)) return "delta0 must be in [0, 90) degrees";
  if (!fin(p.Cd) || p.Cd < 0) return "Cd must be finite and >= 0";
  if (!fin(p.N_R) || !(p.N_R > 0)) return "N_R must be finite and > 0";
  if (!fin(p.varTheta) || p.varTheta < 0) return "varTheta must be finite and >= 0";
  if (!fin(p.phiS0) || p.phiS0 < 0 || p.phiS0 > 1) return "phiS0 must be in [0, 1]";
  if (mode == 1 && !(p.phiS0 > 0)) return "phiS0 must be > 0 in inflow mode (the injection divides by it)";
  if (!fin(p.cfl) || !(p.cfl > 0) || p.cfl > 1) return "cfl must be in (0, 1]";
  if (!fin(p.rho_ratio) || p.rho_ratio < 0 || !(p.rho_ratio < 1)) return "rho_ratio must be in [0, 1)";
  if (!fin(p.dt_min) || !(p.dt_min > 0)) return "dt_min must be finite and > 0";
  if (!fin(p.dt_max) || p.dt_max < p.dt_min) return "dt_max must be finite and >= dt_min";
  if (p.precision != M2P_FP64 && p.precision != M2P_FP32) return "precision must be M2P_FP64 or M2P_FP32";
  if (p.dry_skip != 0 && p.dry_skip != 1) return "dry_skip must be 0 or 1";
  if (p.metric_form != 0 && p.metric_form != 1) return "metric_form must be 0 or 1";
  return "";
}

// A 64-bit hash of the scenario's topography, parameters and hydrograph tables: FNV-1a's
// step (xor, then multiply by the odd FNV prime, a bijection) on 64-bit words in four
// interleaved lanes, so that a 1024^2 grid hashes in well under a millisecond (a serial
// byte-wise chain took about 10 ms); a difference in any one word always changes the
// result.  It binds a checkpoint to the scenario that wrote it.
uint64_t scenario_hash(const m2p_ctx* c) {
  constexpr uint64_t kBasis = 1469598103934665603ull, kPrime = 1099511628211ull;
  uint64_t h[4] = {kBasis, kBasis ^ 1, kBasis ^ 2, kBasis ^ 3};
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    size_t i = 0;
    for (; i + 32 <= n; i += 32)
      for (int k = 0; k < 4; k++) {
        uint64_t w;
        std::memcpy(&w, b + i + 8 * k, 8);
        h[k] = (h[k] ^ w) * kPrime;
      }
    for (; i < n; i++) h[0] = (h[0] ^ b[i]) * kPrime;
    for (int k = 0; k < 4; k++) h[k] = (h[k] ^ n) * kPrime;   // the length ends each part
  };
  const double pv[] = {c->par.total_sim, c->par.each_out, c->par.delta0, c->par.Cd, c->par.N_R,
                       c->par.varTheta,  c->par.phiS0,    c->par.cfl,    c->par.rho_ratio,
                       c->par.dt_min,    c->par.dt_max,   c->dx};
  mix(pv, sizeof pv);
  mix(c->topo.data(), c->topo.size() * sizeof(double));
  mix(c->table.data(), c->table.size() * sizeof(double));
  mix(c->loc.data(), c->loc.size() * sizeof(double));
  mix(c->itime.data(), c->itime.size() * sizeof(double));
  uint64_t r = kBasis;
  for (int k = 0; k < 4; k++) r = (r ^ h[k]) * kPrime;
  return r ^ (r >> 29);
}

// A context that holds only an error message, for the caller to read and close.
int reject(m2p_ctx** out, int code, const std::string& msg) {
  m2p_ctx* c = new m2p_ctx();
  c->err = msg;
  *out = c;
  return code;
}

}  // namespace

extern "C" {

void m2p_default_params(int mode, m2p_params* p) {
  // par_List of the shipped examples (MoSES_2PDF_landslide/par_List:1-7) and the kernel
  // constants (adNOC_2Pb_kernels.cuh:16; adNOC_2Pb_kernels.cu:1710, 2014-2017).
  p->total_sim = mode == 0 ? 180 : 720;
  p->each_out = mode == 0 ? 3 : 10;
  p->delta0 = 16.0;
  p->Cd = 6.0;
  p->N_R = 268.;
  p->varTheta = 5.0;
  p->phiS0 = 0.5;
  p->cfl = 0.1;
  p->rho_ratio = 1.420 / 2.60;
  p->dt_min = mode == 0 ? 0.001 : 0.002;
  p->dt_max = 2.0;
  p->precision = M2P_FP64;
  p->dry_skip = 1;
  p->metric_form = 0;
}

int m2p_open_grid(int mode, int NX, int NY, double cellsize, const double* topo,
                  const double* depth, const double* phi, const m2p_params* par, int nsamp,
                  int nloc, const double* t, const double* h, const double* u, const double* v,
                  const double* loc, m2p_ctx** out) {
  if (!out) return M2P_EINVAL;
  *out = nullptr;
  if ((mode != 0 && mode != 1) || NX < 2 || NY < 2 || !topo || !par || !(cellsize > 0))
    return M2P_EINVAL;
  if (mode == 0 && !depth) return M2P_EINVAL;
  if (mode == 1 && (nsamp < 1 || nloc < 1 || !t || !h || !u || !v || !loc)) return M2P_EINVAL;
  {
    const std::string why = check_params(mode, *par);
    if (!why.empty()) return reject(out, M2P_EINVAL, "invalid parameters: " + why);
  }
  m2p_ctx* c = new m2p_ctx();
  c->mode = mode;
  c->prec = par->precision;
  c->es = c->prec == M2P_FP64 ? sizeof(double) : sizeof(float);
  c->NX = NX;
  c->NY = NY;
  c->nxd = NX + 2 * MD;
  c->nyd = NY + 2 * MD;
  c->P = (long long)c->nxd * c->nyd;
  c->dx = cellsize * 10;  // L/R:181-182; the cells are square (the CFL max relies on it)
  c->dy = c->dx;
  c->par = *par;
  c->topo = pad_plane(NX, NY, topo);
  c->depth = (mode == 0) ? pad_plane(NX, NY, depth) : std::vector<double>((size_t)c->P, 0.0);
  if (phi && mode == 0) {
    // per-cell solid fraction: padded like the depth, but not scaled
    c->phi = pad_plane(NX, NY, phi, false);
  }
  // output times tf[] (L/R:382-387)
  const int outputStep = (par->each_out > 0) ? (int)(par->total_sim / par->each_out) : 0;
  c->tf.assign(outputStep + 1, 0.0);
  for (int ii = 1; ii <= outputStep; ii++) c->tf[ii] = ii * par->each_out * 10;
  if (mode == 1) {
    // hydrograph tables (I/R:297-331, 394-428)
    c->nsamp = nsamp;
    c->nloc = nloc;
    c->itime.resize(nsamp);
    c->table.assign((size_t)3 * nsamp * nloc, 0.0);
    for (int j = 0; j < nsamp; j++) {
      c->itime[j] = (double)((float)t[j] * 10);
      for (int i = 0; i < nloc; i++) {
        const size_t q = (size_t)j * nloc + i;
        c->table[0 * (size_t)nsamp * nloc + q] = 10.0 * h[q];
        c->table[1 * (size_t)nsamp * nloc + q] = u[q];
        c->table[2 * (size_t)nsamp * nloc + q] = v[q];
      }
    }
    c->loc.assign(loc, loc + (size_t)3 * nloc);
    c->dire.assign((size_t)3 * nloc, 0);
    for (int j = 0; j < nloc; j++) {
      const double Y = loc[j * 3 + 2];
      c->dire[j * 3 + 0] = (int)loc[j * 3 + 0];
      if (Y == 0 || Y == NY) {
        c->dire[j * 3 + 1] = -5;
        c->dire[j * 3 + 2] = 0;
      } else {
        c->dire[j * 3 + 1] = 0;
        c->dire[j * 3 + 2] = -5;
      }
      // the injected strip must lie inside the padded grid
      const int X0 = (int)loc[j * 3 + 1], Y0 = (int)Y;
      const int w = c->dire[j * 3 + 1] + MD + 3, hgt = c->dire[j * 3 + 2] + MD + 3;
      if (X0 < 0 || Y0 < 0 || X0 + w > c->nxd || Y0 + hgt > c->nyd) {
        delete c;
        return M2P_EINVAL;
      }
    }
  }
  c->scen_hash = scenario_hash(c);
  int rc = setup(c);
  if (rc) {
    // keep the context so the caller can read the message
    release(c);
    *out = c;
    return rc;
  }
  *out = c;
  return 0;
}

int m2p_open(const char* par_list, int mode, const char* overlay, m2p_ctx** out) {
  if (!out || !par_list || (mode != 0 && mode != 1)) return M2P_EINVAL;
  *out = nullptr;
  static thread_local m2p_ctx errctx;
  m2p::ParList pl;
  std::string e = m2p::read_par_list(par_list, mode, pl);
  auto bad = [&](int code, const std::string& msg) {
    m2p_ctx* c = new m2p_ctx();
    c->err = msg;
    *out = c;
    return code;
  };
  if (!e.empty()) return bad(M2P_EIO, e);
  m2p::AscGrid dem;
  e = m2p::read_asc(pl.dem, true, dem);
  if (!e.empty()) return bad(M2P_EIO, e);
  m2p_params p;
  m2p_default_params(mode, &p);
  p.total_sim = pl.TotalSim;
  p.each_out = pl.EachOut;
  p.delta0 = pl.delta0;
  p.Cd = pl.Cd;
  p.N_R = pl.N_R;
  p.varTheta = pl.varTheta;
  p.phiS0 = pl.phiS0;
  if (parse_overlay(overlay, p))
    return bad(M2P_EINVAL, std::string("bad overlay string: ") + overlay);
  int rc;
  m2p_ctx* c = nullptr;
  if (mode == 0) {
    std::vector<double> depth;
    e = m2p::read_asc_values(pl.init, dem.NX, dem.NY, false, depth);
    if (!e.empty()) return bad(M2P_EIO, e);
    rc = m2p_open_grid(0, dem.NX, dem.NY, dem.cellsize, dem.v.data(), depth.data(), nullptr, &p, 0,
                       0, nullptr, nullptr, nullptr, nullptr, nullptr, &c);
  } else {
    m2p::Hydro hy;
    e = m2p::read_hydro(pl, hy);
    if (!e.empty()) return bad(M2P_EIO, e);
    rc = m2p_open_grid(1, dem.NX, dem.NY, dem.cellsize, dem.v.data(), nullptr, nullptr, &p, hy.nsamp,
                       hy.nloc, hy.t.data(), hy.h.data(), hy.u.data(), hy.v.data(), hy.loc.data(), &c);
  }
  (void)errctx;
  if (!c) return bad(rc ? rc : M2P_EINVAL, "invalid grid or inflow table");
  if (rc && !c->s) {   // an error-only context from m2p_open_grid: hand its message on
    *out = c;
    return rc;
  }
  c->xll = dem.xll;
  c->yll = dem.yll;
  *out = c;
  return rc;
}

int m2p_step(m2p_ctx* c, int nsteps) {
  if (!c || !c->s) return M2P_ESTATE;
  IN_BATCH(c);
  if (nsteps < 0) return fail(c, M2P_EINVAL, "negative step count");
  for (int n = 0; n < nsteps; n++) {
    enqueue_step(c, 0);
    if (c->prof && c->pending.size() > 4096) resolve_profile(c);
  }
  HIPCHK(c, hipGetLastError());
  return 0;
}

int m2p_run_to(m2p_ctx* c, double t_end, int max_steps, int* steps_done) {
  if (!c || !c->s) return M2P_ESTATE;
  IN_BATCH(c);
  if (!(t_end > 0) || !std::isfinite(t_end) || max_steps < 0)
    return fail(c, M2P_EINVAL, "t_end must be finite and > 0, max_steps >= 0");
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  const int nt0 = st.nt;
  if (st.time < t_end && max_steps > 0) {
    // the call's own end time and budget; the halting rules of m2p_set_stop stay in force
    // beside them and are left as they are
    st.call_t_end = t_end;
    st.call_halt = (max_steps > INT_MAX - nt0) ? INT_MAX : nt0 + max_steps;
    HIPCHK(c, hipMemcpyAsync(c->dSt, &st, sizeof st, hipMemcpyHostToDevice, c->s));
    // chunks of steps; the kernels of a halted scenario return at once, so a chunk may run
    // past the end
    int issued = 0;
    while (issued < max_steps) {
      const int n = std::min(64, max_steps - issued);
      for (int k = 0; k < n; k++) enqueue_step(c, 0);
      issued += n;
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, c->s));
      HIPCHK(c, hipStreamSynchronize(c->s));
      if (st.skip || st.nt >= st.halt_at || st.nt >= st.call_halt) break;
    }
    st.call_t_end = -1.0;
    st.call_halt = INT_MAX;
    // halted by the call's rules only: running again is allowed
    if (st.skip && st.nt < st.halt_at) st.skip = 0;
    HIPCHK(c, hipMemcpyAsync(c->dSt, &st, sizeof st, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
  }
  if (steps_done) *steps_done = st.nt - nt0;
  return 0;
}

int m2p_sync(m2p_ctx* c) {
  if (!c || !c->s) return M2P_ESTATE;
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

}  // extern "C"

namespace {

// The time loop of RunKernels::run (L/R:306-356, 646-1330) with its output, from the
// context's current state, for at most max_steps steps.  *finished: the last output time
// was reached.
int run_loop(m2p_ctx* c, const char* result_dir, int max_steps, double* wall_s, int* finished) {
  if (!c || !c->s || !result_dir) return M2P_ESTATE;
  IN_BATCH(c);
  const auto t0 = std::chrono::steady_clock::now();
  const std::string dir(result_dir);
  const size_t P = (size_t)c->P;
  m2p::OutputGrid og{c->NX, c->NY, c->nxd, c->nyd};
  {
    std::vector<double> bx(P), by(P), cv(P), s0(P), s1(P);
    for (int r = 0; r < c->nyd; r++)
      for (int q = 0; q < c->nxd; q++) {
        bx[(size_t)r * c->nxd + q] = coord_host(q, c->NX, c->dx);
        by[(size_t)r * c->nxd + q] = coord_host(r, c->NY, c->dy);
      }
    HIPCHK(c, hipMemcpy(cv.data(), c->dM + m2p::MC_C * P, P * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(s0.data(), c->dM + m2p::MC_S0 * P, P * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(s1.data(), c->dM + m2p::MC_S1 * P, P * sizeof(double), hipMemcpyDeviceToHost));
    std::string e = m2p::write_dem(dir, og, bx, by, c->topo, cv, s0, s1);
    if (e.empty()) e = m2p::write_init(dir, og, c->depth, c->par.phiS0);
    if (e.empty()) e = m2p::write_info(dir, c->NX, c->NY, c->dx, c->dy, c->xll, c->yll, (int)c->tf.size());
    if (!e.empty()) return fail(c, M2P_EIO, e);
  }
  const int ntf = (int)c->tf.size();
  std::vector<double>& outtime = c->outtime;   // kept in the context: a checkpoint carries it
  outtime.resize(ntf + 1, 0.0);
  std::vector<double> u(7 * P);
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  int nt;
  bool stop = (ntf <= 1) || st.iter >= ntf;   // a resumed run may already be complete
  const int nt_end = st.nt + max_steps;
  for (nt = st.nt + 1; !stop && nt < 100000 && nt <= nt_end; nt++) {
    enqueue_step(c, ntf);
    HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    if (st.schreiben) {
      if (int rc = download_state(c, c->dU0, 7, u.data())) return rc;
      outtime[st.iter - 1] = st.step_start;
      std::string e = m2p::write_time(dir, outtime, st.iter);
      if (e.empty()) e = m2p::write_result(dir, st.iter, og, u.data());
      if (!e.empty()) return fail(c, M2P_EIO, e);
      if (st.iter == ntf) stop = true;
      const int zero = 0;
      HIPCHK(c, hipMemcpyAsync((char*)c->dSt + offsetof(DevState, schreiben), &zero, sizeof zero,
                               hipMemcpyHostToDevice, c->s));
    }
    std::fprintf(stdout, "\r\t\tSimulation progress: %3.0f %%",
                 (st.step_start * 0.1 / c->par.total_sim) * 100);
    std::fflush(stdout);
  }
  HIPCHK(c, hipStreamSynchronize(c->s));
  if (finished) *finished = stop ? 1 : 0;
  if (stop || nt >= 100000) {
    std::printf("\nTotal time : %g sec  ", st.time / 10);
    std::printf("\nTotal number of steps: %d\n", nt);
  }
  if (wall_s)
    *wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

}  // namespace

extern "C" {

int m2p_run(m2p_ctx* c, const char* result_dir, double* wall_s) {
  return run_loop(c, result_dir, 100000, wall_s, nullptr);
}

int m2p_run_steps(m2p_ctx* c, const char* result_dir, int max_steps, int* finished, double* wall_s) {
  if (c && max_steps < 0) return fail(c, M2P_EINVAL, "negative step count");
  return run_loop(c, result_dir, max_steps, wall_s, finished);
}



int m2p_get_state(m2p_ctx* c, double* u, double* t, int* nt) {
  if (!c || !c->s) return M2P_ESTATE;
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, strm(c)));
  HIPCHK(c, hipStreamSynchronize(strm(c)));
  if (u)
    if (int rc = download_state(c, c->dU0, 7, u)) return rc;
  if (t) *t = st.time;
  if (nt) *nt = st.nt;
  return 0;
}

int m2p_set_state(m2p_ctx* c, const double* u) {
  if (!c || !c->s) return M2P_ESTATE;
  IN_BATCH(c);
  if (!u) return fail(c, M2P_EINVAL, "null state");
  if (int rc = upload_state(c, c->dU0, u)) return rc;
  {
    // a new state: no non-finite value has been met in it yet
    const int nf[2] = {0, -1};
    HIPCHK(c, hipMemcpyAsync((char*)c->dSt + offsetof(DevState, nonfinite), nf, sizeof nf,
                             hipMemcpyHostToDevice, c->s));
  }
  HIPCHK(c, hipMemcpyAsync(c->dU1, c->dU0, 7 * (size_t)c->P * c->es, hipMemcpyDeviceToDevice, c->s));
  if (c->dLIM) HIPCHK(c, hipMemsetAsync(c->dLIM, 0, (size_t)c->P, c->s));
  zero_ghost_momenta(c, c->s);
  rescan_dry(c, c->dry, c->s);   // the flags must describe the new state
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->s));
  return 0;
}

int m2p_dims(const m2p_ctx* c, int* nx, int* ny, int* nxd, int* nyd) {
  if (!c) return M2P_ESTATE;
  if (nx) *nx = c->NX;
  if (ny) *ny = c->NY;
  if (nxd) *nxd = c->nxd;
  if (nyd) *nyd = c->nyd;
  return 0;
}

int m2p_mass(m2p_ctx* c, double* solid, double* fluid) {
  if (!c || !c->s) return M2P_ESTATE;
  const int nb = m2p::mass_blocks();
  std::vector<double> part(4 * (size_t)nb);
  hipStream_t s = strm(c);
  if (c->prec == M2P_FP32)
    m2p::launch_mass(c->g, (const float*)c->dU0, c->dM + m2p::MC_DET * c->P, c->dMass, s);
  else
    m2p::launch_mass(c->g, (const double*)c->dU0, c->dM + m2p::MC_DET * c->P, c->dMass, s);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(part.data(), c->dMass, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  // double-double sum of the block partials, in block order
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
  if (solid) *solid = a + al;
  if (fluid) *fluid = b + bl;
  return 0;
}

int m2p_scalars(m2p_ctx* c, double* dt, double* dmax, int* iter) {
  if (!c || !c->s) return M2P_ESTATE;
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, strm(c)));
  HIPCHK(c, hipStreamSynchronize(strm(c)));
  if (dt) *dt = st.dt_corr;
  if (dmax) *dmax = st.dmax;
  if (iter) *iter = st.iter;
  return 0;
}

int m2p_get_field(m2p_ctx* c, const char* name, double* out) {
  if (!c || !c->s) return M2P_ESTATE;
  if (!name || !out) return fail(c, M2P_EINVAL, "null argument");
  static const std::map<std::string, std::pair<int, int>> tab = {
      {"cvalue", {m2p::MC_C, 1}},    {"svec", {m2p::MC_S0, 2}},     {"Detmin", {m2p::MC_DET, 1}},
      {"invJ11", {m2p::MC_I11, 1}},  {"invJ12", {m2p::MC_I12, 1}},  {"invJ13", {m2p::MC_I13, 1}},
      {"invJ21", {m2p::MC_I21, 1}},  {"invJ22", {m2p::MC_I22, 1}},  {"invJ23", {m2p::MC_I23, 1}},
      {"dxdxi11", {m2p::MC_X11, 1}}, {"dxdxi12", {m2p::MC_X12, 1}}, {"dxdxi21", {m2p::MC_X21, 1}},
      {"dxdxi22", {m2p::MC_X22, 1}}, {"J13dxi", {m2p::MC_J13X, 1}}, {"J23dxi", {m2p::MC_J23X, 1}},
      {"J33dxi", {m2p::MC_J33X, 1}}, {"J13det", {m2p::MC_J13E, 1}}, {"J23det", {m2p::MC_J23E, 1}},
      {"J33det", {m2p::MC_J33E, 1}},
      {"invJ11_avgEW", {m2p::MC_FX + m2p::FQ_I11, 1}}, {"invJ12_avgEW", {m2p::MC_FX + m2p::FQ_I12, 1}},
      {"invJ13_avgEW", {m2p::MC_FX + m2p::FQ_I13, 1}}, {"invJ21_avgEW", {m2p::MC_FX + m2p::FQ_I21, 1}},
      {"invJ22_avgEW", {m2p::MC_FX + m2p::FQ_I22, 1}}, {"invJ23_avgEW", {m2p::MC_FX + m2p::FQ_I23, 1}},
      {"Detmin_avgEW", {m2p::MC_FX + m2p::FQ_DET, 1}}, {"dxdxi11_avgEW", {m2p::MC_FX + m2p::FQ_XA, 1}},
      {"dxdxi21_avgEW", {m2p::MC_FX + m2p::FQ_XB, 1}},
      {"invJ11_avgSN", {m2p::MC_FY + m2p::FQ_I11, 1}}, {"invJ12_avgSN", {m2p::MC_FY + m2p::FQ_I12, 1}},
      {"invJ13_avgSN", {m2p::MC_FY + m2p::FQ_I13, 1}}, {"invJ21_avgSN", {m2p::MC_FY + m2p::FQ_I21, 1}},
      {"invJ22_avgSN", {m2p::MC_FY + m2p::FQ_I22, 1}}, {"invJ23_avgSN", {m2p::MC_FY + m2p::FQ_I23, 1}},
      {"Detmin_avgSN", {m2p::MC_FY + m2p::FQ_DET, 1}}, {"dxdxi12_avgSN", {m2p::MC_FY + m2p::FQ_XA, 1}},
      {"dxdxi22_avgSN", {m2p::MC_FY + m2p::FQ_XB, 1}}};
  const size_t P = (size_t)c->P;
  std::string n(name);
  if (n == "uone45") {
    if (int rc = download_state(c, c->dONE, 2, out)) return rc;
    return 2;
  }
  if (n == "metric_uniform") {
    out[0] = c->uni_metric ? 1.0 : 0.0;
    return 1;
  }
  auto it = tab.find(n);
  if (it == tab.end()) return fail(c, M2P_EINVAL, "unknown field " + n);
  if (c->prec == M2P_FP32 && it->second.first >= m2p::MC_CELL_END)
    return fail(c, M2P_EINVAL, "the FP32 variant keeps face averages only as floats: " + n);
  HIPCHK(c, hipStreamSynchronize(strm(c)));
  HIPCHK(c, hipMemcpy(out, c->dM + (size_t)it->second.first * P,
                      (size_t)it->second.second * P * sizeof(double), hipMemcpyDeviceToHost));
  return it->second.second;
}

int m2p_profile(m2p_ctx* c, int enable) {
  if (!c) return M2P_ESTATE;
  static const char* const fam[] = {"inflow", "ghost", "cfl", "dt", "stage0", "stage1", "vmax"};
  if (enable < 0 || enable > 2 + 6) return fail(c, M2P_EINVAL, "profile: enable must be 0, 1 or 2..8");
  resolve_profile(c);
  c->prof = enable != 0;
  c->prof_only = enable >= 2 ? fam[enable - 2] : nullptr;
  c->pstats.clear();
  return 0;
}

int m2p_profile_get(m2p_ctx* c, const char* kernel, double* ms, int* launches) {
  if (!c || !kernel) return M2P_ESTATE;
  resolve_profile(c);
  auto it = c->pstats.find(kernel);
  if (it == c->pstats.end()) {
    if (ms) *ms = 0;
    if (launches) *launches = 0;
    return 0;
  }
  if (ms) *ms = it->second.second ? it->second.first / it->second.second : 0.0;
  if (launches) *launches = it->second.second;
  return 0;
}

int m2p_set_stop(m2p_ctx* c, double t_end, double vstop, int max_steps) {
  if (!c || !c->s) return M2P_ESTATE;
  hipStream_t s = strm(c);
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  st.t_end = t_end > 0 ? t_end : -1.0;
  st.vstop2 = vstop > 0 ? vstop * vstop : 0.0;
  st.halt_at = max_steps < 0 ? INT_MAX : (max_steps > INT_MAX - st.nt ? INT_MAX : st.nt + max_steps);
  st.skip = 0;
  st.moved = 0;
  st.v2_nt = -1;
  HIPCHK(c, hipMemcpyAsync(c->dSt, &st, sizeof st, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipStreamSynchronize(s));
  c->vstop_on = vstop > 0;
  if (c->batch) m2p_batch_mark_dirty(c->batch);
  return 0;
}

int m2p_set_dry_skip(m2p_ctx* c, int on) {
  if (!c || !c->s) return M2P_ESTATE;
  if (on != 0 && on != 1) return fail(c, M2P_EINVAL, "dry_skip must be 0 or 1");
  hipStream_t s = strm(c);
  c->par.dry_skip = on;
  c->dry.nseg = on && m2p::dry_capable(c->g) ? m2p::dry_segments(c->g) : 0;
  // while off, no kernel keeps the flags, so they are set again from the state
  m2p::DryMap all = c->dry;
  all.nseg = m2p::dry_segments(c->g);
  rescan_dry(c, all, s);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(s));
  if (c->batch) m2p_batch_mark_dirty(c->batch);
  return 0;
}

int m2p_status(m2p_ctx* c, double* t, int* nt, int* halted) {
  if (!c || !c->s) return M2P_ESTATE;
  hipStream_t s = strm(c);
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (t) *t = st.time;
  if (nt) *nt = st.nt;
  if (halted) *halted = (st.skip || st.nt >= st.halt_at) ? 1 : 0;
  return 0;
}

int m2p_nonfinite(m2p_ctx* c, int* flag, int* first_nt) {
  if (!c || !c->s) return M2P_ESTATE;
  hipStream_t s = strm(c);
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (flag) *flag = st.nonfinite;
  if (first_nt) *first_nt = st.nonfinite ? st.nonfinite_nt : -1;
  return 0;
}

int m2p_footprint(int mode, int NX, int NY, int precision, int nsamp, int nloc,
                  long long* steady, long long* peak) {
  if (NX < 2 || NY < 2 || (precision != M2P_FP64 && precision != M2P_FP32)) return M2P_EINVAL;
  const long long P = (long long)(NX + 2 * MD) * (NY + 2 * MD);
  const long long full = (long long)m2p::MC_COUNT * P * 8, cells = (long long)m2p::MC_CELL_END * P * 8;
  long long b = 4LL * m2p::mass_blocks() * 8 + (long long)sizeof(DevState) + 4096;
  b += (P / 256 + P / 189 + 4LL * (NX + NY) + 64) * 8;   // per-block partial maxima
  b += 2 * 7LL * (NX + NY + 4 * MD);   // ghost-fill flags
  // dry flags; plans, sized for the most tiles (3-row tiles)
  b += 2LL * NY * ((NX + 62) / 63) + 4 * (16 + 16LL * m2p::plan_per(((NX + 62) / 63) * ((NY + 2) / 3)));
  if (precision == M2P_FP64)
    b += full + 16 * P * 8;
  else
    b += cells + 16 * P * 4 + (long long)m2p::MC_COUNT * P * 4 + P;
  if (mode == 1) b += (3LL * nsamp * nloc + 3LL * nloc + nsamp) * 8 + 3LL * nloc * 4;
  // set-up temporaries: topography, depth, per-cell phi, and the FP64 initial state (FP32)
  const long long tmp = 3 * P * 8 + (precision == M2P_FP32 ? 7 * P * 8 : 0);
  if (steady) *steady = b;
  if (peak) *peak = b + tmp;
  return 0;
}

// ---- checkpoint / resume ------------------------------------------------------------

namespace {

constexpr char kMagic[8] = {'M', '2', 'P', 'C', 'K', 'P', 'T', '1'};

uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; i++) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct CkptHeader {
  char magic[8];
  int32_t version, mode, NX, NY, prec, ntf;
  int64_t P;
  uint32_t devstate_bytes, lim_bytes_per_cell;
  uint64_t scen_hash;   // scenario_hash of the context that wrote it
};
constexpr int32_t kCkptVersion = 2;

}  // namespace

int m2p_checkpoint(m2p_ctx* c, const char* path) {
  if (!c || !c->s) return M2P_ESTATE;
  if (!path) return fail(c, M2P_EINVAL, "null path");
  hipStream_t s = strm(c);
  const size_t P = (size_t)c->P, nst = 7 * P * c->es;
  std::vector<char> u0(nst), u1(nst), lim(c->dLIM ? P : 0);
  DevState st{};
  HIPCHK(c, hipMemcpyAsync(&st, c->dSt, sizeof st, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(u0.data(), c->dU0, nst, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(u1.data(), c->dU1, nst, hipMemcpyDeviceToHost, s));
  if (c->dLIM) HIPCHK(c, hipMemcpyAsync(lim.data(), c->dLIM, P, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  CkptHeader h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = kCkptVersion;
  h.mode = c->mode;
  h.NX = c->NX;
  h.NY = c->NY;
  h.prec = c->prec;
  h.ntf = (int32_t)c->tf.size();
  h.P = c->P;
  h.devstate_bytes = sizeof(DevState);
  h.lim_bytes_per_cell = c->dLIM ? 1 : 0;
  h.scen_hash = c->scen_hash;
  std::vector<double> ot(c->outtime);
  ot.resize(c->tf.size() + 1, 0.0);
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return fail(c, M2P_EIO, "can't write " + tmp);
  uint64_t hsum = 1469598103934665603ull;
  auto put = [&](const void* p, size_t n) {
    hsum = fnv1a(hsum, p, n);
    return std::fwrite(p, 1, n, f) == n;
  };
  bool ok = put(&h, sizeof h) && put(&st, sizeof st) && put(ot.data(), ot.size() * sizeof(double)) &&
            put(u0.data(), nst) && put(u1.data(), nst) && (lim.empty() || put(lim.data(), P));
  ok = ok && std::fwrite(&hsum, 1, sizeof hsum, f) == sizeof hsum;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return fail(c, M2P_EIO, std::string("can't write ") + path);
  }
  return 0;
}

int m2p_restore(m2p_ctx* c, const char* path) {
  if (!c || !c->s) return M2P_ESTATE;
  IN_BATCH(c);
  if (!path) return fail(c, M2P_EINVAL, "null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, M2P_EIO, std::string("can't open ") + path);
  uint64_t hsum = 1469598103934665603ull;
  auto get = [&](void* p, size_t n) {
    if (std::fread(p, 1, n, f) != n) return false;
    hsum = fnv1a(hsum, p, n);
    return true;
  };
  CkptHeader h{};
  DevState st{};
  const size_t P = (size_t)c->P, nst = 7 * P * c->es;
  std::string why;
  if (!get(&h, sizeof h) || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kCkptVersion)
    why = "not a checkpoint of this format";
  else if (h.mode != c->mode || h.NX != c->NX || h.NY != c->NY || h.prec != c->prec ||
           h.P != c->P || h.ntf != (int32_t)c->tf.size() || h.devstate_bytes != sizeof(DevState) ||
           h.lim_bytes_per_cell != (c->dLIM ? 1u : 0u))
    why = "checkpoint does not match this scenario (mode, grid, precision or output times)";
  else if (h.scen_hash != c->scen_hash)
    why = "checkpoint does not match this scenario (topography, parameters or inflow tables)";
  std::vector<double> ot(c->tf.size() + 1);
  std::vector<char> u0(nst), u1(nst), lim(c->dLIM ? P : 0);
  if (why.empty() && !(get(&st, sizeof st) && get(ot.data(), ot.size() * sizeof(double)) &&
                       get(u0.data(), nst) && get(u1.data(), nst) && (lim.empty() || get(lim.data(), P))))
    why = "checkpoint is truncated";
  uint64_t stored = 0;
  if (why.empty() && (std::fread(&stored, 1, sizeof stored, f) != sizeof stored || stored != hsum))
    why = "checkpoint checksum mismatch";
  std::fclose(f);
  if (!why.empty()) return fail(c, M2P_EINVAL, why + ": " + path);
  HIPCHK(c, hipMemcpyAsync(c->dSt, &st, sizeof st, hipMemcpyHostToDevice, c->s));
  HIPCHK(c, hipMemcpyAsync(c->dU0, u0.data(), nst, hipMemcpyHostToDevice, c->s));
  HIPCHK(c, hipMemcpyAsync(c->dU1, u1.data(), nst, hipMemcpyHostToDevice, c->s));
  if (c->dLIM) HIPCHK(c, hipMemcpyAsync(c->dLIM, lim.data(), P, hipMemcpyHostToDevice, c->s));
  zero_ghost_momenta(c, c->s);
  rescan_dry(c, c->dry, c->s);
  HIPCHK(c, hipStreamSynchronize(c->s));
  c->outtime = ot;
  c->vstop_on = st.vstop2 > 0;
  return 0;
}

const char* m2p_last_error(const m2p_ctx* c) {
  if (c) return c->err.c_str();
  const char* e = m2p::context_free_error();
  return *e ? e : "null context";
}

void m2p_close(m2p_ctx* c) {
  if (!c) return;
  if (c->batch) (void)m2p_batch_remove(c->batch, c);
  release(c);
  delete c;
}

}  // extern "C"
