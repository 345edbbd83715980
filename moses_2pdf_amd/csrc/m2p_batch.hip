// m2p_batch.hip -- batched ensembles behind the C ABI (m2p_batch_*): scenarios of uneven
// grids, each an m2p_ctx with its own buffers and DevState, stepped together by one set of
// launches per time step, captured in a HIP graph, with the stage kernel's tile shape,
// occupancy cap and block schedule picked by an autotuner that times the candidates on
// the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/moses2pdf.h"
#include "m2p_batch.h"
#include "m2p_ctx.h"
#include "m2p_tiles.h"

namespace m2p {

static const StageVariant kVar64[] = {{7, 4}, {3, 4}, {7, 3}};
static const StageVariant kVar32[] = {{7, 6}, {7, 4}, {3, 6}, {15, 4}, {3, 8}, {7, 8}};

int stage_variants(int prec, const StageVariant** out) {
  if (prec == M2P_FP32) {
    *out = kVar32;
    return (int)(sizeof kVar32 / sizeof kVar32[0]);
  }
  *out = kVar64;
  return (int)(sizeof kVar64 / sizeof kVar64[0]);
}

int ring_blocks(const Grid& g) { return (ring_threads(g) + 63) / 64; }
int box_blocks_x(const Grid& g) { return (g.nx + 63) / 64; }
int box_blocks_y(const Grid& g) { return (g.ny + 3) / 4; }

}  // namespace m2p

struct m2p_batch {
  int prec = M2P_FP64;
  hipStream_t s = nullptr;
  std::vector<m2p_ctx*> mem;   // in order of addition
  // schedule
  int ty = 7, w = 4, remap = 1, order = 1, graph_steps = 16;
  int cfl_blocks = m2p::BATCH_CFL_BLOCKS;
  int plan_cap = m2p::BATCH_PLAN_CAP;
  bool dirty = true;   // members changed: the block maps must be rebuilt
  int built_ty = 0, built_order = -1, built_w = 0;   // schedule the device tables are for
  // device tables
  m2p::Member* dMem = nullptr;
  size_t capMem = 0;
  int* dMap = nullptr;
  size_t capMap = 0;
  double* dPart = nullptr;
  size_t capPart = 0;
  int* dPlan = nullptr;   // stage plans under dry-tile skipping
  size_t capPlan = 0;
  // m2p_batch_status: the members' DevState pointers and a gathered copy of the states
  DevState** dStPtr = nullptr;
  DevState* dStAll = nullptr;
  size_t capSt = 0;
  m2p::BatchDev B{};
  bool any_inflow = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_n = 0;
  std::string err, report;
};

hipStream_t m2p_batch_stream(m2p_batch* b) { return b->s; }
void m2p_batch_mark_dirty(m2p_batch* b) { b->dirty = true; }

namespace {

int bfail(m2p_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

#define BCHK(b, call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return bfail((b), M2P_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

void drop_graph(m2p_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_n = 0;
}

bool variant_ok(int prec, int ty, int w) {
  const m2p::StageVariant* v;
  const int n = m2p::stage_variants(prec, &v);
  for (int i = 0; i < n; i++)
    if (v[i].ty == ty && v[i].w == w) return true;
  return false;
}

// Member table and block maps for the current members and schedule.
int rebuild(m2p_batch* b) {
  drop_graph(b);
  if (!variant_ok(b->prec, b->ty, b->w))
    return bfail(b, M2P_EINVAL, "no stage kernel compiled for ty=" + std::to_string(b->ty) +
                                    " waves=" + std::to_string(b->w));
  const int n = (int)b->mem.size();
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  if (b->order == 1)  // largest grids first: the long tiles of big members start early
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int c) { return b->mem[a]->P > b->mem[c]->P; });
  std::vector<m2p::Member> M(n);
  std::vector<int> mg, mb, mr, ms;
  b->any_inflow = false;
  bool any_vstop = false, any_dry = false, all_uni = true;
  for (int k = 0; k < n; k++) {
    m2p_ctx* c = b->mem[ord[k]];
    m2p::Member& m = M[k];
    m = m2p::Member{};
    m.g = c->g;
    m.ph = c->ph;
    m.dp.cfl = c->par.cfl;
    m.dp.dt_min = c->par.dt_min;
    m.dp.dt_max = c->par.dt_max;
    m.dp.dx = c->dx;
    m.dp.dy = c->dy;
    m.dp.ntf = 0;
    m.dp.nsamp = c->mode == 1 ? c->nsamp : 0;
    m.mode = c->mode;
    m.nloc = c->nloc;
    m.nsamp = c->nsamp;
    m.nbx = m2p::box_blocks_x(c->g);
    m.phiS0 = c->par.phiS0;
    m.U0 = c->dU0;
    m.U1 = c->dU1;
    m.ONE = c->dONE;
    m.MT = c->dMT;
    m.M = c->dM;
    m.LIM = c->dLIM;
    m.dry = c->dry;
    m.st = c->dSt;
    m.tf = c->dTf;
    m.itime = c->dItime;
    m.table = c->dTable;
    m.loc = c->dLoc;
    m.dire = c->dDire;
    if (c->mode == 1) b->any_inflow = true;
    if (c->vstop_on) any_vstop = true;
    if (c->dry.nseg > 0) any_dry = true;
    if (!c->uni_metric) all_uni = false;
    m.box_n = m.nbx * m2p::box_blocks_y(c->g);
    m.ring_n = m2p::ring_blocks(c->g);
    m.stage_n = m2p::stage_tiles(c->g, b->ty);
    m.ghost0 = (int)mg.size();
    mg.insert(mg.end(), (c->nyd + c->nxd + m2p::GHOST_LINES - 1) / m2p::GHOST_LINES, k);
    m.box0 = (int)mb.size();
    mb.insert(mb.end(), (size_t)m.nbx * m2p::box_blocks_y(c->g), k);
    m.ring0 = (int)mr.size();
    mr.insert(mr.end(), m2p::ring_blocks(c->g), k);
    m.stage0 = (int)ms.size();
    ms.insert(ms.end(), m2p::stage_tiles(c->g, b->ty), k);
  }
  const size_t nmap = mg.size() + mb.size() + mr.size() + ms.size();
  if (ms.size() > (size_t)INT_MAX / 2) return bfail(b, M2P_EINVAL, "batch too large");
  if ((size_t)n > b->capMem) {
    if (b->dMem) (void)hipFree(b->dMem);
    b->dMem = nullptr;
    b->capMem = std::max<size_t>(n, 2 * b->capMem);
    BCHK(b, hipMalloc((void**)&b->dMem, b->capMem * sizeof(m2p::Member)));
  }
  if (nmap > b->capMap) {
    if (b->dMap) (void)hipFree(b->dMap);
    b->dMap = nullptr;
    b->capMap = std::max(nmap, 2 * b->capMap);
    BCHK(b, hipMalloc((void**)&b->dMap, b->capMap * sizeof(int)));
  }
  const size_t npart = mb.size() + mr.size() + ms.size();
  if (npart > b->capPart) {
    if (b->dPart) (void)hipFree(b->dPart);
    b->dPart = nullptr;
    b->capPart = std::max(npart, 2 * b->capPart);
    BCHK(b, hipMalloc((void**)&b->dPart, b->capPart * sizeof(double)));
  }
  const int per8 = m2p::plan_per((int)ms.size());
  const size_t nplan = m2p::PLAN_COUNTERS + 16 * (size_t)per8;
  if (nplan > b->capPlan) {
    if (b->dPlan) (void)hipFree(b->dPlan);
    b->dPlan = nullptr;
    b->capPlan = std::max(nplan, 2 * b->capPlan);
    BCHK(b, hipMalloc((void**)&b->dPlan, b->capPlan * sizeof(int)));
    BCHK(b, hipMemsetAsync(b->dPlan, 0, b->capPlan * sizeof(int), b->s));
  }
  std::vector<int> all;
  all.reserve(nmap);
  for (auto* v : {&mg, &mb, &mr, &ms}) all.insert(all.end(), v->begin(), v->end());
  if (n) BCHK(b, hipMemcpyAsync(b->dMem, M.data(), n * sizeof(m2p::Member), hipMemcpyHostToDevice, b->s));
  if (nmap) BCHK(b, hipMemcpyAsync(b->dMap, all.data(), nmap * sizeof(int), hipMemcpyHostToDevice, b->s));
  BCHK(b, hipStreamSynchronize(b->s));
  m2p::BatchDev& B = b->B;
  B.mem = b->dMem;
  B.part = b->dPart;
  B.mapGhost = b->dMap;
  B.mapBox = b->dMap + mg.size();
  B.mapRing = b->dMap + mg.size() + mb.size();
  B.mapStage = b->dMap + mg.size() + mb.size() + mr.size();
  B.nmem = n;
  B.nGhost = (int)mg.size();
  B.nBox = (int)mb.size();
  B.nRing = (int)mr.size();
  B.nStage = (int)ms.size();
  B.remap = b->remap;
  B.cfl_blocks = b->cfl_blocks;
  B.plan_cap = b->plan_cap;
  B.vmax = any_vstop ? 1 : 0;
  B.plan_on = any_dry ? 1 : 0;
  B.uni = (all_uni && n > 0) ? 1 : 0;
  B.plan_cnt = b->dPlan;
  for (int io = 0; io < 2; io++) {
    B.plan[io].cnt = b->dPlan + 8 * io;
    B.plan[io].list = b->dPlan + m2p::PLAN_COUNTERS + 8 * (size_t)per8 * io;
    B.plan[io].per = per8;
  }
  b->dirty = false;
  b->built_ty = b->ty;
  b->built_order = b->order;
  b->built_w = b->w;
  return 0;
}

// Brings the device tables in line with the members and the schedule.  The block maps are
// rebuilt when the members, the tile height or the member order changed; the occupancy
// cap, XCD banding, CFL block count and plan cap only change kernel arguments (and drop a
// captured graph, which holds them).
int refresh(m2p_batch* b) {
  if (b->dirty || b->ty != b->built_ty || b->order != b->built_order) return rebuild(b);
  if (b->w != b->built_w || b->remap != b->B.remap || b->cfl_blocks != b->B.cfl_blocks ||
      b->plan_cap != b->B.plan_cap) {
    if (!variant_ok(b->prec, b->ty, b->w))
      return bfail(b, M2P_EINVAL, "no stage kernel compiled for ty=" + std::to_string(b->ty) +
                                      " waves=" + std::to_string(b->w));
    drop_graph(b);
    b->B.remap = b->remap;
    b->B.cfl_blocks = b->cfl_blocks;
    b->B.plan_cap = b->plan_cap;
    b->built_w = b->w;
  }
  return 0;
}

hipError_t enqueue(m2p_batch* b, int nsteps) {
  if (b->prec == M2P_FP32)
    return m2p::launch_batch_steps<float>(b->B, b->ty, b->w, b->any_inflow, nsteps, b->s);
  return m2p::launch_batch_steps<double>(b->B, b->ty, b->w, b->any_inflow, nsteps, b->s);
}

int ensure_graph(m2p_batch* b) {
  if (b->exec && b->graph_n == b->graph_steps) return 0;
  drop_graph(b);
  BCHK(b, hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
  const hipError_t e = enqueue(b, b->graph_steps);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->s, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(b, M2P_EHIP, std::string("graph capture: ") +
                                  hipGetErrorString(e != hipSuccess ? e : e2));
  }
  b->graph = g;
  BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_n = b->graph_steps;
  return 0;
}

int step_batch(m2p_batch* b, int nsteps, bool use_graph) {
  if (int rc = refresh(b)) return rc;
  if (b->mem.empty() || nsteps <= 0) return 0;
  int left = nsteps;
  if (use_graph && b->graph_steps > 0 && left >= b->graph_steps) {
    if (int rc = ensure_graph(b)) return rc;
    for (; left >= b->graph_steps; left -= b->graph_steps) BCHK(b, hipGraphLaunch(b->exec, b->s));
  }
  if (left) BCHK(b, enqueue(b, left));
  return 0;
}

// Saved member state for the autotuner's trial steps: every member's state buffers in one
// allocation (each part 256-byte aligned), and the copy lists that save and restore them
// (one launch each).
struct Snapshot {
  void* buf = nullptr;
  m2p::CopyItem* items = nullptr;   // device: [n] save items, then [n] restore items
  int n = 0;
  ~Snapshot() {
    if (buf) (void)hipFree(buf);
    if (items) (void)hipFree(items);
  }
};

int save(m2p_batch* b, Snapshot& sn, bool restore) {
  if (!restore) {
    struct Part {
      void* dev;
      size_t n, off;
    };
    std::vector<Part> parts;
    size_t total = 0;
    for (m2p_ctx* c : b->mem) {
      const size_t P = (size_t)c->P, es = c->es;
      const Part q[] = {{c->dU0, 7 * P * es, 0}, {c->dU1, 7 * P * es, 0}, {c->dONE, 2 * P * es, 0},
                        {c->dLIM, c->dLIM ? P : 0, 0}, {c->dSt, sizeof(DevState), 0},
                        {c->dZ, c->dZ ? 2 * (size_t)m2p::dry_flag_count(c->g) : 0, 0},
                        {c->dGF, c->dGF ? 2 * (size_t)m2p::ghost_flag_count(c->g) : 0, 0}};
      for (const Part& x : q) {
        if (x.n == 0 || !x.dev) continue;
        parts.push_back({x.dev, x.n, total});
        total += (x.n + 255) / 256 * 256;
      }
    }
    sn.n = (int)parts.size();
    BCHK(b, hipMalloc(&sn.buf, total ? total : 256));
    std::vector<m2p::CopyItem> it(2 * parts.size());
    unsigned char* base = (unsigned char*)sn.buf;
    for (size_t i = 0; i < parts.size(); i++) {
      it[i] = {(const unsigned char*)parts[i].dev, base + parts[i].off, parts[i].n};
      it[parts.size() + i] = {base + parts[i].off, (unsigned char*)parts[i].dev, parts[i].n};
    }
    BCHK(b, hipMalloc((void**)&sn.items, it.size() * sizeof(m2p::CopyItem) + 1));
    BCHK(b, hipMemcpy(sn.items, it.data(), it.size() * sizeof(m2p::CopyItem), hipMemcpyHostToDevice));
  }
  m2p::launch_copy_list(sn.items + (restore ? sn.n : 0), sn.n, b->s);
  BCHK(b, hipGetLastError());
  BCHK(b, hipStreamSynchronize(b->s));
  return 0;
}

}  // namespace

extern "C" {

int m2p_batch_create(int precision, m2p_batch** out) {
  if (!out || (precision != M2P_FP64 && precision != M2P_FP32)) return M2P_EINVAL;
  m2p_batch* b = new m2p_batch();
  b->prec = precision;
  // untuned default: 3-row tiles (the tuner's usual pick for sparse batches; a 64-request
  // FP32 serving stream 4% faster than with 7-row tiles at 6 waves), 8 waves in FP32
  b->ty = 3;
  b->w = precision == M2P_FP32 ? 8 : 4;
  const hipError_t e = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking);
  *out = b;
  if (e != hipSuccess) return bfail(b, M2P_EHIP, hipGetErrorString(e));
  return 0;
}

int m2p_batch_add(m2p_batch* b, m2p_ctx* c) {
  if (!b || !b->s || !c || !c->s) return M2P_ESTATE;
  if (c->batch) return bfail(b, M2P_ESTATE, "scenario is already in a batch");
  if (c->prec != b->prec) return bfail(b, M2P_EINVAL, "scenario precision differs from the batch's");
  // the scenario's own queued work must be done before the batch stream touches it
  BCHK(b, hipStreamSynchronize(c->s));
  c->batch = b;
  b->mem.push_back(c);
  b->dirty = true;
  return (int)b->mem.size() - 1;
}

int m2p_batch_remove(m2p_batch* b, m2p_ctx* c) {
  if (!b || !c) return M2P_ESTATE;
  auto it = std::find(b->mem.begin(), b->mem.end(), c);
  if (it == b->mem.end()) return bfail(b, M2P_EINVAL, "scenario is not in this batch");
  BCHK(b, hipStreamSynchronize(b->s));
  b->mem.erase(it);
  c->batch = nullptr;
  b->dirty = true;
  return 0;
}

int m2p_batch_size(const m2p_batch* b) { return b ? (int)b->mem.size() : M2P_ESTATE; }

int m2p_batch_status(m2p_batch* b, int n, double* t, int* nt, int* halted, int* nonfinite,
                     int* nonfinite_nt) {
  if (!b || !b->s) return M2P_ESTATE;
  if (n != (int)b->mem.size()) return bfail(b, M2P_EINVAL, "n must equal the number of members");
  if (n == 0) return 0;
  if ((size_t)n > b->capSt) {
    if (b->dStPtr) (void)hipFree(b->dStPtr);
    if (b->dStAll) (void)hipFree(b->dStAll);
    b->dStPtr = nullptr;
    b->dStAll = nullptr;
    b->capSt = std::max((size_t)n, 2 * b->capSt);
    BCHK(b, hipMalloc((void**)&b->dStPtr, b->capSt * sizeof(DevState*)));
    BCHK(b, hipMalloc((void**)&b->dStAll, b->capSt * sizeof(DevState)));
  }
  std::vector<DevState*> ptr((size_t)n);
  for (int k = 0; k < n; k++) ptr[k] = b->mem[k]->dSt;
  std::vector<DevState> st((size_t)n);
  BCHK(b, hipMemcpyAsync(b->dStPtr, ptr.data(), n * sizeof(DevState*), hipMemcpyHostToDevice, b->s));
  m2p::launch_gather_states(b->dStPtr, b->dStAll, n, b->s);
  BCHK(b, hipGetLastError());
  BCHK(b, hipMemcpyAsync(st.data(), b->dStAll, n * sizeof(DevState), hipMemcpyDeviceToHost, b->s));
  BCHK(b, hipStreamSynchronize(b->s));
  for (int k = 0; k < n; k++) {
    if (t) t[k] = st[k].time;
    if (nt) nt[k] = st[k].nt;
    if (halted) halted[k] = (st[k].skip || st[k].nt >= st[k].halt_at) ? 1 : 0;
    if (nonfinite) nonfinite[k] = st[k].nonfinite;
    if (nonfinite_nt) nonfinite_nt[k] = st[k].nonfinite ? st[k].nonfinite_nt : -1;
  }
  return 0;
}

int m2p_batch_step(m2p_batch* b, int nsteps) {
  if (!b || !b->s) return M2P_ESTATE;
  if (nsteps < 0) return bfail(b, M2P_EINVAL, "negative step count");
  return step_batch(b, nsteps, true);
}

int m2p_batch_sync(m2p_batch* b) {
  if (!b || !b->s) return M2P_ESTATE;
  BCHK(b, hipStreamSynchronize(b->s));
  return 0;
}

int m2p_batch_set(m2p_batch* b, const char* key, int value) {
  if (!b || !key) return M2P_ESTATE;
  const std::string k(key);
  if (k == "ty" && value > 0) {
    b->ty = value;   // the (ty, waves) pair is checked when the batch is next stepped
  } else if (k == "waves" && value > 0) {
    b->w = value;
  } else if (k == "remap" && (value == 0 || value == 1)) {
    b->remap = value;
  } else if (k == "order" && (value == 0 || value == 1)) {
    b->order = value;
  } else if (k == "graph_steps" && value >= 0) {
    b->graph_steps = value;
  } else if (k == "cfl_blocks" && value >= 64) {
    b->cfl_blocks = value;
  } else if (k == "plan_cap" && value >= 1) {
    b->plan_cap = value;
  } else {
    return bfail(b, M2P_EINVAL, "unknown key or bad value: " + k);
  }
  return 0;   // the next step brings the device tables in line (refresh)
}

int m2p_batch_get(const m2p_batch* b, const char* key) {
  if (!b || !key) return M2P_ESTATE;
  const std::string k(key);
  if (k == "ty") return b->ty;
  if (k == "waves") return b->w;
  if (k == "remap") return b->remap;
  if (k == "order") return b->order;
  if (k == "graph_steps") return b->graph_steps;
  if (k == "cfl_blocks") return b->cfl_blocks;
  if (k == "plan_cap") return b->plan_cap;
  if (k == "stage_blocks") return b->B.nStage;
  return M2P_EINVAL;
}

int m2p_batch_tune(m2p_batch* b, int trial_steps, char* report, int report_len) {
  if (!b || !b->s) return M2P_ESTATE;
  if (trial_steps < 1) return bfail(b, M2P_EINVAL, "trial_steps must be >= 1");
  if (b->mem.empty()) return bfail(b, M2P_ESTATE, "empty batch");
  Snapshot sn0;   // the caller's state, put back at the end
  int rc = save(b, sn0, false);
  // The trials start from the state one step on.  Before its first step a landslide batch
  // has no dry segment (phi starts at phiS0 on every cell), so each trial's warm-up step
  // would compute every tile; one step clears phi on the dry cells and sets exact flags.
  Snapshot sn;
  if (!rc) rc = step_batch(b, 1, false);
  if (!rc) rc = save(b, sn, false);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = bfail(b, M2P_EHIP, "hipEventCreate");
  // one step to warm up, trial_steps timed, then every member's state restored
  auto trial = [&](float* ms) {
    *ms = 0.f;
    if ((rc = step_batch(b, 1, false))) return;
    if (hipEventRecord(e0, b->s) != hipSuccess) rc = bfail(b, M2P_EHIP, "hipEventRecord");
    if (!rc) rc = step_batch(b, trial_steps, false);
    if (!rc && hipEventRecord(e1, b->s) != hipSuccess) rc = bfail(b, M2P_EHIP, "hipEventRecord");
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = bfail(b, M2P_EHIP, "hipEventSynchronize");
    if (!rc) (void)hipEventElapsedTime(ms, e0, e1);
    *ms /= trial_steps;
    if (!rc) rc = save(b, sn, true);
  };
  // A staged search, each trial one restored run: (1) every compiled stage schedule (tile
  // height x waves) with the member order and XCD block map on; (2) for the fastest, the
  // order off, then the map off, each kept if faster; (3) the CFL launch's block count,
  // doubling and halving from the default while that is faster.  The full grid (every
  // schedule with every order and map, every block count) took three times the trials and
  // on the 256-member ensemble chose the same schedule; the dimensions searched in stages
  // moved a trial by under 1% there.
  const int ty0 = b->ty, w0 = b->w, remap0 = b->remap, order0 = b->order, cb0 = b->cfl_blocks;
  const m2p::StageVariant* v;
  const int nv = m2p::stage_variants(b->prec, &v);
  std::vector<int> vi(nv);
  std::iota(vi.begin(), vi.end(), 0);
  std::stable_sort(vi.begin(), vi.end(), [&](int x, int y) { return v[x].ty < v[y].ty; });
  b->cfl_blocks = m2p::BATCH_CFL_BLOCKS;   // the stage schedules are timed at the default
  std::string rep;
  char line[160];
  float best_ms = 0.f;
  int bty = 0, bw = 0, bremap = 1, border = 1;
  auto run = [&](int ty, int w, int remap, int order) {
    b->ty = ty;
    b->w = w;
    b->remap = remap;
    b->order = order;
    float ms = 0.f;
    trial(&ms);
    std::snprintf(line, sizeof line, "ty=%d waves=%d remap=%d order=%d: %.4f ms/step\n", ty, w, remap,
                  order, ms);
    rep += line;
    if (!rc && (bty == 0 || ms < best_ms)) {
      best_ms = ms;
      bty = ty;
      bw = w;
      bremap = remap;
      border = order;
    }
  };
  for (int i : vi)
    if (!rc) run(v[i].ty, v[i].w, 1, 1);
  if (!rc) run(bty, bw, bremap, 0);
  if (!rc) run(bty, bw, 0, border);
  b->ty = bty;
  b->w = bw;
  b->remap = bremap;
  b->order = border;
  // then the CFL launch's block count, under the chosen stage schedule
  int best_cb = m2p::BATCH_CFL_BLOCKS;
  for (int dir : {2, -2}) {
    for (int cb = dir > 0 ? best_cb * 2 : best_cb / 2; cb >= 2048 && cb <= 32768 && !rc;
         cb = dir > 0 ? cb * 2 : cb / 2) {
      b->cfl_blocks = cb;
      float ms;
      trial(&ms);
      std::snprintf(line, sizeof line, "cfl_blocks=%d: %.4f ms/step\n", cb, ms);
      rep += line;
      if (rc || !(ms < best_ms)) break;
      best_ms = ms;
      best_cb = cb;
    }
    if (best_cb != m2p::BATCH_CFL_BLOCKS) break;   // doubling helped: halving will not
  }
  b->cfl_blocks = best_cb;
  if (rc) {   // a failed search leaves the caller's schedule
    b->ty = ty0;
    b->w = w0;
    b->remap = remap0;
    b->order = order0;
    b->cfl_blocks = cb0;
  }
  {
    const int rc0 = save(b, sn0, true);
    if (!rc) rc = rc0;
  }
  if (!rc) {
    std::snprintf(line, sizeof line,
                  "chosen: ty=%d waves=%d remap=%d order=%d cfl_blocks=%d (%.4f ms/step)\n", b->ty, b->w,
                  b->remap, b->order, b->cfl_blocks, best_ms);
    rep += line;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  b->report = rep;
  if (report && report_len > 0) {
    std::snprintf(report, (size_t)report_len, "%s", rep.c_str());
  }
  return rc;
}

const char* m2p_batch_last_error(const m2p_batch* b) { return b ? b->err.c_str() : "null batch"; }

void m2p_batch_destroy(m2p_batch* b) {
  if (!b) return;
  if (b->s) (void)hipStreamSynchronize(b->s);
  for (m2p_ctx* c : b->mem) c->batch = nullptr;
  drop_graph(b);
  if (b->dStPtr) (void)hipFree(b->dStPtr);
  if (b->dStAll) (void)hipFree(b->dStAll);
  if (b->dMem) (void)hipFree(b->dMem);
  if (b->dMap) (void)hipFree(b->dMap);
  if (b->dPart) (void)hipFree(b->dPart);
  if (b->dPlan) (void)hipFree(b->dPlan);
  if (b->s) (void)hipStreamDestroy(b->s);
  delete b;
}

}  // extern "C"
