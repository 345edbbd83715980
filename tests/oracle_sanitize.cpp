// oracle_sanitize.cpp -- drives the CPU oracle (oracle/m2p_oracle.cpp, test infrastructure)
// through every entry point on a small landslide and a small inflow case, for
// tests/test_sanitizers.py, which builds it with AddressSanitizer and UndefinedBehavior-
// Sanitizer (host code only; no GPU).  Exit 0 when every step ran and the states are finite.
#include <cmath>
#include <cstdio>
#include <vector>

extern "C" {
void orc_pad_input(int NX, int NY, const double* raw, double* out);
void* orc_new(int mode, int nx, int ny, double dx_scaled, const double* topo, const double* depth,
              const double* phi, const double* par, int n_tf, const double* tf, int nloc,
              int nsamp, const double* inflow, const double* loc, const int* dire,
              const double* inflowTime);
void orc_free(void* h);
int orc_step(void* h, int n);
void orc_faces(void* h);
void orc_stage_terms(void* h);
void orc_inject(void* h);
int orc_nxd(void* h);
int orc_nyd(void* h);
int orc_get(void* h, const char* name, double* out);
void orc_set_u(void* h, const double* in);
int orc_threads(int n);
double orc_scalar(void* h, const char* name);
}

static const char* kNames[] = {
    "u", "uone", "uzero", "bfkt", "posx", "posy", "dxdxi11", "dxdxi12", "dxdxi21", "dxdxi22",
    "dbdx", "dbdy", "cvalue", "svec", "Detmin", "invJ11", "invJ12", "invJ13", "invJ21", "invJ22",
    "invJ23", "invJ31", "invJ32", "invJ33", "J13dxi", "J23dxi", "J33dxi", "J13det", "J23det",
    "J33det", "tande", "Jacb31", "Jacb32", "ux", "uy", "uE", "uW", "uN", "uS", "Hpx", "Hpy",
    "Ppx", "Ppy", "PDx", "PDy", "s", "Val", "v", "apEW", "apSN", "dudxE", "Ac", "Npress1"};

static int run_case(int mode) {
  const int NX = 40, NY = 32;
  const double cell = 20.0;
  std::vector<double> topo(NX * NY), depth(NX * NY, 0.0);
  for (int j = 0; j < NY; j++)
    for (int i = 0; i < NX; i++) {
      const double x = i * cell, y = j * cell, xc = NX * cell / 2;
      topo[j * NX + i] = mode == 0 ? 900.0 - 0.4 * x + 20.0 * std::exp(-((x - 500) * (x - 500) + (y - 300) * (y - 300)) / 20000.0)
                                   : 800.0 - 0.3 * y + 0.5 * std::fabs(x - xc);
      const double r2 = (x - 200) * (x - 200) + (y - 320) * (y - 320);
      if (mode == 0 && r2 < 100.0 * 100.0) depth[j * NX + i] = 10.0 * (1.0 - r2 / 10000.0);
    }
  const int nxd = NX + 6, nyd = NY + 6;
  std::vector<double> ptopo(nxd * nyd), pdepth(nxd * nyd);
  orc_pad_input(NX, NY, topo.data(), ptopo.data());
  orc_pad_input(NX, NY, depth.data(), pdepth.data());
  const double par[9] = {20.0, 6.0, 268.0, 5.0, 0.5, 0.1, 1.42 / 2.6, mode == 0 ? 0.001 : 0.002, 2.0};
  const double tf[4] = {0.0, 20.0, 40.0, 60.0};
  // inflow: 4 locations on the upstream edge (Y = 0), 5 samples
  const int nloc = 4, nsamp = 5;
  std::vector<double> table(3 * nsamp * nloc), loc(3 * nloc), itime(nsamp);
  std::vector<int> dire(3 * nloc);
  for (int k = 0; k < nsamp; k++) {
    itime[k] = 100.0 * k;
    for (int l = 0; l < nloc; l++) {
      table[0 * nsamp * nloc + k * nloc + l] = 10.0 * (1.0 + k);
      table[1 * nsamp * nloc + k * nloc + l] = 0.0;
      table[2 * nsamp * nloc + k * nloc + l] = 3.0;
    }
  }
  for (int l = 0; l < nloc; l++) {
    loc[3 * l] = l + 1;
    loc[3 * l + 1] = NX / 2 + 1 + l;
    loc[3 * l + 2] = 0;
    dire[3 * l] = l + 1;
    dire[3 * l + 1] = -5;
    dire[3 * l + 2] = 0;
  }
  void* h = orc_new(mode, NX, NY, cell * 10, ptopo.data(), mode == 0 ? pdepth.data() : nullptr,
                    nullptr, par, 4, tf, mode == 1 ? nloc : 0, mode == 1 ? nsamp : 0,
                    table.data(), loc.data(), dire.data(), itime.data());
  if (orc_nxd(h) != nxd || orc_nyd(h) != nyd) return 10;
  const size_t P = (size_t)nxd * nyd;
  std::vector<double> buf(7 * P);
  orc_inject(h);
  orc_faces(h);
  orc_stage_terms(h);
  int steps = orc_step(h, 60);
  for (const char* n : kNames)
    if (orc_get(h, n, buf.data()) < 0) return 11;
  orc_get(h, "u", buf.data());
  for (double x : buf)
    if (!std::isfinite(x)) return 12;
  orc_set_u(h, buf.data());
  steps += orc_step(h, 20);
  std::printf("mode %d: %d steps, time %.6f, iter %g\n", mode, steps, orc_scalar(h, "time"),
              orc_scalar(h, "iter"));
  orc_free(h);
  return steps > 0 ? 0 : 13;
}

int main() {
  orc_threads(1);
  int rc = run_case(0);
  if (!rc) rc = run_case(1);
  return rc;
}
