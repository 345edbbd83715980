// io_check.cpp -- drives the product's file readers and writers (m2p_io.cpp) without a
// GPU, for tests/test_io.py.  Usage:
//   io_check par <mode> <par_List>
//   io_check asc <path> <float:0|1>
//   io_check hydro <par_List>          (paths in the par_List relative to the cwd)
//   io_check write <dir>               (writes result files for a synthetic state)
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "../moses_2pdf_amd/csrc/m2p_io.h"

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  std::string cmd = argv[1];
  if (cmd == "par") {
    m2p::ParList p;
    std::string e = m2p::read_par_list(argv[3], std::atoi(argv[2]), p);
    if (!e.empty()) { std::printf("ERR %s\n", e.c_str()); return 1; }
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n%s\n%s\n%s\n%s\n%s\n%s\n", p.TotalSim,
                p.EachOut, p.delta0, p.Cd, p.N_R, p.varTheta, p.phiS0, p.dem.c_str(), p.init.c_str(),
                p.inH.c_str(), p.inU.c_str(), p.inV.c_str(), p.inLoc.c_str());
    return 0;
  }
  if (cmd == "asc") {
    m2p::AscGrid g;
    std::string e = m2p::read_asc(argv[2], std::atoi(argv[3]) != 0, g);
    if (!e.empty()) { std::printf("ERR %s\n", e.c_str()); return 1; }
    std::printf("%d %d %.17g %.17g %.17g\n", g.NX, g.NY, g.xll, g.yll, g.cellsize);
    for (double v : g.v) std::printf("%.17g\n", v);
    return 0;
  }
  if (cmd == "hydro") {
    m2p::ParList p;
    std::string e = m2p::read_par_list(argv[2], 1, p);
    if (e.empty()) {
      m2p::Hydro h;
      e = m2p::read_hydro(p, h);
      if (e.empty()) {
        std::printf("%d %d\n", h.nsamp, h.nloc);
        for (int j = 0; j < h.nsamp; j++) std::printf("%.17g\n", h.t[j]);
        for (double v : h.h) std::printf("%.17g\n", v);
        for (double v : h.u) std::printf("%.17g\n", v);
        for (double v : h.v) std::printf("%.17g\n", v);
        for (double v : h.loc) std::printf("%.17g\n", v);
        return 0;
      }
    }
    std::printf("ERR %s\n", e.c_str());
    return 1;
  }
  if (cmd == "write") {
    const int nx = 4, ny = 3, nxd = nx + 6, nyd = ny + 6;
    const size_t P = (size_t)nxd * nyd;
    std::vector<double> u(7 * P);
    for (size_t i = 0; i < u.size(); i++) u[i] = std::sin(0.37 * (double)i) * 3.0;
    for (size_t i = 0; i < P; i++) {
      if (i % 5 == 0) u[i] = 0.0;           // dry solid
      if (i % 7 == 0) u[3 * P + i] = 1e-6;  // below the 1e-5 output threshold
    }
    m2p::OutputGrid og{nx, ny, nxd, nyd};
    std::vector<double> t = {0.0, 30.25, 60.5};
    std::string e = m2p::write_result(argv[2], 7, og, u.data());
    if (e.empty()) e = m2p::write_result(argv[2], 1234, og, u.data());   // past 999
    if (e.empty()) e = m2p::write_time(argv[2], t, 3);
    if (e.empty()) e = m2p::write_info(argv[2], 185, 111, 200.0, 200.0, 212862.0625, 2561457.5, 61);
    std::vector<double> depth(P);
    for (size_t i = 0; i < P; i++) depth[i] = 0.25 * (double)i;
    if (e.empty()) e = m2p::write_init(argv[2], og, depth, 0.55);
    if (!e.empty()) { std::printf("ERR %s\n", e.c_str()); return 1; }
    return 0;
  }
  return 2;
}
