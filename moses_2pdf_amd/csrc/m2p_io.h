// m2p_io.h -- the reference's on-disk formats: par_List, ESRI ASCII grids, hydrograph
// tables, and the result2Pb text outputs.
#pragma once

#include <string>
#include <vector>

namespace m2p {

struct ParList {
  double TotalSim = 0, EachOut = 0, delta0 = 0, Cd = 0, N_R = 0, varTheta = 0, phiS0 = 0;
  std::string dem, init;                 // landslide
  std::string inH, inU, inV, inLoc;      // inflow
};

// RunKernels::readinParamFile (L/R:11-48, I/R:11-57).  Returns "" or an error message.
std::string read_par_list(const std::string& path, int mode, ParList& out);

struct AscGrid {
  int NX = 0, NY = 0;
  double xll = 0, yll = 0, cellsize = 0;
  std::vector<double> v;  // [j*NX+i]
};

// dataInitialization's reader (L/R:128-278): header lines 0-4 (ncols, nrows, xllcorner,
// yllcorner, cellsize parsed with stoi/stof), data from line 5, or line 6 when line 5
// starts with NODATA_value; values parsed with stof (as_float) or stod.
std::string read_asc(const std::string& path, bool as_float, AscGrid& out);
// Only the NODATA test of a grid's line 5 (the init file's header is otherwise unused).
std::string read_asc_values(const std::string& path, int NX, int NY, bool as_float,
                            std::vector<double>& out);

struct Hydro {
  int nsamp = 0, nloc = 0;
  std::vector<double> t, h, u, v;  // t[nsamp]; h,u,v[nsamp*nloc]
  std::vector<double> loc;         // [nloc*3]
};

// Inflow tables (I/R:170-188, 297-331): one header line skipped, tab-separated values
// parsed with stof.
std::string read_hydro(const ParList& pl, Hydro& out);

// result2Pb writers (L/R:543-596, 1260-1324).
struct OutputGrid {
  int nx, ny, nxd, nyd;
};
std::string write_dem(const std::string& dir, const OutputGrid& g, const std::vector<double>& bx,
                      const std::vector<double>& by, const std::vector<double>& topo,
                      const std::vector<double>& cvalue, const std::vector<double>& s0,
                      const std::vector<double>& s1);
std::string write_init(const std::string& dir, const OutputGrid& g, const std::vector<double>& depth,
                       double phiS0);
std::string write_info(const std::string& dir, int NX, int NY, double dx, double dy, double xll,
                       double yll, int nout);
std::string write_time(const std::string& dir, const std::vector<double>& outtime, int n);
std::string write_result(const std::string& dir, int iter, const OutputGrid& g, const double* u);

}  // namespace m2p
