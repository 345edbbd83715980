// m2p_io.cpp -- readers and writers for the reference's file formats.
//
// Parse precision follows RunKernels (MoSES_2PDF_landslide/adNOC_2Pb_runKernels.cu, L/R;
// inflow fork I/R): topography, cellsize and corners through stof (float), initial depth
// through stod, hydrograph values through stof.
#include "m2p_io.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>

namespace m2p {

namespace {

// readParamLine (L/R:50-71): the line up to '#' or newline.
bool param_line(FILE* f, std::string& out) {
  char line[1024];
  if (!std::fgets(line, sizeof line, f)) return false;
  for (char* c = line; *c; c++)
    if (*c == '#' || *c == '\n') {
      *c = 0;
      break;
    }
  out = line;
  return true;
}

// split() (L/R:73-88): strtok semantics, runs of delimiters collapse.
std::vector<std::string> split(const std::string& s, const char* delim) {
  std::vector<std::string> sv;
  std::vector<char> buf(s.begin(), s.end());
  buf.push_back(0);
  for (char* p = std::strtok(buf.data(), delim); p; p = std::strtok(nullptr, delim)) sv.push_back(p);
  return sv;
}

bool read_lines(const std::string& path, std::vector<std::string>& lines) {
  std::ifstream f(path);
  if (f.fail()) return false;
  std::string s;
  while (std::getline(f, s)) lines.push_back(s);
  return true;
}

bool to_float(const std::string& s, double& v) {  // std::stof
  char* end = nullptr;
  errno = 0;
  const float x = std::strtof(s.c_str(), &end);
  if (end == s.c_str()) return false;
  v = (double)x;
  return true;
}

bool to_double(const std::string& s, double& v) {  // std::stod
  char* end = nullptr;
  const double x = std::strtod(s.c_str(), &end);
  if (end == s.c_str()) return false;
  v = x;
  return true;
}

bool to_int(const std::string& s, int& v) {  // std::stoi
  char* end = nullptr;
  const long x = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str()) return false;
  v = (int)x;
  return true;
}

std::string join(const std::string& dir, const char* name) {
  if (dir.empty()) return name;
  if (dir.back() == '/') return dir + name;
  return dir + "/" + name;
}

}  // namespace

std::string read_par_list(const std::string& path, int mode, ParList& p) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return "Parameter file " + path + " is missing";
  std::string line;
  double* nums[7] = {&p.TotalSim, &p.EachOut, &p.delta0, &p.Cd, &p.N_R, &p.varTheta, &p.phiS0};
  for (int i = 0; i < 7; i++) {
    if (!param_line(f, line) || std::sscanf(line.c_str(), "%lf", nums[i]) != 1) {
      std::fclose(f);
      return "par_List line " + std::to_string(i + 1) + " is not a number";
    }
  }
  std::vector<std::string*> paths;
  if (mode == 0)
    paths = {&p.dem, &p.init};
  else
    paths = {&p.dem, &p.inH, &p.inU, &p.inV, &p.inLoc};
  for (size_t i = 0; i < paths.size(); i++) {
    char buf[1024];
    if (!param_line(f, line) || std::sscanf(line.c_str(), "%1023s", buf) != 1) {
      std::fclose(f);
      return "par_List line " + std::to_string(8 + i) + " has no file name";
    }
    *paths[i] = buf;
  }
  std::fclose(f);
  if (p.EachOut <= 0) return "par_List: output interval must be positive";
  return "";
}

std::string read_asc(const std::string& path, bool as_float, AscGrid& g) {
  std::vector<std::string> L;
  if (!read_lines(path, L)) return "can't open grid file " + path;
  if (L.size() < 6) return "grid file " + path + " has no 6-line header";
  int start = 5;
  for (int i = 0; i < 6; i++) {
    std::vector<std::string> t = split(L[i], " ");
    if (t.empty()) return "grid file " + path + ": empty header line";
    bool ok = true;
    if (i == 0) ok = t.size() > 1 && to_int(t[1], g.NX);
    if (i == 1) ok = t.size() > 1 && to_int(t[1], g.NY);
    if (i == 2) ok = t.size() > 1 && to_float(t[1], g.xll);
    if (i == 3) ok = t.size() > 1 && to_float(t[1], g.yll);
    if (i == 4) ok = t.size() > 1 && to_float(t[1], g.cellsize);
    if (i == 5) start = (t[0] == "NODATA_value") ? 6 : 5;
    if (!ok) return "grid file " + path + ": bad header line " + std::to_string(i + 1);
  }
  if (g.NX < 2 || g.NY < 2) return "grid file " + path + ": grid must be at least 2x2";
  g.v.assign((size_t)g.NX * g.NY, 0.0);
  for (int j = 0; j < g.NY; j++) {
    if ((size_t)(j + start) >= L.size()) return "grid file " + path + ": too few rows";
    std::vector<std::string> t = split(L[j + start], " ");
    if ((int)t.size() < g.NX) return "grid file " + path + ": short row " + std::to_string(j + 1);
    for (int i = 0; i < g.NX; i++) {
      double v;
      if (!(as_float ? to_float(t[i], v) : to_double(t[i], v)))
        return "grid file " + path + ": bad value on row " + std::to_string(j + 1);
      g.v[(size_t)j * g.NX + i] = v;
    }
  }
  return "";
}

std::string read_asc_values(const std::string& path, int NX, int NY, bool as_float,
                            std::vector<double>& out) {
  std::vector<std::string> L;
  if (!read_lines(path, L)) return "can't open grid file " + path;
  if (L.size() < 6) return "grid file " + path + " has no 6-line header";
  std::vector<std::string> t5 = split(L[5], " ");
  const int start = (!t5.empty() && t5[0] == "NODATA_value") ? 6 : 5;
  out.assign((size_t)NX * NY, 0.0);
  for (int j = 0; j < NY; j++) {
    if ((size_t)(j + start) >= L.size()) return "grid file " + path + ": too few rows";
    std::vector<std::string> t = split(L[j + start], " ");
    if ((int)t.size() < NX) return "grid file " + path + ": short row " + std::to_string(j + 1);
    for (int i = 0; i < NX; i++) {
      double v;
      if (!(as_float ? to_float(t[i], v) : to_double(t[i], v)))
        return "grid file " + path + ": bad value on row " + std::to_string(j + 1);
      out[(size_t)j * NX + i] = v;
    }
  }
  return "";
}

std::string read_hydro(const ParList& pl, Hydro& h) {
  std::vector<std::string> LH, LU, LV, LL;
  if (!read_lines(pl.inH, LH)) return "can't open inflow depth file " + pl.inH;
  if (!read_lines(pl.inU, LU)) return "can't open inflow velocity (U) file " + pl.inU;
  if (!read_lines(pl.inV, LV)) return "can't open inflow velocity (V) file " + pl.inV;
  if (!read_lines(pl.inLoc, LL)) return "can't open inflow location file " + pl.inLoc;
  if (LH.size() < 2 || LL.size() < 2) return "inflow tables are empty";
  h.nsamp = (int)LH.size() - 1;
  h.nloc = (int)LL.size() - 1;
  if ((int)LU.size() - 1 < h.nsamp || (int)LV.size() - 1 < h.nsamp)
    return "inflow U/V tables are shorter than the depth table";
  h.t.assign(h.nsamp, 0.0);
  h.h.assign((size_t)h.nsamp * h.nloc, 0.0);
  h.u = h.h;
  h.v = h.h;
  for (int j = 0; j < h.nsamp; j++) {
    std::vector<std::string> a = split(LH[j + 1], "\t"), b = split(LU[j + 1], "\t"),
                             c = split(LV[j + 1], "\t");
    if ((int)a.size() < h.nloc + 1 || (int)b.size() < h.nloc + 1 || (int)c.size() < h.nloc + 1)
      return "inflow table row " + std::to_string(j + 1) + " is short";
    if (!to_float(a[0], h.t[j])) return "inflow table: bad time";
    for (int i = 0; i < h.nloc; i++) {
      const size_t q = (size_t)j * h.nloc + i;
      if (!to_float(a[i + 1], h.h[q]) || !to_float(b[i + 1], h.u[q]) || !to_float(c[i + 1], h.v[q]))
        return "inflow table: bad value";
    }
  }
  h.loc.assign((size_t)h.nloc * 3, 0.0);
  for (int j = 0; j < h.nloc; j++) {
    std::vector<std::string> a = split(LL[j + 1], "\t");
    if (a.size() < 3) return "inflow location row " + std::to_string(j + 1) + " is short";
    for (int i = 0; i < 3; i++)
      if (!to_float(a[i], h.loc[(size_t)j * 3 + i])) return "inflow location: bad value";
  }
  return "";
}

std::string write_dem(const std::string& dir, const OutputGrid& g, const std::vector<double>& bx,
                      const std::vector<double>& by, const std::vector<double>& topo,
                      const std::vector<double>& cvalue, const std::vector<double>& s0,
                      const std::vector<double>& s1) {
  const std::string p = join(dir, "DEM.dat");
  FILE* f = std::fopen(p.c_str(), "w");
  if (!f) return "can't open \"" + dir + "\" folder (need the result directory)";
  std::fprintf(f, "VARIABLES = \"x\", \"y\", \"z\", \"c\", \"S1\", \"S2\"\n ");
  for (int i = 3; i < g.nxd - 3; i++)
    for (int j = 3; j < g.nyd - 3; j++) {
      const size_t q = (size_t)j * g.nxd + i;
      std::fprintf(f, "%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\n", bx[q] * 0.1, by[q] * 0.1,
                   topo[q] * 0.1, cvalue[q], s0[q], s1[q]);
    }
  std::fclose(f);
  return "";
}

std::string write_init(const std::string& dir, const OutputGrid& g, const std::vector<double>& depth,
                       double phiS0) {
  const std::string p = join(dir, "001.dat");
  FILE* f = std::fopen(p.c_str(), "w");
  if (!f) return "can't open \"" + dir + "\" folder";
  std::fprintf(f, "VARIABLES = \"H\", \"phi\", \"Us\", \"Uf\", \"Vs\", \"Vf\"\n ");
  for (int i = 3; i < g.nxd - 3; i++)
    for (int j = 3; j < g.nyd - 3; j++)
      std::fprintf(f, "%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\n",
                   0.1 * depth[(size_t)j * g.nxd + i], phiS0, 0.0, 0.0, 0.0, 0.0);
  std::fclose(f);
  return "";
}

std::string write_info(const std::string& dir, int NX, int NY, double dx, double dy, double xll,
                       double yll, int nout) {
  const std::string p = join(dir, "Info.dat");
  FILE* f = std::fopen(p.c_str(), "w");
  if (!f) return "can't open \"" + dir + "\" folder";
  std::fprintf(f, "VARIABLES = \"x-point\", \"y-point\", \"dx\", \"dy\", \"xllcorner\", \"yllcorner\", \"TotalStep\"\n ");
  std::fprintf(f, "\t%d\t\t %d\t %10.2f\t %10.2f\t %10.4f\t %10.4f\t %d", NX, NY, (dx * 0.1),
               (dy * 0.1), xll, yll, nout);
  std::fclose(f);
  return "";
}

std::string write_time(const std::string& dir, const std::vector<double>& outtime, int n) {
  const std::string p = join(dir, "Time.dat");
  FILE* f = std::fopen(p.c_str(), "w");
  if (!f) return "can't write " + p;
  for (int nn = 0; nn < n; nn++) std::fprintf(f, "%20.4f", outtime[nn]);
  std::fclose(f);
  return "";
}

std::string write_result(const std::string& dir, int iter, const OutputGrid& g, const double* u) {
  char name[32];
  std::snprintf(name, sizeof name, "%03d.dat", iter);
  const std::string p = join(dir, name);
  FILE* f = std::fopen(p.c_str(), "w");
  if (!f) return "can't write " + p;
  const size_t P = (size_t)g.nxd * g.nyd;
  std::fprintf(f, "VARIABLES = \"H\", \"phi\", \"Us\", \"Uf\", \"Vs\", \"Vf\"\n ");
  for (int i = 3; i < g.nxd - 3; i++)
    for (int j = 3; j < g.nyd - 3; j++) {
      const size_t q = (size_t)j * g.nxd + i;
      const double hs = u[0 * P + q], hf = u[3 * P + q];
      const double H = 0.1 * (hs + hf);
      double us, vs, phi, uf, vf;
      if (hs > 0.00001) {
        vs = u[2 * P + q] / hs;
        us = u[1 * P + q] / hs;
        phi = u[6 * P + q];
      } else {
        vs = 0.0;
        us = 0.0;
        phi = 0.0;
      }
      if (hf > 0.00001) {
        vf = u[5 * P + q] / hf;
        uf = u[4 * P + q] / hf;
      } else {
        vf = 0.0;
        uf = 0.0;
      }
      std::fprintf(f, "%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\t%10.4f\n", H, phi, us, uf, vs, vf);
    }
  std::fclose(f);
  return "";
}

}  // namespace m2p
