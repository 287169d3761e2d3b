// io.cpp -- the reference's file formats (see io.hpp for the citations).
#include "io.hpp"

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

namespace cbns_io {

namespace {

std::string trim(const std::string& s) {
  const char* ws = " \t\r\n";
  const size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}

int bc_of(const std::string& v) {
  if (v == "periodic" || v == "1") return 1;
  return 0;
}

// keys of python/rae2822.ini in the order python/parser_ini.py writes them
const char* const kPositional[17] = {"grid_fname", "checkpoint_fname", "CFL",   "t_cur",
                                     "t_end",      "Pr",               "Pr_t",  "gamma",
                                     "Mach",       "Re",               "T_inf", "angle_attack",
                                     "nstep_max",  "checkpoint_freq",  "i_begin", "i_end",
                                     "is_restart"};

bool apply(Input& in, const std::string& key, const std::string& val, std::string& err) {
  std::istringstream is(val);
  auto num = [&](double& d) { return static_cast<bool>(is >> d); };
  auto lng = [&](long& l) { return static_cast<bool>(is >> l); };
  auto integ = [&](int& i) { return static_cast<bool>(is >> i); };
  bool ok = true;
  if (key == "grid_fname") in.grid_fname = val;
  else if (key == "checkpoint_fname") in.checkpoint_fname = val;
  else if (key == "vtk_fname") in.vtk_fname = val;
  else if (key == "CFL") ok = num(in.CFL);
  else if (key == "t_cur") ok = num(in.t_cur);
  else if (key == "t_end") ok = num(in.t_end);
  else if (key == "Pr") ok = num(in.Pr);
  else if (key == "Pr_t") ok = num(in.Pr_t);
  else if (key == "gamma") ok = num(in.gamma);
  else if (key == "Mach") ok = num(in.Mach);
  else if (key == "Re") ok = num(in.Re);
  else if (key == "T_inf") ok = num(in.T_inf);
  else if (key == "angle_attack") ok = num(in.angle_attack);
  else if (key == "nstep_max") ok = lng(in.nstep_max);
  else if (key == "checkpoint_freq") ok = lng(in.checkpoint_freq);
  else if (key == "i_begin") ok = integ(in.i_begin);
  else if (key == "i_end") ok = integ(in.i_end);
  else if (key == "is_restart") {
    int b = 0;
    ok = integ(b);
    in.is_restart = b != 0;
  } else if (key == "viscous") ok = integ(in.viscous);
  else if (key == "visc_centered") ok = integ(in.visc_centered);
  else if (key == "bc_xi") in.bc[0] = bc_of(val);
  else if (key == "bc_eta") in.bc[1] = bc_of(val);
  else if (key == "bc_zeta") in.bc[2] = bc_of(val);
  else if (key == "period_xi" || key == "period_eta" || key == "period_zeta") {
    const int a = key == "period_xi" ? 0 : (key == "period_eta" ? 1 : 2);
    ok = num(in.period[3 * a]) && num(in.period[3 * a + 1]) && num(in.period[3 * a + 2]);
  }
  // unknown keys are ignored, as the reference's positional reader would
  if (!ok) err = "bad value for " + key + ": '" + val + "'";
  return ok;
}

}  // namespace

bool read_input(const std::string& path, Input& in, std::string& err) {
  std::ifstream fh(path);
  if (!fh) {
    err = "cannot open " + path;
    return false;
  }
  std::stringstream buf;
  buf << fh.rdbuf();
  const std::string text = buf.str();
  if (text.find('=') == std::string::npos) {
    // positional input.txt (src/Block3d.cpp:44-92)
    std::istringstream is(text);
    std::string tok;
    int n = 0;
    while (n < 17 && is >> tok) {
      if (!apply(in, kPositional[n], tok, err)) return false;
      n++;
    }
    if (n < 17) {
      err = "input.txt: expected 17 values, found " + std::to_string(n);
      return false;
    }
    return true;
  }
  std::istringstream is(text);
  std::string line;
  std::map<std::string, bool> seen;
  while (std::getline(is, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '#' || line[0] == ';' || line[0] == '[') continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
    if (!apply(in, key, val, err)) return false;
    seen[key] = true;
  }
  for (const char* k : {"grid_fname", "CFL", "t_end", "gamma", "Mach", "nstep_max"})
    if (!seen.count(k)) {
      err = std::string("missing key ") + k;
      return false;
    }
  return true;
}

bool read_plot3d(const std::string& path, int& IM, int& JM, int& KM, std::vector<double>& x,
                 std::vector<double>& y, std::vector<double>& z, std::string& err) {
  std::ifstream fh(path);
  if (!fh) {
    err = "cannot open " + path;
    return false;
  }
  if (!(fh >> IM >> JM >> KM) || IM < 1 || JM < 1 || KM < 1) {
    err = "bad Plot3D header in " + path;
    return false;
  }
  const size_t n = (size_t)IM * JM * KM;
  std::vector<double>* arr[3] = {&x, &y, &z};
  for (auto* a : arr) {
    a->resize(n);
    for (size_t q = 0; q < n; q++)
      if (!(fh >> (*a)[q])) {
        err = "Plot3D file " + path + " ends early";
        return false;
      }
  }
  return true;
}

bool write_plot3d(const std::string& path, int IM, int JM, int KM, const double* x,
                  const double* y, const double* z) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "%d %d %d\n", IM, JM, KM);
  const size_t n = (size_t)IM * JM * KM;
  for (const double* a : {x, y, z})
    for (size_t q = 0; q < n; q++) std::fprintf(f, "%.17g\n", a[q]);
  return std::fclose(f) == 0;
}

bool write_checkpoint(const std::string& path, double t_cur, uint32_t IM_G, uint32_t JM_G,
                      uint32_t KM_G, const double* Q) {
  std::ofstream fh(path, std::ios::out | std::ios::binary);
  if (!fh) return false;
  fh.write(reinterpret_cast<const char*>(&t_cur), sizeof(double));
  fh.write(reinterpret_cast<const char*>(&IM_G), sizeof(uint32_t));
  fh.write(reinterpret_cast<const char*>(&JM_G), sizeof(uint32_t));
  fh.write(reinterpret_cast<const char*>(&KM_G), sizeof(uint32_t));
  const size_t n = (size_t)5 * IM_G * JM_G * KM_G;
  fh.write(reinterpret_cast<const char*>(Q), (std::streamsize)(n * sizeof(double)));
  return static_cast<bool>(fh);
}

bool read_checkpoint(const std::string& path, uint32_t IM_G, uint32_t JM_G, uint32_t KM_G,
                     double& t_cur, double* Q, std::string& err) {
  std::ifstream fh(path, std::ios::in | std::ios::binary);
  if (!fh) {
    err = "cannot open " + path;
    return false;
  }
  double t = 0.0;
  uint32_t d[3] = {0, 0, 0};
  fh.read(reinterpret_cast<char*>(&t), sizeof(double));
  fh.read(reinterpret_cast<char*>(d), 3 * sizeof(uint32_t));
  if (!fh) {
    err = "checkpoint " + path + " is truncated";
    return false;
  }
  if (d[0] != IM_G || d[1] != JM_G || d[2] != KM_G) {
    err = "The dimensions do not match.";
    return false;
  }
  const size_t n = (size_t)5 * IM_G * JM_G * KM_G;
  fh.read(reinterpret_cast<char*>(Q), (std::streamsize)(n * sizeof(double)));
  if (!fh) {
    err = "checkpoint " + path + " is truncated";
    return false;
  }
  t_cur = t;
  return true;
}

bool write_vtk(const std::string& path, int IM, int JM, int KM, const double* x, const double* y,
               const double* z, const double* rho, const double* u, const double* v,
               const double* w, const double* p) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  const int n = IM * JM * KM;
  std::fprintf(f, "# vtk DataFile Version 3.0\n");
  std::fprintf(f, "vtk output\n");
  std::fprintf(f, "ASCII\n");
  std::fprintf(f, "DATASET STRUCTURED_GRID\n");
  std::fprintf(f, "DIMENSIONS %d %d %d\n", IM, JM, KM);
  std::fprintf(f, "POINTS %d double\n", n);
  for (int q = 0; q < n; q++) std::fprintf(f, "%E %E %E\n", x[q], y[q], z[q]);
  const int NG = 2, RX = IM + 2 * NG, RY = JM + 2 * NG;
  auto field = [&](const char* name, const double* a) {
    std::fprintf(f, "%s 1 %d double\n", name, n);
    for (int k = 0; k < KM; k++)
      for (int j = 0; j < JM; j++)
        for (int i = 0; i < IM; i++)
          std::fprintf(f, "%E\n", a[(NG + i) + RX * ((NG + j) + (size_t)RY * (NG + k))]);
  };
  std::fprintf(f, "POINT_DATA %d\n", n);
  std::fprintf(f, "FIELD FieldData 5\n");
  field("rho", rho);
  field("u", u);
  field("v", v);
  field("w", w);
  field("p", p);
  return std::fclose(f) == 0;
}

}  // namespace cbns_io

// C entry points for tools and tests (host only, no GPU)
extern "C" {
int cbns_io_write_vtk(const char* path, int IM, int JM, int KM, const double* x, const double* y,
                      const double* z, const double* rho, const double* u, const double* v,
                      const double* w, const double* p) {
  return cbns_io::write_vtk(path, IM, JM, KM, x, y, z, rho, u, v, w, p) ? 0 : 1;
}
int cbns_io_write_plot3d(const char* path, int IM, int JM, int KM, const double* x,
                         const double* y, const double* z) {
  return cbns_io::write_plot3d(path, IM, JM, KM, x, y, z) ? 0 : 1;
}
int cbns_io_read_plot3d_dims(const char* path, int* dims) {
  std::vector<double> x, y, z;
  std::string err;
  return cbns_io::read_plot3d(path, dims[0], dims[1], dims[2], x, y, z, err) ? 0 : 1;
}
int cbns_io_read_plot3d(const char* path, double* x, double* y, double* z) {
  std::vector<double> a, b, c;
  std::string err;
  int IM, JM, KM;
  if (!cbns_io::read_plot3d(path, IM, JM, KM, a, b, c, err)) return 1;
  for (size_t q = 0; q < a.size(); q++) {
    x[q] = a[q];
    y[q] = b[q];
    z[q] = c[q];
  }
  return 0;
}
int cbns_io_write_checkpoint(const char* path, double t_cur, uint32_t IM_G, uint32_t JM_G,
                             uint32_t KM_G, const double* Q) {
  return cbns_io::write_checkpoint(path, t_cur, IM_G, JM_G, KM_G, Q) ? 0 : 1;
}
int cbns_io_read_checkpoint(const char* path, uint32_t IM_G, uint32_t JM_G, uint32_t KM_G,
                            double* t_cur, double* Q) {
  std::string err;
  return cbns_io::read_checkpoint(path, IM_G, JM_G, KM_G, *t_cur, Q, err) ? 0 : 1;
}
// parses an input file into 17 reference values + extras; returns 0 on success
int cbns_io_read_input(const char* path, double* num, int* ints, char* grid, char* ckpt, int len) {
  cbns_io::Input in;
  std::string err;
  if (!cbns_io::read_input(path, in, err)) return 1;
  const double nv[] = {in.CFL, in.t_cur, in.t_end, in.Pr, in.Pr_t, in.gamma,
                       in.Mach, in.Re, in.T_inf, in.angle_attack};
  for (int q = 0; q < 10; q++) num[q] = nv[q];
  for (int q = 0; q < 9; q++) num[10 + q] = in.period[q];
  const long iv[] = {in.nstep_max, in.checkpoint_freq, in.i_begin, in.i_end, in.is_restart,
                     in.viscous, in.visc_centered, in.bc[0], in.bc[1], in.bc[2]};
  for (int q = 0; q < 10; q++) ints[q] = (int)iv[q];
  std::snprintf(grid, (size_t)len, "%s", in.grid_fname.c_str());
  std::snprintf(ckpt, (size_t)len, "%s", in.checkpoint_fname.c_str());
  return 0;
}
}
