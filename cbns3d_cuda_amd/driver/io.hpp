// io.hpp -- the reference's on-disk formats, host-side C++ (no HIP):
//   input parameters   .ini with the keys of python/rae2822.ini:1-26, or the
//                      positional input.txt that python/parser_ini.py writes
//                      (read order src/Block3d.cpp:44-92)
//   mesh               Plot3D ASCII single grid: "IM JM KM" then x, y, z with
//                      i fastest (src/Block3d.cpp:151-202)
//   checkpoint         binary: double t_cur, uint32 IM_G, JM_G, KM_G, then
//                      5*IM_G*JM_G*KM_G doubles of Q, component fastest, 2
//                      ghost layers, native endian (src/Block3d.cpp:411-474)
//   result             legacy-VTK ASCII STRUCTURED_GRID with rho,u,v,w,p
//                      (src/Block3d.cpp:357-409)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cbns_io {

struct Input {
  std::string grid_fname, checkpoint_fname;
  double CFL = 0.5, t_cur = 0.0, t_end = 0.0, Pr = 0.72, Pr_t = 0.9, gamma = 1.4;
  double Mach = 0.5, Re = 1.0e4, T_inf = 288.15, angle_attack = 0.0;
  long nstep_max = 0, checkpoint_freq = 1;
  int i_begin = 0, i_end = 0;
  bool is_restart = true;  // the reference's meaning: 1 = fresh start
  // keys this build adds (absent from the reference's files -> its defaults)
  int viscous = 1;         // the reference's IS_INVISCID build sets 0
  int visc_centered = 0;
  int bc[3] = {0, 0, 0};   // 0 = the reference's RAE2822 rule, 1 = periodic
  double period[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  std::string vtk_fname = "output.vtk";
};

// Parses an .ini file (sections ignored, "key = value") or, if the file has
// no '=', the positional input.txt.  Returns false and fills err on failure.
bool read_input(const std::string& path, Input& in, std::string& err);

bool read_plot3d(const std::string& path, int& IM, int& JM, int& KM, std::vector<double>& x,
                 std::vector<double>& y, std::vector<double>& z, std::string& err);
bool write_plot3d(const std::string& path, int IM, int JM, int KM, const double* x,
                  const double* y, const double* z);

bool write_checkpoint(const std::string& path, double t_cur, uint32_t IM_G, uint32_t JM_G,
                      uint32_t KM_G, const double* Q);
// dimensions must match (the reference exits on mismatch, src/Block3d.cpp:452-456)
bool read_checkpoint(const std::string& path, uint32_t IM_G, uint32_t JM_G, uint32_t KM_G,
                     double& t_cur, double* Q, std::string& err);

// prim fields are 2-ghost-layer arrays (IM+4)(JM+4)(KM+4) like get_primitive
bool write_vtk(const std::string& path, int IM, int JM, int KM, const double* x, const double* y,
               const double* z, const double* rho, const double* u, const double* v,
               const double* w, const double* p);

}  // namespace cbns_io
