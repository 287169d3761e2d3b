// block3d.cpp -- the solver driver over libcbns: the flow of Block3d::solve
// (src/Block3d.cpp:204-313) with its input, mesh, checkpoint and VTK formats
// and its stdout lines, on one MI355X block through the C ABI.
//
//   cbns3d [input]      input = .ini (python/rae2822.ini keys, plus bc_*,
//                       period_*, viscous, visc_centered) or the positional
//                       input.txt (default: input.txt in the working directory)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/cbns.h"
#include "io.hpp"

namespace {

int fail(const std::string& msg) {
  std::cout << msg << std::endl;
  return 1;
}

void echo(const cbns_io::Input& in) {
  // src/Block3d.cpp:44-93
  std::cout << "grid_fname : " << in.grid_fname << std::endl;
  std::cout << "checkpoint_fname : " << in.checkpoint_fname << std::endl;
  std::cout << "CFL : " << in.CFL << std::endl;
  std::cout << "t_cur : " << in.t_cur << std::endl;
  std::cout << "t_end : " << in.t_end << std::endl;
  std::cout << "Pr : " << in.Pr << std::endl;
  std::cout << "Pr_t : " << in.Pr_t << std::endl;
  std::cout << "gamma : " << in.gamma << std::endl;
  std::cout << "Mach : " << in.Mach << std::endl;
  std::cout << "Re : " << in.Re << std::endl;
  std::cout << "T_inf : " << in.T_inf << std::endl;
  std::cout << "angle_attack : " << in.angle_attack << std::endl;
  std::cout << "nstep_max : " << in.nstep_max << std::endl;
  std::cout << "checkpoint_freq : " << in.checkpoint_freq << std::endl;
  std::cout << "i_begin : " << in.i_begin << std::endl;
  std::cout << "i_end : " << in.i_end << std::endl;
  std::cout << "is_restart : " << in.is_restart << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string path = argc > 1 ? argv[1] : "input.txt";
  cbns_io::Input in;
  std::string err;
  if (!cbns_io::read_input(path, in, err)) return fail("Exception opening/reading file\n" + err);
  echo(in);

  int IM = 0, JM = 0, KM = 0;
  std::vector<double> x, y, z;
  if (!cbns_io::read_plot3d(in.grid_fname, IM, JM, KM, x, y, z, err))
    return fail("Exception opening/reading file\n" + err);
  std::cout << IM << ", " << JM << ", " << KM << "\n";

  cbns_params p{};
  p.IM = IM; p.JM = JM; p.KM = KM;
  p.i_begin = in.i_begin; p.i_end = in.i_end;
  p.viscous = in.viscous; p.visc_centered = in.visc_centered;
  for (int a = 0; a < 3; a++) p.bc[a] = in.bc[a];
  p.CFL = in.CFL; p.gamma = in.gamma; p.Mach = in.Mach; p.Re = in.Re; p.Pr = in.Pr;
  p.Pr_t = in.Pr_t; p.T_inf = in.T_inf; p.angle_attack = in.angle_attack;

  cbns_block* blk = nullptr;
  int rc = cbns_create(&p, &blk);
  if (rc) return fail("cbns_create failed, code " + std::to_string(rc));
  auto chk = [&](int r, const char* what) {
    if (r) {
      std::cout << "Error: " << what << ": " << cbns_last_error(blk) << std::endl;
      cbns_destroy(blk);
      std::exit(1);
    }
  };
  const bool any_periodic = in.bc[0] || in.bc[1] || in.bc[2];
  chk(cbns_set_mesh(blk, x.data(), y.data(), z.data(), any_periodic ? in.period : nullptr),
      "calc_metrics");

  const uint32_t IM_G = IM + 4, JM_G = JM + 4, KM_G = KM + 4;
  std::vector<double> Q((size_t)5 * IM_G * JM_G * KM_G);
  double t_cur = in.t_cur;
  if (in.is_restart) {
    chk(cbns_initial_condition(blk), "initial_condition");
    chk(cbns_calc_conservative(blk), "calc_conservative");
  } else {
    if (!cbns_io::read_checkpoint(in.checkpoint_fname, IM_G, JM_G, KM_G, t_cur, Q.data(), err))
      return fail(err);
    chk(cbns_set_conservative(blk, Q.data()), "set_conservative");
    if (in.viscous) chk(cbns_calc_primitive(blk), "calc_primitive");
  }
  chk(cbns_set_bc_calc_primitive(blk, CBNS_BUF_Q), "set_bc_calc_primitive");

  auto finished = [&]() {
    // src/Block3d.cpp:476-488
    if (t_cur + 1.0e-8 > in.t_end) {
      std::cout << "Finished!\n";
      return true;
    }
    return false;
  };
  const auto start = std::chrono::high_resolution_clock::now();
  for (long n_t = 0; n_t < in.nstep_max; n_t++) {
    if (finished()) break;
    double dt = 0.0;
    chk(cbns_rk3_step(blk, in.t_end, &t_cur, &dt), "rk3 step");
    std::cout << n_t << ", " << dt << ", " << t_cur << std::endl;
    if (in.checkpoint_freq > 0 && 0 == (n_t + 1) % in.checkpoint_freq) {
      chk(cbns_get_conservative(blk, Q.data()), "get_conservative");
      if (!cbns_io::write_checkpoint(in.checkpoint_fname, t_cur, IM_G, JM_G, KM_G, Q.data()))
        std::cout << "Cannot open file!" << std::endl;
    }
  }
  const auto end = std::chrono::high_resolution_clock::now();
  std::cout << "Elapsed Time = " << std::chrono::duration<double>(end - start).count() << "\n";

  if (0 < in.nstep_max) {
    chk(cbns_get_conservative(blk, Q.data()), "get_conservative");
    if (!cbns_io::write_checkpoint(in.checkpoint_fname, t_cur, IM_G, JM_G, KM_G, Q.data()))
      std::cout << "Cannot open file!" << std::endl;
  }
  const size_t ng = (size_t)IM_G * JM_G * KM_G;
  std::vector<double> rho(ng), u(ng), v(ng), w(ng), pr(ng);
  chk(cbns_get_primitive(blk, rho.data(), u.data(), v.data(), w.data(), pr.data()),
      "get_primitive");
  if (cbns_io::write_vtk(in.vtk_fname, IM, JM, KM, x.data(), y.data(), z.data(), rho.data(),
                         u.data(), v.data(), w.data(), pr.data()))
    std::cout << "The output is successful!\n";
  cbns_destroy(blk);
  return 0;
}
