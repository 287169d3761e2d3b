// cbns_internal.hpp -- block state, kernel constants and launch helpers shared
// by the HIP translation units of libcbns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cbns.h"

namespace cbns {

constexpr int G = CBNS_G;
constexpr int NEQ = 5;

// Everything a kernel needs about the block, passed by value (no __constant__
// symbol: several blocks may live in one process).
struct Consts {
  int IM, JM, KM;           // nodes
  int PX, PY, PZ;           // ghosted extents
  int PXY;                  // PX*PY
  int S;                    // doubles per field
  int bc[3];
  int viscous, visc_centered;
  int i_begin, i_end;
  double CFL, gamma, gam1, gam1_inv, gM2, Mach2, Re_inv, Pr_inv, p_inf;
  double C_T_inf, C_dt_v;
  double qcoef;             // -Pr_inv * gam1_inv / Mach2 (src/viscous_terms_kernel.cu:98)
};

__host__ __device__ inline int idx(const Consts& c, int i, int j, int k) {
  return (i + G) + c.PX * ((j + G) + c.PY * (k + G));
}
__host__ __device__ inline int axis_n(const Consts& c, int a) {
  return a == 0 ? c.IM : (a == 1 ? c.JM : c.KM);
}
__host__ __device__ inline int stride(const Consts& c, int a) {
  return a == 0 ? 1 : (a == 1 ? c.PX : c.PXY);
}
__host__ __device__ inline bool periodic(const Consts& c, int a) {
  return c.bc[a] == CBNS_BC_PERIODIC;
}
// RK-updated node range along an axis (src/runge_kutta_kernel.cu:24-34 for
// reference axes; every node for periodic axes)
__host__ __device__ inline int upd_lo(const Consts& c, int a) { return periodic(c, a) ? 0 : 1; }
__host__ __device__ inline int upd_hi(const Consts& c, int a) {
  return periodic(c, a) ? axis_n(c, a) - 1 : axis_n(c, a) - 2;
}

// metric slots
enum { XI_X = 0, XI_Y, XI_Z, ETA_X, ETA_Y, ETA_Z, ZETA_X, ZETA_Y, ZETA_Z, JAC };
enum { P_RHO = 0, P_U, P_V, P_W, P_P, P_T };

// A 3-D box of nodes [lo, lo+n) flattened onto a 1-D grid.
struct Box {
  int lo[3];
  int n[3];
  __host__ __device__ long count() const { return (long)n[0] * n[1] * n[2]; }
};
__device__ inline bool box_node(const Box& b, long t, int& i, int& j, int& k) {
  if (t >= b.count()) return false;
  i = b.lo[0] + (int)(t % b.n[0]);
  long r = t / b.n[0];
  j = b.lo[1] + (int)(r % b.n[1]);
  k = b.lo[2] + (int)(r / b.n[1]);
  return true;
}
inline unsigned box_blocks(const Box& b, int threads) {
  return (unsigned)((b.count() + threads - 1) / threads);
}

}  // namespace cbns

// Per-kernel device timing with HIP events on the block's stream (only while
// profiling is switched on; see cbns_set_profiling).
struct cbns_prof {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> live[CBNS_K_COUNT];
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
};

struct cbns_block {
  cbns_prof prof;
  cbns_params p{};
  cbns::Consts c{};
  int device = 0;
  hipStream_t stream = nullptr;
  bool have_mesh = false;
  int form = CBNS_FORM_FP64;
  // device arrays
  double* met = nullptr;     // 10*S
  double* Q = nullptr;       // 5*S
  double* qp = nullptr;      // 5*S, the reference's Q_p
  double* qspare = nullptr;  // 5*S, ping-pong partner of qp
  double* prim = nullptr;    // 6*S
  double* mu = nullptr;      // S
  double* rhs = nullptr;     // 5*S, fused-stage accumulator
  double* flux = nullptr;    // 15*S (lazy, unfused path)
  double* dfv = nullptr;     // 4*S  (lazy)
  double* vtmp = nullptr;    // 12*S (lazy): viscous fluxes Ev,Fv,Gv
  double* red = nullptr;     // reduction scratch
  int red_n = 0;
  double* h_scalar = nullptr;  // pinned host scalars [4]
  std::string err;
};

namespace cbns {

// status helpers (defined in cbns_api.hip)
int fail(cbns_block* b, int code, const std::string& msg);
int hip_fail(cbns_block* b, hipError_t e, const char* what);

#define CBNS_HIP(b, call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return ::cbns::hip_fail((b), e_, #call); \
  } while (0)

// event bracket around one kernel launch (no-op unless profiling)
void prof_begin(cbns_block* b, int id, std::pair<hipEvent_t, hipEvent_t>& ev);
void prof_end(cbns_block* b, int id, std::pair<hipEvent_t, hipEvent_t>& ev);
struct ProfScope {
  cbns_block* b;
  int id;
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  ProfScope(cbns_block* b_, int id_) : b(b_), id(id_) { prof_begin(b, id, ev); }
  ~ProfScope() { prof_end(b, id, ev); }
};

// kernel launchers (one translation unit each)
int launch_metrics(cbns_block* b, const double* dx, const double* dy, const double* dz,
                   const double* period, int* bad_jac);
int launch_gradient_nodes(hipStream_t s, int IM, int JM, int KM, const double* f, double* fx,
                          double* fy, double* fz);
int launch_bc_primitive(cbns_block* b, double* Q);
int launch_primitive_all(cbns_block* b, const double* Q);
int launch_conservative_all(cbns_block* b, double* Q);
int launch_freestream(cbns_block* b);
int launch_numerical_flux(cbns_block* b, const double* Q);
int launch_viscous_unfused(cbns_block* b);
int launch_viscous(cbns_block* b, double* vf, double* mu, double* out, int mode);
int launch_update_rk3(cbns_block* b, double dt, int stage);
int launch_calc_dt(cbns_block* b, double* dt_host);
int launch_fused_stage(cbns_block* b, double dt, int stage, const double* qin, double* qout);
int launch_fill(hipStream_t s, double* p, long n, double v);

}  // namespace cbns
