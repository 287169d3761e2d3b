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

// Spectral-radius sum of a node, src/time_step_kernel.cu:43-94, from its
// primitives (T_l used only on viscous blocks) and its nine metric terms m
// (XI_X..ZETA_Z); one rounding per operation (no contraction) wherever it is
// inlined, so the time step matches the oracle bit for bit.
__device__ inline double node_lambda_v(const Consts& c, double rr, double uu, double vv, double ww,
                                       double pp, double T_l, const double* m) {
#pragma clang fp contract(off)
  const double cs = sqrt(c.gamma * pp / rr);
  double factor_v = 0.0;
  if (c.viscous) {
    const double mu = (1.0 + c.C_T_inf) / (T_l + c.C_T_inf) * sqrt(T_l * T_l * T_l);
    factor_v = 2.0 * mu * c.C_dt_v * sqrt(uu * uu + vv * vv + ww * ww) / cs / rr;
  }
  double sum = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double n_x = m[3 * d + 0], n_y = m[3 * d + 1], n_z = m[3 * d + 2];
    const double qn = uu * n_x + vv * n_y + ww * n_z;
    const double n_abs = sqrt(n_x * n_x + n_y * n_y + n_z * n_z);
    double tt = fabs(qn) + cs * n_abs;
    if (c.viscous) tt += n_abs * n_abs * factor_v;
    sum = d == 0 ? tt : sum + tt;
  }
  return sum;
}
// the same, read from node o of the primitive and metric fields
__device__ inline double node_lambda(const Consts& c, const double* __restrict__ prim,
                                     const double* __restrict__ met, int o) {
  const long S = c.S;
  double m[9];
#pragma unroll
  for (int q = 0; q < 9; q++) m[q] = met[q * S + o];
  return node_lambda_v(c, prim[o], prim[S + o], prim[2 * S + o], prim[3 * S + o], prim[4 * S + o],
                       c.viscous ? prim[5 * S + o] : 0.0, m);
}

// Maximum over a wavefront (every lane takes part) and over a workgroup of
// `threads` threads (tid: the linear thread index; the result is valid in
// wavefront 0); lds holds threads/64 doubles.
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double block_max(double v, double* lds, int tid, int threads) {
  v = wave_max(v);
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  if (wid == 0) {
    v = lane < threads / 64 ? lds[lane] : -INFINITY;
    v = wave_max(v);
  }
  return v;
}

// A 3-D box of nodes [lo, lo+n).
struct Box {
  int lo[3];
  int n[3];
  __host__ __device__ long count() const { return (long)n[0] * n[1] * n[2]; }
};
// Small boxes (boundary rules) run on a flattened 1-D grid; every count here
// is below 2^31 (field sizes are checked at cbns_create), so the decode uses
// 32-bit integer division.
__device__ inline bool box_node(const Box& b, long t, int& i, int& j, int& k) {
  if (t >= b.count()) return false;
  const unsigned u = (unsigned)t;
  const unsigned r = u / (unsigned)b.n[0];
  i = b.lo[0] + (int)(u - r * (unsigned)b.n[0]);
  const unsigned q = r / (unsigned)b.n[1];
  j = b.lo[1] + (int)(r - q * (unsigned)b.n[1]);
  k = b.lo[2] + (int)q;
  return true;
}
inline unsigned box_blocks(const Box& b, int threads) {
  return (unsigned)((b.count() + threads - 1) / threads);
}
// Large boxes run on a 3-D grid: 64 lanes along i, 4 rows along j, one
// plane per grid z, so no index division at all.
constexpr int BOX_X = 64, BOX_Y = 4;
inline dim3 box_grid(const Box& b) {
  return dim3((unsigned)((b.n[0] + BOX_X - 1) / BOX_X), (unsigned)((b.n[1] + BOX_Y - 1) / BOX_Y),
              (unsigned)b.n[2]);
}
inline dim3 box_block() { return dim3(BOX_X, BOX_Y, 1); }
__device__ inline bool box_node3(const Box& b, int& i, int& j, int& k) {
  const int ti = blockIdx.x * BOX_X + threadIdx.x, tj = blockIdx.y * BOX_Y + threadIdx.y;
  if (ti >= b.n[0] || tj >= b.n[1]) return false;
  i = b.lo[0] + ti;
  j = b.lo[1] + tj;
  k = b.lo[2] + (int)blockIdx.z;
  return true;
}

// The faces a stage's flux kernels compute along axis a (face m, between
// nodes m and m+1, stored at node m): -1..N-1 on a periodic axis, 0..N-2 on
// a reference axis, and the update range along the other two axes; `extra`
// widens the range along a by that many faces on either side.
inline Box face_box(const Consts& c, int a, int extra = 0) {
  Box box;
  for (int d = 0; d < 3; d++) {
    box.lo[d] = upd_lo(c, d);
    box.n[d] = upd_hi(c, d) - upd_lo(c, d) + 1;
  }
  const int n = axis_n(c, a);
  box.lo[a] = (periodic(c, a) ? -1 : 0) - extra;
  box.n[a] = n - 1 + (periodic(c, a) ? 2 : 0) + 2 * extra;
  return box;
}

// Tile march: 64 x 4 tiles of (i, j), each thread walking a chunk of
// `chunk` nodes along k.  Workgroups are handed to the 8 XCDs round-robin and
// every XCD has its own L2, so the flat block id is remapped to give each XCD
// one contiguous run of tiles (neighbouring rows and chunks then share that
// XCD's L2).  The grid is rounded up to a multiple of 8 blocks; blocks past
// the tile count return.
#ifndef CBNS_N_XCD
#define CBNS_N_XCD 1
#endif
constexpr int N_XCD = CBNS_N_XCD;
struct TileMarch {
  Box box;
  int ntx, nty, nkc, chunk, rows, xh;
};
// rows: workgroup rows along j (the workgroup is 64 x rows threads).
// xh > 0: each wavefront (one row) also holds the xh nodes on either side of
// its tile along i, so tiles are 64 - 2 xh nodes wide and the tile's
// xi-stencils read those neighbours by lane shuffles (tile_march_lane).
inline TileMarch tile_march(const Box& b, int chunk, int rows = BOX_Y, int xh = 0) {
  TileMarch t;
  t.box = b;
  t.rows = rows;
  t.xh = xh;
  t.ntx = (b.n[0] + BOX_X - 2 * xh - 1) / (BOX_X - 2 * xh);
  t.nty = (b.n[1] + rows - 1) / rows;
  t.nkc = (b.n[2] + chunk - 1) / chunk;
  t.chunk = chunk;
  return t;
}
inline unsigned tile_march_blocks(const TileMarch& t) {
  const unsigned nb = (unsigned)t.ntx * t.nty * t.nkc;
  return (nb + N_XCD - 1) / N_XCD * N_XCD;
}
__device__ inline bool tile_march_node(const TileMarch& t, int& i, int& j, int& k0, int& k1) {
  const unsigned nb = (unsigned)t.ntx * t.nty * t.nkc;
  const unsigned per = (nb + N_XCD - 1) / N_XCD;
  const unsigned lb = (blockIdx.x % N_XCD) * per + blockIdx.x / N_XCD;
  if (lb >= nb) return false;
  const unsigned tx = lb % t.ntx, r = lb / t.ntx;
  const unsigned ty = r % t.nty, kc = r / t.nty;
  const int ti = tx * BOX_X + threadIdx.x, tj = ty * t.rows + threadIdx.y;
  if (ti >= t.box.n[0] || tj >= t.box.n[1]) return false;
  i = t.box.lo[0] + ti;
  j = t.box.lo[1] + tj;
  k0 = t.box.lo[2] + (int)kc * t.chunk;
  k1 = min(k0 + t.chunk, t.box.lo[2] + t.box.n[2]);
  return true;
}
// The lane of a tile march with xi-halo xh: returns false (for the whole
// wavefront) past the box along j or past the tile count.  i is the lane's
// node, lo[0] - xh <= i; `has`: i < lo[0] + n[0] + xh (the node exists and
// its values are loaded); `own`: the node is in the tile and in the box.
__device__ inline bool tile_march_lane(const TileMarch& t, int& i, int& j, int& k0, int& k1,
                                       bool& has, bool& own) {
  const unsigned nb = (unsigned)t.ntx * t.nty * t.nkc;
  const unsigned per = (nb + N_XCD - 1) / N_XCD;
  const unsigned lb = (blockIdx.x % N_XCD) * per + blockIdx.x / N_XCD;
  if (lb >= nb) return false;
  const unsigned tx = lb % t.ntx, r = lb / t.ntx;
  const unsigned ty = r % t.nty, kc = r / t.nty;
  const int lane = (int)threadIdx.x;
  const int ti = (int)tx * (BOX_X - 2 * t.xh) - t.xh + lane, tj = ty * t.rows + threadIdx.y;
  if (tj >= t.box.n[1]) return false;
  has = ti < t.box.n[0] + t.xh;
  own = lane >= t.xh && lane < BOX_X - t.xh && ti < t.box.n[0];
  i = t.box.lo[0] + ti;
  j = t.box.lo[1] + tj;
  k0 = t.box.lo[2] + (int)kc * t.chunk;
  k1 = min(k0 + t.chunk, t.box.lo[2] + t.box.n[2]);
  return true;
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
  double* flux = nullptr;    // 15*S (lazy: reference-decomposed calls)
  double* dfv = nullptr;     // 4*S  (lazy)
  double* vtmp = nullptr;    // 12*S (lazy): viscous fluxes Ev,Fv,Gv
  double* red = nullptr;     // reduction scratch: [0..1] calc_dt, [2..3] folded dt, [4..] partials
  double* lam_part = nullptr;  // (lazy, viscous) spectral-radius maxima of the flux pass
  // cbns_rk3_step of a viscous block: stage 1's viscous flux pass also forms
  // the step's dt (limited to dt_cap) on the device, where the stages read it
  bool dt_fold = false;
  // inside cbns_rk3_step: stages 1-2 skip the primitives nothing reads
  bool step_internal = false;
  double dt_cap = 0.0;
  int red_n = 0;
  double* h_scalar = nullptr;  // pinned host scalars [4]
  // learned hybrid (cbns_set_indicator)
  bool have_indicator = false;
  bool flags_valid = false;
  bool ind_count_on = false;
  bool ind_per_stage = false;
  int ind_every = 1;             // steps between indicator passes
  int ind_age = 0;               // steps since the last pass
  float ind_thresh = 0.0f;
  float* ind_P = nullptr;        // network parameters (fp32)
  void* ind_pack = nullptr;      // bf16 operand tiles (3 x 256 x 8 bytes)
  unsigned* ind_count = nullptr; // troubled faces per axis [3]
  uint8_t* flags = nullptr;      // 3*S troubled-face bytes, face m at node m
  // hybrid: the troubled faces (a flagged face or a flagged neighbour along
  // its axis) listed per axis whenever the flags change; every stage forms
  // their characteristic-WENO fluxes in one pass (k_weno_list) and the
  // marches, which reconstruct linearly, take those from tflux
  int* tslot = nullptr;          // 3*S: slot of each listed face in tflux (face m at node m)
  int* tlist = nullptr;          // listed faces (node offsets); axis a from toff[a]
  double* tflux = nullptr;       // 5 x ttotal: WENO flux of each listed face
  unsigned* tcount = nullptr;    // [3] faces listed per axis (device)
  unsigned h_tcount[3] = {0, 0, 0};
  long tcap[3] = {0, 0, 0};      // list capacity per axis
  long toff[3] = {0, 0, 0};      // first slot of each axis
  long ttotal = 0;               // slots in all (tflux's field stride)
  bool tfull = false;            // a list overflowed: the marches run WENO at every face
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
void prof_begin(cbns_block* b, int id, std::pair<hipEvent_t, hipEvent_t>& ev, hipStream_t s);
void prof_end(cbns_block* b, int id, std::pair<hipEvent_t, hipEvent_t>& ev, hipStream_t s);
struct ProfScope {
  cbns_block* b;
  int id;
  hipStream_t s;
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  ProfScope(cbns_block* b_, int id_, hipStream_t s_ = nullptr);
  ~ProfScope() { prof_end(b, id, ev, s); }
};

// kernel launchers (one translation unit each)
int launch_metrics(cbns_block* b, const double* dx, const double* dy, const double* dz,
                   const double* period, int* bad_jac);
int launch_gradient_nodes(hipStream_t s, int IM, int JM, int KM, const double* f, double* fx,
                          double* fy, double* fz);
int launch_bc_primitive(cbns_block* b, double* Q);
// pmask: primitive fields whose ghost images the periodic wrap refreshes
// (bit n: field n of rho, u, v, w, p, T; -1: all of the block's)
int launch_bc_only(cbns_block* b, double* Q, int pmask = -1);
// Primitives a fused stage's output needs: all of them, except inside
// cbns_rk3_step, where the first two stages' outputs feed only the next
// stage (the marches derive what they need from Q; the viscous flux pass
// reads u, v, w, T) and the third stage writes all of them.
inline int stage_prim_mask(const cbns_block* b, int stage) {
  const int pall = b->c.viscous ? 0x3F : 0x1F;
  return b->step_internal && stage < 3 ? (b->c.viscous ? 0x2E : 0) : pall;
}
int launch_primitive_all(cbns_block* b, const double* Q);
int launch_conservative_all(cbns_block* b, double* Q);
int launch_freestream(cbns_block* b);
int launch_numerical_flux(cbns_block* b, const double* Q);
int launch_viscous_unfused(cbns_block* b);
// parts: 1 the flux pass, 2 the divergence pass, 3 both
int launch_viscous(cbns_block* b, double* vf, double* out, int mode, double* lam_part = nullptr,
                   int parts = 3);
int launch_viscous_f32(cbns_block* b, double* scratch, double* rhs, double* lam_part = nullptr,
                       int parts = 3);
long visc_flux_blocks(const Consts& c);
int launch_dt_from_partials(cbns_block* b, const double* partial, int n, double cap,
                            double* out);
int launch_update_rk3(cbns_block* b, double dt, int stage);
int launch_calc_dt(cbns_block* b, double* dt_host);
// workgroups (partial maxima) of the time-step pass, and the pass itself
// with dt (limited to cap) and the maximum left in out[0..1] on the device
long time_step_blocks(const Consts& c);
int launch_dt_device(cbns_block* b, double cap, double* out);
int launch_sutherland(cbns_block* b);
int launch_fused_stage(cbns_block* b, double dt, int stage, const double* qin, double* qout);
bool fused_stage_uses_face_flux();
int launch_fill(hipStream_t s, double* p, long n, double v);
long face_count(const cbns_block* b, int axis);
int launch_face_samples(cbns_block* b, int axis, const double* Q, float* feat, float* label);
int launch_indicator(cbns_block* b, const double* Q);
// after the flags change: list the troubled faces per axis and read the
// counts back (sets tfull when a list overflows its capacity)
int launch_face_lists(cbns_block* b);
int upload_indicator(cbns_block* b, const float* params);

}  // namespace cbns
