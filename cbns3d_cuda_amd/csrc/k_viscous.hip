// k_viscous.hip -- viscous terms in two passes instead of the reference's six
// (gradient_wg_kernel x4, viscous_terms_kernel, viscous_flux_kernel,
// viscous_contribution_kernel; src/calc_viscous_terms.cu:16-88 and
// src/calc_viscous_flux_contribution.cu:16-69):
//
//   k_visc_fluxes  per node of the viscous range [-2, N+1]^3: 4th-order
//                  derivatives of u,v,w,T along xi,eta,zeta, Sutherland mu,
//                  stress tau and heat flux q (from the Cartesian velocity
//                  and temperature gradients), contracted with J grad xi_d
//                  into the 12 values Ev,Fv,Gv (4 each).  The 12 gradients,
//                  6 stresses and 3 heat fluxes never reach HBM.
//   k_visc_div     4th-order central divergence of Ev,Fv,Gv at the update
//                  nodes, at the reference's transverse offset (-1) or
//                  centred; written to dfv (reference decomposition) or
//                  subtracted from a right-hand side.  The fused stage
//                  does without it: its marches subtract the fluxes' face
//                  values from their face fluxes (k_fused.hip, visc_face).
// Both walk along k in chunks (TileMarch): the five zeta-planes a 4th-order
// stencil needs (u,v,w,T in the flux pass, Gv in the divergence) stay in a
// register window, so each plane value is loaded once per chunk instead of
// five times (a plane of one field is 0.56 MB at 256^3: the +-2 planes of
// four fields do not fit a 4 MB XCD L2 and would be re-read from HBM).
// In the flux pass a wavefront holds its 60-node tile along i and two nodes
// on either side; the xi-derivatives take their neighbours from the adjacent
// lanes by wavefront shuffles (LDS crossbar) instead of 16 loads per node,
// and the eta-derivatives from a plane buffer in LDS that holds the
// workgroup's rows and two rows on either side.
// With its stencils loaded, the flux pass waited on the vector-memory
// pipeline (address and data units ~90 % busy at 256^3) rather than on HBM
// bytes; the shuffles and the plane buffer cut its loads per node from 46 to
// 18.  The divergence pass keeps its loads: there the shuffles cost 30
// VGPRs, two of its six waves per SIMD and 20 % of its time, and a plane
// buffer for Fv saved nothing (it is bound by its HBM bytes).
#include "cbns_internal.hpp"

namespace cbns {

namespace {

#ifndef VISC_CHUNK
#define VISC_CHUNK 8  // divergence-pass nodes along k per thread
#endif
#ifndef VISC_ROWS
#define VISC_ROWS 8  // divergence-pass workgroup rows along j (64 x VISC_ROWS threads)
#endif
constexpr int VROWS = VISC_ROWS;
#ifndef VISC_FLUX_ROWS
#define VISC_FLUX_ROWS 4  // flux-pass workgroup rows along j
#endif
#ifndef VISC_FLUX_CHUNK
#define VISC_FLUX_CHUNK 8  // flux-pass nodes along k per thread
#endif
constexpr int FROWS = VISC_FLUX_ROWS;

// 4th-order derivative along axis a at node m of the viscous range:
// one-sided at the two outermost layers of a reference axis
// (src/gradient_kernel.cu:162-253), central with wrapped ghosts otherwise.
__device__ __forceinline__ double dgrad(const Consts& c, int a, int m, const double* __restrict__ f,
                                        int st) {
  if (!periodic(c, a)) {
    const int p = m + 2, ne = axis_n(c, a) + 4;
    if (0 == p)
      return -25.0 / 12.0 * f[0] + 4.0 * f[st] - 3.0 * f[2 * st] + (4.0 / 3.0) * f[3 * st] -
             0.25 * f[4 * st];
    if (1 == p)
      return -5.0 / 6.0 * f[0] - 0.25 * f[-st] + 1.5 * f[st] - 0.5 * f[2 * st] +
             (1.0 / 12.0) * f[3 * st];
    if (ne - 2 == p)
      return (5.0 / 6.0) * f[0] - 1.0 / 12.0 * f[-3 * st] + 0.5 * f[-2 * st] - 1.5 * f[-st] +
             0.25 * f[st];
    if (ne - 1 == p)
      return (25.0 / 12.0) * f[0] + 0.25 * f[-4 * st] - 4.0 / 3.0 * f[-3 * st] +
             3.0 * f[-2 * st] - 4.0 * f[-st];
  }
  return (1.0 / 12.0) * f[-2 * st] - 2.0 / 3.0 * f[-st] + (2.0 / 3.0) * f[st] -
         1.0 / 12.0 * f[2 * st];
}

// xi-neighbours of a lane's value: n = v at i-2, i-1, i+1, i+2 (meaningful
// on the lanes that own their node; the shuffles wrap at the wave's ends)
template <class V>
__device__ __forceinline__ void xi_neighbours(V v, V n[4]) {
  const int lane = threadIdx.x;
  n[0] = __shfl(v, lane - 2);
  n[1] = __shfl(v, lane - 1);
  n[2] = __shfl(v, lane + 1);
  n[3] = __shfl(v, lane + 2);
}

// d/dxi of v at node i from the neighbouring lanes (every lane takes part);
// the one-sided layers of a reference axis read memory (dgrad, from f at the
// node) where `node` says the lane's row is in the array
__device__ __forceinline__ double dxi(const Consts& c, int i, double v,
                                      const double* __restrict__ f, bool node) {
  double nb[4];
  xi_neighbours(v, nb);
  const double d =
      (1.0 / 12.0) * nb[0] - 2.0 / 3.0 * nb[1] + (2.0 / 3.0) * nb[2] - 1.0 / 12.0 * nb[3];
  // (only the four outer nodes -2, -1, IM, IM+1 of the viscous range: the
  // halo lanes beyond them hold no node of the range)
  if (node && !periodic(c, 0) && (i == -2 || i == -1 || i == c.IM || i == c.IM + 1))
    return dgrad(c, 0, i, f, 1);
  return d;
}

// 4th-order central difference from a 5-plane window (same arithmetic as
// dgrad's central branch)
__device__ __forceinline__ double dwin(const double w[5]) {
  return (1.0 / 12.0) * w[0] - 2.0 / 3.0 * w[1] + (2.0 / 3.0) * w[3] - 1.0 / 12.0 * w[4];
}

// The node's three viscous flux vectors E[d] = J grad xi_d . [tau; u.tau - q]
// (src/viscous_terms_kernel.cu:59-105, src/viscous_flux_kernel.cu:57-110),
// from the Cartesian gradients
//   g[a][b] = d u_a / d x_b = sum_d (d u_a / d xi_d) (d xi_d / d x_b)
// of u, v, w (a = 0..2) and T (a = 3):
//   tau = mu/Re (g + g^T - 2/3 tr(g) I),  q = -mu/(Re Pr (gamma-1) M^2) grad T.
// du[a][d]: d u_a / d xi_d; uvw: the node's velocity; T_l: its temperature;
// m: its metrics (XI_X..JAC).  T = double (plain form) or float (precision
// form: the index derivatives, metrics and their products with J are rounded
// to float, mu is formed in double).  Returns mu (unscaled).
template <class T>
__device__ __forceinline__ double visc_node_fluxes(const Consts& c, const double du[4][3],
                                                   const double uvw[3], double T_l,
                                                   const double m[10], T E[3][4]) {
  T g[4][3];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const T d0 = (T)du[a][0], d1 = (T)du[a][1], d2 = (T)du[a][2];
#pragma unroll
    for (int b = 0; b < 3; b++) g[a][b] = d0 * (T)m[b] + d1 * (T)m[3 + b] + d2 * (T)m[6 + b];
  }
  const double mu_d = (1.0 + c.C_T_inf) / (T_l + c.C_T_inf) * sqrt(T_l * T_l * T_l);
  const T mr = (T)(mu_d * c.Re_inv);
  const T third2 = T(2.0 / 3.0) * (g[0][0] + g[1][1] + g[2][2]);
  T tau[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    tau[a][a] = mr * (T(2) * g[a][a] - third2);
#pragma unroll
    for (int b = a + 1; b < 3; b++) tau[a][b] = tau[b][a] = mr * (g[a][b] + g[b][a]);
  }
  const T mq = mr * (T)c.qcoef;
  T e[3];  // u . tau - q
#pragma unroll
  for (int b = 0; b < 3; b++)
    e[b] = (T)uvw[0] * tau[0][b] + (T)uvw[1] * tau[1][b] + (T)uvw[2] * tau[2][b] - mq * g[3][b];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const T nx = (T)(m[3 * d + 0] * m[JAC]), ny = (T)(m[3 * d + 1] * m[JAC]);
    const T nz = (T)(m[3 * d + 2] * m[JAC]);
#pragma unroll
    for (int a = 0; a < 3; a++) E[d][a] = nx * tau[0][a] + ny * tau[1][a] + nz * tau[2][a];
    E[d][3] = nx * e[0] + ny * e[1] + nz * e[2];
  }
  return mu_d;
}

// Fluxes of node (i, j, k) stored to vf (Ev, Fv, Gv: 4 fields each).
// win[f][s]: field f (u,v,w,T) at k-2+s; dx[f], de[f]: d/dxi, d/deta of
// field f; m: the node's metrics.
template <class T>
__device__ __forceinline__ void visc_fluxes_node(const Consts& c, int i, int j, int k,
                                                 const double win[4][5],
                                                 const double dx[4], const double de[4],
                                                 const double m[10],
                                                 const double* __restrict__ prim,
                                                 T* __restrict__ vf) {
  const long S = c.S;
  const int o = idx(c, i, j, k);
  const int sz = c.PXY;
  // zeta: the register window, except at the one-sided outer layers of a
  // reference axis
  const bool zc = periodic(c, 2) || (k >= 0 && k <= c.KM - 1);
  const int pf[4] = {P_U, P_V, P_W, P_T};
  double du[4][3];
#pragma unroll
  for (int f = 0; f < 4; f++) {
    du[f][0] = dx[f];
    du[f][1] = de[f];
    du[f][2] = zc ? dwin(win[f]) : dgrad(c, 2, k, prim + pf[f] * S + o, sz);
  }
  const double uvw[3] = {win[0][2], win[1][2], win[2][2]};
  T E[3][4];
  visc_node_fluxes<T>(c, du, uvw, win[3][2], m, E);
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int e = 0; e < 4; e++) vf[(long)(4 * d + e) * S + o] = E[d][e];
}

// The workgroup's rows (64 x FROWS lanes, one row of j per wavefront) and
// the two rows on either side along j: at each plane k the lanes put their
// u,v,w,T into a plane buffer in LDS (the halo rows loaded by the first four
// wavefronts), and the eta-derivatives are differenced from it instead of
// from four loads per field.  Two buffers alternate with k, so one barrier
// per plane orders the writes of plane k+1 after the reads of plane k-1.
constexpr int VPLANE_ROWS = FROWS + 4;
static_assert(FROWS >= 4, "the halo rows are loaded by the first four wavefronts");

// Where a workgroup's lane of the flux pass sits: like tile_march_lane, but
// rows past the box along j stay (they load the values the last rows'
// stencils read and take part in the barriers).  false for the whole
// workgroup past the tile count.
__device__ inline bool visc_flux_lane(const TileMarch& t, int& i, int& j, int& k0, int& k1,
                                      bool& has, bool& own, bool& row_in) {
  const unsigned nb = (unsigned)t.ntx * t.nty * t.nkc;
  const unsigned per = (nb + N_XCD - 1) / N_XCD;
  const unsigned lb = (blockIdx.x % N_XCD) * per + blockIdx.x / N_XCD;
  if (lb >= nb) return false;
  const unsigned tx = lb % t.ntx, r = lb / t.ntx;
  const unsigned ty = r % t.nty, kc = r / t.nty;
  const int lane = (int)threadIdx.x;
  const int ti = (int)tx * (BOX_X - 2 * t.xh) - t.xh + lane, tj = ty * t.rows + threadIdx.y;
  row_in = tj < t.box.n[1];
  has = ti < t.box.n[0] + t.xh;
  own = row_in && lane >= t.xh && lane < BOX_X - t.xh && ti < t.box.n[0];
  i = t.box.lo[0] + ti;
  j = t.box.lo[1] + tj;
  k0 = t.box.lo[2] + (int)kc * t.chunk;
  k1 = min(k0 + t.chunk, t.box.lo[2] + t.box.n[2]);
  return true;
}

// LAM: also the spectral-radius maximum of the workgroup's nodes in
// [0, IM) x [0, JM) x [0, KM) (node_lambda, the time step's per-node term),
// written to lam_part[workgroup]; NaN if any node's is NaN.
template <class T, bool LAM>
__global__ void __launch_bounds__(64 * FROWS) k_visc_fluxes(Consts c, TileMarch tm,
                                                     const double* __restrict__ prim,
                                                     const double* __restrict__ met,
                                                     T* __restrict__ vf,
                                                     double* __restrict__ lam_part) {
  __shared__ double plane[2][4][VPLANE_ROWS][BOX_X];
  int i = 0, j = 0, k0 = 0, k1 = 0;
  bool has = false, own = false, row_in = false;
  if (!visc_flux_lane(tm, i, j, k0, k1, has, own, row_in)) return;
  const int lane = threadIdx.x, ty = threadIdx.y;
  // the lane's row is loaded if it lies in the array (j >= -2 > -G always)
  const bool has_row = has && j < c.JM + G;
  // halo row of this wavefront (the first four): j0-2, j0-1, j0+FROWS, j0+FROWS+1
  const int hr = ty < 2 ? ty : ty + FROWS;
  const int jh = j - ty - 2 + hr;
  const bool halo = ty < 4, has_halo = halo && has && jh < c.JM + G;
  // eta-stencil from the plane buffer, except at the one-sided outer layers
  // of a reference axis
  const bool ec = periodic(c, 1) || (j >= 0 && j <= c.JM - 1);
  double lam = -INFINITY;
  int nan_seen = 0;
  const bool lam_line = LAM && own && i >= 0 && i < c.IM && j >= 0 && j < c.JM;
  const long S = c.S;
  const int sz = c.PXY;
  const double* f0[4] = {prim + P_U * S, prim + P_V * S, prim + P_W * S, prim + P_T * S};
  const int base = idx(c, i, j, 0), hbase = idx(c, i, jh, 0);
  double win[4][5];
#pragma unroll
  for (int f = 0; f < 4; f++)
#pragma unroll
    for (int s = 0; s < 4; s++) win[f][s + 1] = has_row ? f0[f][base + (k0 - 2 + s) * sz] : 0.0;
  // the halo rows are fetched one plane ahead (plane k1 <= KM + 2 lies in
  // the array), so the barrier does not wait on a load of its own iteration
  // (the metrics too cost a wavefront per SIMD and ran slower)
  double hn[4];
#pragma unroll
  for (int f = 0; f < 4; f++) hn[f] = has_halo ? f0[f][hbase + k0 * sz] : 0.0;
  for (int k = k0; k < k1; k++) {
    double dx[4], de[4], hv[4];
    const int pb = k & 1;
#pragma unroll
    for (int f = 0; f < 4; f++) {
      hv[f] = hn[f];
      hn[f] = has_halo ? f0[f][hbase + (k + 1) * sz] : 0.0;
    }
#pragma unroll
    for (int f = 0; f < 4; f++) {
#pragma unroll
      for (int s = 0; s < 4; s++) win[f][s] = win[f][s + 1];
      win[f][4] = has_row ? f0[f][base + (k + 2) * sz] : 0.0;
      dx[f] = dxi(c, i, win[f][2], f0[f] + base + k * sz, has_row);
      plane[pb][f][ty + 2][lane] = win[f][2];
      if (halo) plane[pb][f][hr][lane] = hv[f];
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 4; f++)
      de[f] = ec ? (1.0 / 12.0) * plane[pb][f][ty][lane] - 2.0 / 3.0 * plane[pb][f][ty + 1][lane] +
                       (2.0 / 3.0) * plane[pb][f][ty + 3][lane] -
                       1.0 / 12.0 * plane[pb][f][ty + 4][lane]
                 : (own ? dgrad(c, 1, j, f0[f] + base + k * sz, c.PX) : 0.0);
    if (own) {
      double m[10];
#pragma unroll
      for (int q = 0; q < 10; q++) m[q] = met[q * S + base + k * sz];
      visc_fluxes_node<T>(c, i, j, k, win, dx, de, m, prim, vf);
      if (LAM && lam_line && k >= 0 && k < c.KM) {
        // u, v, w, T and the metrics from the registers, rho and p loaded
        const int o = base + k * sz;
        const double sl = node_lambda_v(c, prim[P_RHO * S + o], win[0][2], win[1][2], win[2][2],
                                        prim[P_P * S + o], win[3][2], m);
        if (isnan(sl)) nan_seen = 1;
        else lam = fmax(lam, sl);
      }
    }
  }
  if (LAM) {
    __shared__ double lds[FROWS];
    const int tid = threadIdx.x + threadIdx.y * BOX_X;
    lam = block_max(lam, lds, tid, 64 * FROWS);
    __syncthreads();
    const int any_nan = __syncthreads_or(nan_seen);
    if (tid == 0) lam_part[blockIdx.x] = any_nan ? NAN : lam;
  }
}

// src/viscous_contribution_kernel.cu:35-67.  mode 0: dfv[e] = div; mode 1:
// out[(e+1)] -= div (the fused stage's right-hand side, momentum and energy).
// gw[e][s]: Gv component e on the zeta line of the stencil at k-2+s.
template <class T>
__device__ __forceinline__ void visc_div_node(const Consts& c, int i, int j, int k,
                                              const T gw[4][5], const T* __restrict__ vf,
                                              double* __restrict__ out, int mode) {
  const long S = c.S;
  const int sh = c.visc_centered ? 0 : -1;
  const T c0 = T(1.0 / 12.0), c1 = T(-2.0 / 3.0);
  const int ox = idx(c, i, j + sh, k + sh), oy = idx(c, i + sh, j, k + sh);
  const int sx = 1, sy = c.PX;
  const int o = idx(c, i, j, k);
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const T* Ev = vf + (long)e * S;
    const T* Fv = vf + (long)(4 + e) * S;
    T df = c0 * Ev[ox - 2 * sx] + c1 * Ev[ox - sx] - c1 * Ev[ox + sx] - c0 * Ev[ox + 2 * sx];
    df += c0 * Fv[oy - 2 * sy] + c1 * Fv[oy - sy] - c1 * Fv[oy + sy] - c0 * Fv[oy + 2 * sy];
    df += c0 * gw[e][0] + c1 * gw[e][1] - c1 * gw[e][3] - c0 * gw[e][4];
    if (mode == 0) out[e * S + o] = (double)df;
    else out[(e + 1) * S + o] -= (double)df;
  }
}

template <class T>
__global__ void __launch_bounds__(64 * VROWS) k_visc_div(Consts c, TileMarch tm,
                                                  const T* __restrict__ vf,
                                                  double* __restrict__ out, int mode) {
  int i, j, k0, k1;
  if (!tile_march_node(tm, i, j, k0, k1)) return;
  const long S = c.S;
  const int sh = c.visc_centered ? 0 : -1;
  const int sz = c.PXY;
  const int base = idx(c, i + sh, j + sh, 0);  // the zeta line of the Gv stencil
  const T* G0 = vf + 8 * S;
  T gw[4][5];
#pragma unroll
  for (int e = 0; e < 4; e++)
#pragma unroll
    for (int s = 0; s < 4; s++) gw[e][s + 1] = G0[e * S + base + (k0 - 2 + s) * sz];
  for (int k = k0; k < k1; k++) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
#pragma unroll
      for (int s = 0; s < 4; s++) gw[e][s] = gw[e][s + 1];
      gw[e][4] = G0[e * S + base + (k + 2) * sz];
    }
    visc_div_node<T>(c, i, j, k, gw, vf, out, mode);
  }
}

}  // namespace

// Workgroups of the viscous flux pass (the entries lam_part needs).
static TileMarch visc_flux_tiles(const Consts& c) {
  Box vb;
  for (int a = 0; a < 3; a++) {
    vb.lo[a] = -2;
    vb.n[a] = axis_n(c, a) + 4;
  }
  return tile_march(vb, VISC_FLUX_CHUNK, FROWS, 2);
}

long visc_flux_blocks(const Consts& c) { return (long)tile_march_blocks(visc_flux_tiles(c)); }

// vf: 12*S scratch (of T).  mode 0 writes dfv (4*S) to
// `out`, mode 1 subtracts the divergence from the right-hand side (5*S).
// lam_part (visc_flux_blocks entries) or null: the flux pass also forms the
// time step's per-workgroup spectral-radius maxima.
template <class T>
int viscous(cbns_block* b, T* vf, double* out, int mode, double* lam_part, int parts) {
  const Consts& c = b->c;
  if (parts & 1) {
    ProfScope ps(b, CBNS_K_VISC_FLUX);
    const TileMarch tm = visc_flux_tiles(c);
    const dim3 grid(tile_march_blocks(tm)), block(BOX_X, FROWS);
    if (lam_part)
      k_visc_fluxes<T, true><<<grid, block, 0, b->stream>>>(c, tm, b->prim, b->met, vf,
                                                            lam_part);
    else
      k_visc_fluxes<T, false><<<grid, block, 0, b->stream>>>(c, tm, b->prim, b->met, vf,
                                                             nullptr);
    CBNS_HIP(b, hipGetLastError());
  }
  if (!(parts & 2)) return CBNS_OK;
  Box ub;
  for (int a = 0; a < 3; a++) {
    ub.lo[a] = upd_lo(c, a);
    ub.n[a] = upd_hi(c, a) - upd_lo(c, a) + 1;
  }
  ProfScope ps(b, CBNS_K_VISC_DIV);
  const TileMarch tm = tile_march(ub, VISC_CHUNK, VROWS);
  k_visc_div<T><<<tile_march_blocks(tm), dim3(BOX_X, VROWS), 0, b->stream>>>(c, tm, vf, out, mode);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

int launch_viscous(cbns_block* b, double* vf, double* out, int mode, double* lam_part, int parts) {
  return viscous<double>(b, vf, out, mode, lam_part, parts);
}

// precision form: the same scratch holds 12*S floats
int launch_viscous_f32(cbns_block* b, double* scratch, double* rhs, double* lam_part, int parts) {
  return viscous<float>(b, reinterpret_cast<float*>(scratch), rhs, 1, lam_part, parts);
}

}  // namespace cbns
