// k_unfused.hip -- the reference's launcher decomposition, one entry point per
// block3d_cuda call: calc_numerical_flux (face fluxes to HBM),
// calc_viscous_terms + calc_viscous_flux_contribution (intermediates to HBM)
// and update_rk3.  The fused sweep kernels (k_fused.hip) replace this chain in
// cbns_rk3_stage; this path exists so callers of the reference API can swap
// it in call for call, and it cross-checks the fused path.
#include "cbns_internal.hpp"
#include "face_math.hpp"

namespace cbns {

namespace {

// src/rec_riemann_kernel.cu:15-247: one thread per face; the face between
// node m and m+1 along `d` is stored at the index of node m.
__global__ void __launch_bounds__(256) k_face_flux(Consts c, Box box, int d,
                                                   const double* __restrict__ Q,
                                                   const double* __restrict__ prim,
                                                   const double* __restrict__ met,
                                                   double* __restrict__ flux) {
  int i, j, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const long S = c.S;
  const int st = stride(c, d);
  const int a = idx(c, i, j, k), bb = a + st;
  const double Ja = met[JAC * S + a], Jb = met[JAC * S + bb];
  const double nx = 0.5 * (met[(3 * d + 0) * S + a] * Ja + met[(3 * d + 0) * S + bb] * Jb);
  const double ny = 0.5 * (met[(3 * d + 1) * S + a] * Ja + met[(3 * d + 1) * S + bb] * Jb);
  const double nz = 0.5 * (met[(3 * d + 2) * S + a] * Ja + met[(3 * d + 2) * S + bb] * Jb);
  double q[6][5], pl[5], pr[5], f[5];
#pragma unroll
  for (int s = 0; s < 6; s++)
#pragma unroll
    for (int n = 0; n < 5; n++) q[s][n] = Q[n * S + a + (s - 2) * st];
#pragma unroll
  for (int n = 0; n < 5; n++) {
    pl[n] = prim[n * S + a];
    pr[n] = prim[n * S + bb];
  }
  const GasT<double> g{c.gamma, c.gam1, c.gam1_inv};
  face_flux<double>(g, q, pl, pr, nx, ny, nz, f);
#pragma unroll
  for (int n = 0; n < 5; n++) flux[(d * 5 + n) * S + a] = f[n];
}

// src/runge_kutta_kernel.cu:13-80
__global__ void k_update_rk3(Consts c, Box box, double dt, int stage, const double* __restrict__ met,
                             const double* __restrict__ flux, const double* __restrict__ dfv,
                             double* __restrict__ Q, double* __restrict__ Qp) {
  int i, j, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const long S = c.S;
  const int o = idx(c, i, j, k);
  const double Jac_a = met[JAC * S + o];
  for (int n = 0; n < NEQ; n++) {
    const double* Ep = flux + (0 * 5 + n) * S;
    const double* Fp = flux + (1 * 5 + n) * S;
    const double* Gp = flux + (2 * 5 + n) * S;
    double df = Ep[o] - Ep[o - 1];
    df += Fp[o] - Fp[o - c.PX];
    df += Gp[o] - Gp[o - c.PXY];
    if (c.viscous && n > 0) df -= dfv[(n - 1) * S + o];
    df /= Jac_a;
    const long qi = n * S + o;
    if (1 == stage) Qp[qi] = Q[qi] - dt * df;
    else if (2 == stage) Qp[qi] = 0.75 * Q[qi] + 0.25 * (Qp[qi] - dt * df);
    else if (3 == stage) Q[qi] = (Q[qi] + 2.0 * (Qp[qi] - dt * df)) / 3.0;
  }
}

// 4th-order derivative along one axis on the viscous node range [-2, N+1]:
// one-sided at the two outermost layers of a reference axis
// (src/gradient_kernel.cu:135-255), central with wrapped ghosts otherwise.
__device__ inline double dgrad(const Consts& c, int a, int m, const double* f, int st) {
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

// gradient_wg_kernel x4 (src/calc_viscous_terms.cu:27-36): vt[(f*3+a)*S]
__global__ void k_visc_grad(Consts c, Box box, const double* __restrict__ prim,
                            double* __restrict__ vt) {
  int i, j, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const long S = c.S;
  const int o = idx(c, i, j, k);
  const int pos[3] = {i, j, k};
  for (int f = 0; f < 4; f++) {
    const double* F = prim + (long)(f == 3 ? P_T : P_U + f) * S + o;
    for (int a = 0; a < 3; a++) vt[(f * 3 + a) * S + o] = dgrad(c, a, pos[a], F, stride(c, a));
  }
}

// viscous_terms_kernel, src/viscous_terms_kernel.cu:59-105.  tau slots
// xx,yy,zz,xy,xz,yz at vt[12..17], q at vt[18..20]; mu (unscaled) to mu[].
__global__ void k_visc_terms(Consts c, Box box, const double* __restrict__ prim,
                             const double* __restrict__ met, double* __restrict__ mu,
                             double* __restrict__ vt) {
  int i, j, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const long S = c.S;
  const int o = idx(c, i, j, k);
  const double T_l = prim[P_T * S + o];
  const double C_T_inf = c.C_T_inf;
  double mu_l = (1.0 + C_T_inf) / (T_l + C_T_inf) * sqrt(T_l * T_l * T_l);
  mu[o] = mu_l;
  const double xi_x_l = met[XI_X * S + o], xi_y_l = met[XI_Y * S + o], xi_z_l = met[XI_Z * S + o];
  const double eta_x_l = met[ETA_X * S + o], eta_y_l = met[ETA_Y * S + o],
               eta_z_l = met[ETA_Z * S + o];
  const double zeta_x_l = met[ZETA_X * S + o], zeta_y_l = met[ZETA_Y * S + o],
               zeta_z_l = met[ZETA_Z * S + o];
  double u_xi_l = vt[0 * S + o], u_eta_l = vt[1 * S + o], u_zeta_l = vt[2 * S + o];
  const double v_xi_l = vt[3 * S + o], v_eta_l = vt[4 * S + o], v_zeta_l = vt[5 * S + o];
  const double w_xi_l = vt[6 * S + o], w_eta_l = vt[7 * S + o], w_zeta_l = vt[8 * S + o];
  mu_l *= c.Re_inv;
  vt[12 * S + o] = (2.0/3.0)*mu_l*(2*eta_x_l*u_eta_l - eta_y_l*v_eta_l - eta_z_l*w_eta_l + 2*u_xi_l*xi_x_l + 2*u_zeta_l*zeta_x_l - v_xi_l*xi_y_l - v_zeta_l*zeta_y_l - w_xi_l*xi_z_l - w_zeta_l*zeta_z_l);
  vt[15 * S + o] = mu_l*(eta_x_l*v_eta_l + eta_y_l*u_eta_l + u_xi_l*xi_y_l + u_zeta_l*zeta_y_l + v_xi_l*xi_x_l + v_zeta_l*zeta_x_l);
  vt[13 * S + o] = (2.0/3.0)*mu_l*(-eta_x_l*u_eta_l + 2*eta_y_l*v_eta_l - eta_z_l*w_eta_l - u_xi_l*xi_x_l - u_zeta_l*zeta_x_l + 2*v_xi_l*xi_y_l + 2*v_zeta_l*zeta_y_l - w_xi_l*xi_z_l - w_zeta_l*zeta_z_l);
  vt[16 * S + o] = mu_l*(eta_x_l*w_eta_l + eta_z_l*u_eta_l + u_xi_l*xi_z_l + u_zeta_l*zeta_z_l + w_xi_l*xi_x_l + w_zeta_l*zeta_x_l);
  vt[17 * S + o] = mu_l*(eta_y_l*w_eta_l + eta_z_l*v_eta_l + v_xi_l*xi_z_l + v_zeta_l*zeta_z_l + w_xi_l*xi_y_l + w_zeta_l*zeta_y_l);
  vt[14 * S + o] = (2.0/3.0)*mu_l*(-eta_x_l*u_eta_l - eta_y_l*v_eta_l + 2*eta_z_l*w_eta_l - u_xi_l*xi_x_l - u_zeta_l*zeta_x_l - v_xi_l*xi_y_l - v_zeta_l*zeta_y_l + 2*w_xi_l*xi_z_l + 2*w_zeta_l*zeta_z_l);
  mu_l *= c.qcoef;
  u_xi_l = vt[9 * S + o];
  u_eta_l = vt[10 * S + o];
  u_zeta_l = vt[11 * S + o];
  vt[18 * S + o] = mu_l * (xi_x_l * u_xi_l + eta_x_l * u_eta_l + zeta_x_l * u_zeta_l);
  vt[19 * S + o] = mu_l * (xi_y_l * u_xi_l + eta_y_l * u_eta_l + zeta_y_l * u_zeta_l);
  vt[20 * S + o] = mu_l * (xi_z_l * u_xi_l + eta_z_l * u_eta_l + zeta_z_l * u_zeta_l);
}

// viscous_flux_kernel, src/viscous_flux_kernel.cu:54-110: Ev,Fv,Gv at vt[21..32]
__global__ void k_visc_flux(Consts c, Box box, const double* __restrict__ prim,
                            const double* __restrict__ met, double* __restrict__ vt) {
  int i, j, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const long S = c.S;
  const int o = idx(c, i, j, k);
  const double Jac_a = met[JAC * S + o];
  const double uu = prim[P_U * S + o], vv = prim[P_V * S + o], ww = prim[P_W * S + o];
  const double txx = vt[12 * S + o], tyy = vt[13 * S + o], tzz = vt[14 * S + o];
  const double txy = vt[15 * S + o], txz = vt[16 * S + o], tyz = vt[17 * S + o];
  const double qx = vt[18 * S + o], qy = vt[19 * S + o], qz = vt[20 * S + o];
  for (int d = 0; d < 3; d++) {
    const double n_x = met[(3 * d + 0) * S + o] * Jac_a;
    const double n_y = met[(3 * d + 1) * S + o] * Jac_a;
    const double n_z = met[(3 * d + 2) * S + o] * Jac_a;
    double* E = vt + (21 + 4 * d) * S + o;
    E[0] = n_x * txx + n_y * txy + n_z * txz;
    E[S] = n_x * txy + n_y * tyy + n_z * tyz;
    E[2 * S] = n_x * txz + n_y * tyz + n_z * tzz;
    E[3 * S] = n_x * (uu * txx + vv * txy + ww * txz - qx) + n_y * (uu * txy + vv * tyy + ww * tyz - qy) +
               n_z * (uu * txz + vv * tyz + ww * tzz - qz);
  }
}

// viscous_contribution_kernel, src/viscous_contribution_kernel.cu:35-67,
// with the reference's transverse offset -1 (visc_centered = 0) or 0.
__global__ void k_visc_contrib(Consts c, Box box, const double* __restrict__ vt,
                               double* __restrict__ dfv) {
  int i, j, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const long S = c.S;
  const int sh = c.visc_centered ? 0 : -1;
  const double c0 = 1.0 / 12.0, c1 = -2.0 / 3.0;
  const int ox = idx(c, i, j + sh, k + sh), oy = idx(c, i + sh, j, k + sh),
            oz = idx(c, i + sh, j + sh, k);
  const int sx = 1, sy = c.PX, sz = c.PXY;
  for (int e = 0; e < 4; e++) {
    const double* Ev = vt + (21 + e) * S;
    const double* Fv = vt + (25 + e) * S;
    const double* Gv = vt + (29 + e) * S;
    double df = c0 * Ev[ox - 2 * sx] + c1 * Ev[ox - sx] - c1 * Ev[ox + sx] - c0 * Ev[ox + 2 * sx];
    df += c0 * Fv[oy - 2 * sy] + c1 * Fv[oy - sy] - c1 * Fv[oy + sy] - c0 * Fv[oy + 2 * sy];
    df += c0 * Gv[oz - 2 * sz] + c1 * Gv[oz - sz] - c1 * Gv[oz + sz] - c0 * Gv[oz + 2 * sz];
    dfv[e * S + idx(c, i, j, k)] = df;
  }
}

Box update_box(const Consts& c) {
  Box b;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = upd_lo(c, a);
    b.n[a] = upd_hi(c, a) - upd_lo(c, a) + 1;
  }
  return b;
}

Box visc_box(const Consts& c) {
  Box b;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = -2;
    b.n[a] = axis_n(c, a) + 4;
  }
  return b;
}

}  // namespace

int launch_numerical_flux(cbns_block* b, const double* Q) {
  const Consts& c = b->c;
  for (int d = 0; d < 3; d++) {
    Box box = update_box(c);
    box.lo[d] = periodic(c, d) ? -1 : 0;
    box.n[d] = axis_n(c, d) - 1 + (periodic(c, d) ? 2 : 0);  // faces -1..N-1 or 0..N-2
    k_face_flux<<<box_blocks(box, 256), 256, 0, b->stream>>>(c, box, d, Q, b->prim, b->met,
                                                             b->flux);
    CBNS_HIP(b, hipGetLastError());
  }
  return CBNS_OK;
}

int launch_viscous_unfused(cbns_block* b) {
  const Consts& c = b->c;
  const Box vb = visc_box(c);
  k_visc_grad<<<box_blocks(vb, 256), 256, 0, b->stream>>>(c, vb, b->prim, b->vtmp);
  CBNS_HIP(b, hipGetLastError());
  k_visc_terms<<<box_blocks(vb, 256), 256, 0, b->stream>>>(c, vb, b->prim, b->met, b->mu, b->vtmp);
  CBNS_HIP(b, hipGetLastError());
  k_visc_flux<<<box_blocks(vb, 256), 256, 0, b->stream>>>(c, vb, b->prim, b->met, b->vtmp);
  CBNS_HIP(b, hipGetLastError());
  const Box ub = update_box(c);
  k_visc_contrib<<<box_blocks(ub, 256), 256, 0, b->stream>>>(c, ub, b->vtmp, b->dfv);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

int launch_update_rk3(cbns_block* b, double dt, int stage) {
  const Box ub = update_box(b->c);
  k_update_rk3<<<box_blocks(ub, 256), 256, 0, b->stream>>>(b->c, ub, dt, stage, b->met, b->flux,
                                                           b->dfv, b->Q, b->qp);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

}  // namespace cbns
