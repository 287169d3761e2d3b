// k_unfused.hip -- the reference's launcher decomposition, one entry point per
// block3d_cuda call: calc_numerical_flux (face fluxes to HBM),
// calc_viscous_flux_contribution (dfv to HBM, k_viscous.hip) and update_rk3.  The fused sweep kernels (k_fused.hip) replace this chain in
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
  if (!box_node3(box, i, j, k)) return;
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
  // the Roe average from the node primitives, as the reference reads them
  Basis<double> B;
  roe_basis(g, pl[0], pl[1], pl[2], pl[3], pl[4], pr[0], pr[1], pr[2], pr[3], pr[4], nx, ny, nz,
            B);
  face_flux_basis<double>(g, q, B, nx, ny, nz, f);
#pragma unroll
  for (int n = 0; n < 5; n++) flux[(d * 5 + n) * S + a] = f[n];
}

// src/runge_kutta_kernel.cu:13-80
__global__ void k_update_rk3(Consts c, Box box, double dt, int stage, const double* __restrict__ met,
                             const double* __restrict__ flux, const double* __restrict__ dfv,
                             double* __restrict__ Q, double* __restrict__ Qp) {
  int i, j, k;
  if (!box_node3(box, i, j, k)) return;
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

Box update_box(const Consts& c) {
  Box b;
  for (int a = 0; a < 3; a++) {
    b.lo[a] = upd_lo(c, a);
    b.n[a] = upd_hi(c, a) - upd_lo(c, a) + 1;
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
    k_face_flux<<<box_grid(box), box_block(), 0, b->stream>>>(c, box, d, Q, b->prim, b->met,
                                                             b->flux);
    CBNS_HIP(b, hipGetLastError());
  }
  return CBNS_OK;
}

int launch_viscous_unfused(cbns_block* b) { return launch_viscous(b, b->vtmp, b->dfv, 0); }

int launch_update_rk3(cbns_block* b, double dt, int stage) {
  const Box ub = update_box(b->c);
  k_update_rk3<<<box_grid(ub), box_block(), 0, b->stream>>>(b->c, ub, dt, stage, b->met, b->flux,
                                                           b->dfv, b->Q, b->qp);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

}  // namespace cbns
