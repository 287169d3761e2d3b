// k_bc.hip -- set_bc_calc_primitive: interior primitives and ghost/boundary
// fill (src/primitive_bc_kernel.cu:13-494).  The reference runs every rule in
// one kernel whose phases are ordered only by block-local barriers (:187,
// :414), so a thread can read a boundary value another block has not written
// yet.  Here each phase is its own launch on the block's stream, in the order
// the barriers intend: interior -> xi faces -> eta=0 writes -> eta=0
// primitives -> eta=JM-1 -> zeta faces -> periodic wraps.  All of it is
// integer index work plus the reference's per-node formulas, and it is
// compared bit for bit with the oracle.
#include "cbns_internal.hpp"

namespace cbns {

namespace {

constexpr int NGR = CBNS_NG_REF;  // ghost depth the reference's rules write

__device__ inline void prim_from_q(const Consts& c, const double* __restrict__ Q,
                                   double* __restrict__ prim, int qi, int pi) {
  const long S = c.S;
  const double rr = Q[qi];
  const double uu = Q[S + qi] / rr;
  const double vv = Q[2 * S + qi] / rr;
  const double ww = Q[3 * S + qi] / rr;
  const double pp = c.gam1 * (Q[4 * S + qi] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  prim[pi] = rr;
  prim[S + pi] = uu;
  prim[2 * S + pi] = vv;
  prim[3 * S + pi] = ww;
  prim[4 * S + pi] = pp;
  if (c.viscous) prim[5 * S + pi] = c.gM2 * pp / rr;
}

__device__ inline void copy_q(const Consts& c, double* __restrict__ Q, int d, int s) {
#pragma unroll
  for (int n = 0; n < NEQ; n++) Q[(long)n * c.S + d] = Q[(long)n * c.S + s];
}

// interior primitives, src/primitive_bc_kernel.cu:56-79
__global__ void k_prim_box(Consts c, Box box, const double* __restrict__ Q,
                           double* __restrict__ prim) {
  int i, j, k;
  if (!box_node3(box, i, j, k)) return;
  const int o = idx(c, i, j, k);
  prim_from_q(c, Q, prim, o, o);
}

// xi = 0 and xi = IM-1 far field, :84-186.  box: (side, j, k)
__global__ void k_bc_xi(Consts c, Box box, double* __restrict__ Q, double* __restrict__ prim) {
  int side, j, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, side, j, k)) return;
  const long S = c.S;
  const int s = idx(c, side == 0 ? 1 : c.IM - 2, j, k);
  const double rho_1 = Q[s], rho_u = Q[S + s], rho_v = Q[2 * S + s], rho_w = Q[3 * S + s];
  const double rho_et =
      c.p_inf / c.gam1 + 0.5 * (rho_u * rho_u + rho_v * rho_v + rho_w * rho_w) / rho_1;
  const double rr = rho_1, uu = rho_u / rr, vv = rho_v / rr, ww = rho_w / rr;
  const double pp = c.gam1 * (rho_et - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  for (int li = 0; li <= NGR; li++) {
    const int d = side == 0 ? idx(c, -li, j, k) : idx(c, c.IM - 1 + li, j, k);
    Q[d] = rho_1;
    Q[S + d] = rho_u;
    Q[2 * S + d] = rho_v;
    Q[3 * S + d] = rho_w;
    Q[4 * S + d] = rho_et;
    prim[d] = rr;
    prim[S + d] = uu;
    prim[2 * S + d] = vv;
    prim[3 * S + d] = ww;
    prim[4 * S + d] = pp;
    if (c.viscous) prim[5 * S + d] = c.gM2 * pp / rr;
  }
}

// eta = 0: wake cut (i < i_begin-1) and wall (i < i_end), :190-352.
// box: (i, k).  Mirror threads of the wake write identical values.
__global__ void k_bc_eta0_write(Consts c, Box box, double* __restrict__ Q,
                                const double* __restrict__ met) {
  int i, k, unused;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, k, unused)) return;
  const long S = c.S;
  const double gam1 = c.gam1;
  if ((long)c.i_begin - 1 > i) {
    const int a = idx(c, i, 1, k);
    double rr = Q[a], uu = Q[S + a] / rr, vv = Q[2 * S + a] / rr, ww = Q[3 * S + a] / rr;
    double pp = gam1 * (Q[4 * S + a] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
    const int b = idx(c, c.IM - 1 - i, 1, k);
    const double rr1 = Q[b], uu1 = Q[S + b] / rr1, vv1 = Q[2 * S + b] / rr1,
                 ww1 = Q[3 * S + b] / rr1;
    const double pp1 = gam1 * (Q[4 * S + b] - 0.5 * rr1 * (uu1 * uu1 + vv1 * vv1 + ww1 * ww1));
    rr += rr1; uu += uu1; vv += vv1; ww += ww1; pp += pp1;
    rr *= 0.5; uu *= 0.5; vv *= 0.5; ww *= 0.5; pp *= 0.5;
    const int d = idx(c, i, 0, k);
    Q[d] = rr;
    Q[S + d] = rr * uu;
    Q[2 * S + d] = rr * vv;
    Q[3 * S + d] = rr * ww;
    Q[4 * S + d] = pp / gam1 + 0.5 * rr * (uu * uu + vv * vv + ww * ww);
    copy_q(c, Q, idx(c, c.IM - 1 - i, 0, k), d);
    for (int lj = 1; lj <= NGR; lj++) {
      copy_q(c, Q, idx(c, i, -lj, k), idx(c, c.IM - 1 - i, lj, k));
      copy_q(c, Q, idx(c, c.IM - 1 - i, -lj, k), idx(c, i, lj, k));
    }
  } else if ((long)c.i_end > i) {
    if (c.viscous) {
      // no-slip wall :264-295
      const int a = idx(c, i, 1, k);
      const double rho_1 = Q[a], rho_u = Q[S + a], rho_v = Q[2 * S + a], rho_w = Q[3 * S + a];
      const double rho_et = Q[4 * S + a];
      const double pp =
          gam1 * (rho_et - 0.5 * (rho_u * rho_u + rho_v * rho_v + rho_w * rho_w) / rho_1);
      const int d = idx(c, i, 0, k);
      Q[d] = rho_1;
      Q[S + d] = 0.0;
      Q[2 * S + d] = 0.0;
      Q[3 * S + d] = 0.0;
      Q[4 * S + d] = pp / gam1;
      for (int lj = 1; lj <= NGR; lj++) {
        const int s = idx(c, i, lj, k), t = idx(c, i, -lj, k);
        Q[t] = Q[s];
        Q[S + t] = -Q[S + s];
        Q[2 * S + t] = -Q[2 * S + s];
        Q[3 * S + t] = -Q[3 * S + s];
        Q[4 * S + t] = Q[4 * S + s];
      }
    } else {
      // slip wall :297-349 (energy formula exactly as the reference writes it)
      const int a = idx(c, i, 1, k);
      const double rr = Q[a];
      double uu = Q[S + a] / rr, vv = Q[2 * S + a] / rr;
      const double ww = Q[3 * S + a] / rr;
      const double pp = gam1 * (Q[4 * S + a] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
      const int m = idx(c, i, 0, k);
      double t_x = met[ETA_Y * S + m];
      double t_y = -met[ETA_X * S + m];
      double t_abs = met[JAC * S + m];
      t_x *= t_abs;
      t_y *= t_abs;
      t_abs = sqrt(t_x * t_x + t_y * t_y);
      t_x /= t_abs;
      t_y /= t_abs;
      for (int lj = 1; lj <= NGR; lj++) {
        const int s = idx(c, i, lj, k), t = idx(c, i, -lj, k);
        t_abs = Q[S + s] * t_x + Q[2 * S + s] * t_y;
        Q[t] = Q[s];
        Q[S + t] = 2 * t_abs * t_x - Q[S + s];
        Q[2 * S + t] = 2 * t_abs * t_y - Q[2 * S + s];
        Q[3 * S + t] = Q[3 * S + s];
        Q[4 * S + t] = Q[4 * S + s] - 0.5 * (Q[S + s] * Q[S + s] + Q[2 * S + s] * Q[2 * S + s] -
                                             Q[S + t] * Q[S + t] + Q[2 * S + t] * Q[2 * S + t]) /
                                          Q[t];
      }
      t_abs = uu * t_x + vv * t_y;
      uu = t_abs * t_x;
      vv = t_abs * t_y;
      const int d = idx(c, i, 0, k);
      Q[d] = rr;
      Q[S + d] = rr * uu;
      Q[2 * S + d] = rr * vv;
      Q[3 * S + d] = rr * ww;
      Q[4 * S + d] = pp / gam1 + 0.5 * rr * (uu * uu + vv * vv + ww * ww);
    }
  }
}

// eta = 0 primitives, :354-376.  box: (i, lj in 0..NGR, k)
__global__ void k_bc_eta0_prim(Consts c, Box box, const double* __restrict__ Q,
                               double* __restrict__ prim) {
  int i, lj, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, lj, k)) return;
  const int o = idx(c, i, -lj, k);
  prim_from_q(c, Q, prim, o, o);
}

// eta = JM-1, :383-413: primitives at JM-1 and JM from Q(JM-1).  box: (i, lj in 0..NGR-1, k)
__global__ void k_bc_etaN(Consts c, Box box, const double* __restrict__ Q,
                          double* __restrict__ prim) {
  int i, lj, k;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, lj, k)) return;
  prim_from_q(c, Q, prim, idx(c, i, c.JM - 1, k), idx(c, i, c.JM - 1 + lj, k));
}

// zeta = 0 and zeta = KM-1 copies, :417-492.  box: (i, j, side)
__global__ void k_bc_zeta(Consts c, Box box, double* __restrict__ Q, double* __restrict__ prim) {
  int i, j, side;
  if (!box_node(box, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, side)) return;
  const int s = side == 0 ? idx(c, i, j, 1) : idx(c, i, j, c.KM - 2);
  for (int lk = 0; lk <= NGR; lk++) {
    const int d = side == 0 ? idx(c, i, j, -lk) : idx(c, i, j, c.KM - 1 + lk);
    copy_q(c, Q, d, s);
  }
  for (int lk = 0; lk <= NGR; lk++) {
    const int d = side == 0 ? idx(c, i, j, -lk) : idx(c, i, j, c.KM - 1 + lk);
    prim_from_q(c, Q, prim, s, d);
  }
}

// Periodic wrap (per: bit a set for each periodic axis a).  Every ghost node
// that lies in the ghost layers of a periodic axis copies Q and the
// primitives of the node at its coordinates wrapped on every periodic axis
// at once.  Those are the values that wrapping one axis after another over
// the full ghosted extent of the other two leaves at every node, edges and
// corners included, once the boundary rules of the other axes have written
// their own ghost layers (they run first); ghost nodes of non-periodic axes
// alone are left as those rules wrote them.  The ghost shell is enumerated
// slab by slab: zeta slabs (full extent), then eta slabs (zeta in range),
// then xi slabs (eta, zeta in range).
__global__ void k_wrap_shell(Consts c, int per, double* __restrict__ Q,
                             double* __restrict__ prim, int pmask) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned PX = c.PX, PY = c.PY, twoG = 2 * G;
  const unsigned nz = twoG * PX * PY, ny = twoG * PX * (unsigned)c.KM,
                 nx = twoG * (unsigned)c.JM * (unsigned)c.KM;
  int i, j, k;
  if (t < nz) {
    const unsigned g = t / (PX * PY), r = t - g * PX * PY;
    k = g < (unsigned)G ? (int)g - G : c.KM + (int)g - G;
    j = (int)(r / PX) - G;
    i = (int)(r - (r / PX) * PX) - G;
  } else if (t < nz + ny) {
    const unsigned u = t - nz;
    const unsigned kk = u / (twoG * PX), r = u - kk * twoG * PX;
    const unsigned g = r / PX;
    k = (int)kk;
    j = g < (unsigned)G ? (int)g - G : c.JM + (int)g - G;
    i = (int)(r - g * PX) - G;
  } else if (t < nz + ny + nx) {
    const unsigned u = t - nz - ny;
    const unsigned kk = u / (twoG * (unsigned)c.JM), r = u - kk * twoG * (unsigned)c.JM;
    const unsigned jj = r / twoG, g = r - jj * twoG;
    k = (int)kk;
    j = (int)jj;
    i = g < (unsigned)G ? (int)g - G : c.IM + (int)g - G;
  } else {
    return;
  }
  const bool gi = (per & 1) && (i < 0 || i >= c.IM);
  const bool gj = (per & 2) && (j < 0 || j >= c.JM);
  const bool gk = (per & 4) && (k < 0 || k >= c.KM);
  if (!(gi || gj || gk)) return;  // a ghost node of non-periodic axes only
  const int si = !(per & 1) ? i : (i < 0 ? i + c.IM : (i >= c.IM ? i - c.IM : i));
  const int sj = !(per & 2) ? j : (j < 0 ? j + c.JM : (j >= c.JM ? j - c.JM : j));
  const int sk = !(per & 4) ? k : (k < 0 ? k + c.KM : (k >= c.KM ? k - c.KM : k));
  const int d = idx(c, i, j, k), s = idx(c, si, sj, sk);
#pragma unroll
  for (int q = 0; q < NEQ; q++) Q[(long)q * c.S + d] = Q[(long)q * c.S + s];
#pragma unroll
  for (int q = 0; q < 6; q++)
    if (pmask >> q & 1) prim[(long)q * c.S + d] = prim[(long)q * c.S + s];
}

template <class K, class... A>
int run(cbns_block* b, const Box& box, K kern, A... args) {
  if (box.count() <= 0) return CBNS_OK;
  kern<<<box_blocks(box, 256), 256, 0, b->stream>>>(b->c, box, args...);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

}  // namespace

// Boundary rules and ghost fill for state Q whose interior primitives are
// already in b->prim (the fused stage writes them with the new state).
int launch_bc_only(cbns_block* b, double* Q, int pmask) {
  const Consts& c = b->c;
  int rc;
  const int jlo = upd_lo(c, 1), jhi = upd_hi(c, 1);
  const int klo = upd_lo(c, 2), khi = upd_hi(c, 2);
  if (!periodic(c, 0)) {
    Box box{{0, jlo, klo}, {2, jhi - jlo + 1, khi - klo + 1}};
    if ((rc = run(b, box, k_bc_xi, Q, b->prim))) return rc;
  }
  if (!periodic(c, 1)) {
    Box bw{{0, klo, 0}, {c.IM, khi - klo + 1, 1}};
    if ((rc = run(b, bw, k_bc_eta0_write, Q, (const double*)b->met))) return rc;
    Box bp{{0, 0, klo}, {c.IM, NGR + 1, khi - klo + 1}};
    if ((rc = run(b, bp, k_bc_eta0_prim, (const double*)Q, b->prim))) return rc;
    Box bn{{0, 0, klo}, {c.IM, NGR, khi - klo + 1}};
    if ((rc = run(b, bn, k_bc_etaN, (const double*)Q, b->prim))) return rc;
  }
  if (!periodic(c, 2)) {
    Box box{{0, 0, 0}, {c.IM, c.JM, 2}};
    if ((rc = run(b, box, k_bc_zeta, Q, b->prim))) return rc;
  }
  if (pmask < 0) pmask = c.viscous ? 0x3F : 0x1F;
  const int per = (periodic(c, 0) ? 1 : 0) | (periodic(c, 1) ? 2 : 0) | (periodic(c, 2) ? 4 : 0);
  if (per) {
    // one launch for all periodic axes (one per axis cost 27 us each on a
    // 256x128x128 block with two periodic axes)
    const long cnt = (long)c.S - (long)c.IM * c.JM * c.KM;
    k_wrap_shell<<<(unsigned)((cnt + 255) / 256), 256, 0, b->stream>>>(c, per, Q, b->prim,
                                                                       pmask);
    CBNS_HIP(b, hipGetLastError());
  }
  return CBNS_OK;
}

int launch_bc_primitive(cbns_block* b, double* Q) {
  const Consts& c = b->c;
  Box box;
  for (int a = 0; a < 3; a++) {
    box.lo[a] = upd_lo(c, a);
    box.n[a] = upd_hi(c, a) - upd_lo(c, a) + 1;
  }
  k_prim_box<<<box_grid(box), box_block(), 0, b->stream>>>(c, box, Q, b->prim);
  CBNS_HIP(b, hipGetLastError());
  return launch_bc_only(b, Q);
}

}  // namespace cbns
