// face_math.hpp -- per-face inviscid flux: Roe-averaged characteristic basis,
// characteristic WENO5-JS reconstruction of the conservative state on both
// sides, Lax-Friedrichs-type flux.  Method of src/rec_riemann_kernel.cu:44-89
// (calc_eigenvectors_roe, rec_weno5 x2, lf_flux), restructured for gfx950:
//   - the six stencil nodes are projected once (the two sides share four);
//   - L (left eigenvectors) is applied in closed form: rows 0..2 share two
//     dot products, rows 3..4 are the tangent projections;
//   - R (right eigenvectors) is applied in closed form;
//   - each WENO5 evaluation uses one division (weights multiplied through by
//     IS1^2 IS2^2 IS3^2, the linear weights and the 1/6 of the candidate
//     stencils folded into the candidates and the normaliser) instead of
//     seven, on first differences relative to the centre value; the left
//     and right evaluations of a field share their differences, two
//     smoothness terms and, in double, one reciprocal;
//   - the Lax-Friedrichs flux reuses 1/rho, |u|^2/2 and E of the
//     reconstructed states (the reference re-derives E from p);
//   - 1/x is formed once and multiplied where the reference divides by the
//     same x repeatedly;
//   - reciprocals and reciprocal square roots start from the hardware
//     v_rcp_f64 / v_rsq_f64 estimate (relative error <= 5.3e-8 on MI355X
//     over 2^20 arguments in [2^-30, 2^31]) and take one Newton step, which
//     leaves <= 2.3e-15 (rcp) and <= 4.2e-15 (rsq); an IEEE f64 divide or
//     sqrt costs ~11-12 VALU instructions on gfx950, this 3-4, and a second
//     step (<= 1.4e-16) would cost 2-3 more in a VALU-bound sweep.
//     Arguments are positive and finite wherever they are used.
// The result equals the reference formula up to rounding (the stage parity
// bar is 1e-12 relative; tests/test_gpu_parity.py).
#pragma once

#include <hip/hip_runtime.h>

namespace cbns {

__device__ __forceinline__ double fast_rcp(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// 1/sqrt(x): y <- y + y (1/2 - x y^2 / 2)
__device__ __forceinline__ double fast_rsq(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return fma(y, fma(-0.5 * x * y, y, 0.5), y);
}
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

template <class T>
struct GasT {
  T gamma, gam1, gam1_inv;
};

// Roe-averaged characteristic basis of a face, src/calc_eigenvectors_roe.cu:28-150.
template <class T>
struct Basis {
  T nx, ny, nz, lx, ly, lz, mx, my, mz;  // orthonormal (n normal, l,m tangents)
  T uu, vv, ww, H, c, q2, qn, ql, qm;    // Roe state
  T K, o;                                // gam1/(2c^2), 0.5/c
};

// Characteristic basis from the face normal and the Roe-averaged state
// (u, v, w, H) (src/calc_eigenvectors_roe.cu:28-150).
template <class T>
__device__ __forceinline__ void basis_from_roe(const GasT<T>& g, T uu, T vv, T ww, T H, T n_x,
                                               T n_y, T n_z, Basis<T>& B) {
  T l_x = T(0), l_y = n_z, l_z = -n_y;
  T m_x, m_y, m_z;
  // |l| > 1e-8 of the reference, compared squared
  T l2 = l_y * l_y + l_z * l_z;
  if (l2 > T(1.0e-16)) {
    m_x = -n_y * n_y - n_z * n_z;
    m_y = n_x * n_y;
    m_z = n_x * n_z;
  } else {
    l_x = -n_z; l_y = T(0); l_z = n_x;
    m_x = n_x * n_y;
    m_y = -n_x * n_x - n_z * n_z;
    m_z = n_y * n_z;
    l2 = l_x * l_x + l_z * l_z;
  }
  const T inv_n = fast_rsq(n_x * n_x + n_y * n_y + n_z * n_z);
  B.nx = n_x * inv_n; B.ny = n_y * inv_n; B.nz = n_z * inv_n;
  const T inv_l = fast_rsq(l2);
  B.lx = l_x * inv_l; B.ly = l_y * inv_l; B.lz = l_z * inv_l;
  // |m|^2 = |l|^2 |n|^2 for either choice of tangents
  const T inv_m = inv_l * inv_n;
  B.mx = m_x * inv_m; B.my = m_y * inv_m; B.mz = m_z * inv_m;

  B.uu = uu; B.vv = vv; B.ww = ww; B.H = H;
  B.q2 = B.uu * B.uu + B.vv * B.vv + B.ww * B.ww;
  const T c2 = g.gam1 * (B.H - T(0.5) * B.q2);
  const T inv_c = fast_rsq(c2);
  B.c = c2 * inv_c;
  B.qn = B.uu * B.nx + B.vv * B.ny + B.ww * B.nz;
  B.ql = B.uu * B.lx + B.vv * B.ly + B.ww * B.lz;
  B.qm = B.uu * B.mx + B.vv * B.my + B.ww * B.mz;
  B.o = T(0.5) * inv_c;
  B.K = T(0.5) * g.gam1 * inv_c * inv_c;
}

// From the two node states as (rho, u, v, w, H): Roe average with
// r = sqrt(rho_R / rho_L).
template <class T>
__device__ __forceinline__ void roe_basis_h(const GasT<T>& g, T rL, T uL, T vL, T wL, T HL,
                                            T rR, T uR, T vR, T wR, T HR, T n_x, T n_y, T n_z,
                                            Basis<T>& B) {
  const T ratio = rR * fast_rsq(rR * rL);  // sqrt(rR / rL)
  const T inv1r = fast_rcp(T(1) + ratio);
  basis_from_roe(g, (uL + ratio * uR) * inv1r, (vL + ratio * vR) * inv1r,
                 (wL + ratio * wR) * inv1r, (HL + ratio * HR) * inv1r, n_x, n_y, n_z, B);
}

// From the two node states as conservative states: with a = 1/sqrt(rho),
// the Roe average sum sqrt(rho) f / sum sqrt(rho) is sum a (rho f) / sum
// sqrt(rho), and rho H = E + p.
template <class T>
__device__ __forceinline__ void roe_basis_q(const GasT<T>& g, const T qL[5], const T qR[5],
                                            T n_x, T n_y, T n_z, Basis<T>& B) {
  const T aL = fast_rsq(qL[0]), aR = fast_rsq(qR[0]);
  const T D = fast_rcp(qL[0] * aL + qR[0] * aR);
  const T wL = aL * D, wR = aR * D;
  // p = gam1 (E - |rho u|^2 / (2 rho)), 1/rho = a^2
  const T pL = g.gam1 * (qL[4] - T(0.5) * (qL[1] * qL[1] + qL[2] * qL[2] + qL[3] * qL[3]) * (aL * aL));
  const T pR = g.gam1 * (qR[4] - T(0.5) * (qR[1] * qR[1] + qR[2] * qR[2] + qR[3] * qR[3]) * (aR * aR));
  basis_from_roe(g, qL[1] * wL + qR[1] * wR, qL[2] * wL + qR[2] * wR, qL[3] * wL + qR[3] * wR,
                 (qL[4] + pL) * wL + (qR[4] + pR) * wR, n_x, n_y, n_z, B);
}

// From (rho, u, v, w, p): H = c^2/gam1 + |u|^2/2 with c^2 = gamma p / rho.
template <class T>
__device__ __forceinline__ void roe_basis(const GasT<T>& g, T rL, T uL, T vL, T wL, T pL,
                                          T rR, T uR, T vR, T wR, T pR, T n_x, T n_y, T n_z,
                                          Basis<T>& B) {
  const T irL = fast_rcp(rL), irR = fast_rcp(rR);
  const T HL = g.gamma * pL * irL * g.gam1_inv + T(0.5) * (uL * uL + vL * vL + wL * wL);
  const T HR = g.gamma * pR * irR * g.gam1_inv + T(0.5) * (uR * uR + vR * vR + wR * wR);
  roe_basis_h(g, rL, uL, vL, wL, HL, rR, uR, vR, wR, HR, n_x, n_y, n_z, B);
}

// A reconstructed face state: primitives plus 1/rho, |u|^2/2 and the
// conservative state they were derived from (momenta, total energy E;
// p = gam1 (E - rho |u|^2/2)).
template <class T>
struct Side {
  T r, u, v, w, p, ir, k, E, ru, rv, rw;
};

// conservative -> face state (src/weno.cu:113-118 / src/primitive_bc_kernel.cu:63-67)
template <class T>
__device__ __forceinline__ Side<T> q_to_side(const GasT<T>& g, const T q[5]) {
  Side<T> s;
  s.r = q[0];
  s.ir = fast_rcp(s.r);
  s.u = q[1] * s.ir;
  s.v = q[2] * s.ir;
  s.w = q[3] * s.ir;
  s.k = T(0.5) * (s.u * s.u + s.v * s.v + s.w * s.w);
  s.E = q[4];
  s.p = g.gam1 * (s.E - s.r * s.k);
  s.ru = q[1];
  s.rv = q[2];
  s.rw = q[3];
  return s;
}

// total enthalpy of a face state, as roe_basis forms it
template <class T>
__device__ __forceinline__ T side_h(const GasT<T>& g, const Side<T>& s) {
  return g.gamma * s.p * s.ir * g.gam1_inv + s.k;
}

// q = R W, then to the face state (src/weno.cu:106-118)
template <class T>
__device__ __forceinline__ Side<T> char_to_side(const GasT<T>& g, const Basis<T>& B,
                                                const T W[5]) {
  const T s = W[0] + W[2];
  const T dl = B.c * (W[2] - W[0]);
  T q[5];
  q[0] = s + W[1];
  q[1] = B.uu * q[0] + B.nx * dl + B.lx * W[3] + B.mx * W[4];
  q[2] = B.vv * q[0] + B.ny * dl + B.ly * W[3] + B.my * W[4];
  q[3] = B.ww * q[0] + B.nz * dl + B.lz * W[3] + B.mz * W[4];
  q[4] = B.H * s + T(0.5) * B.q2 * W[1] + B.qn * dl + B.ql * W[3] + B.qm * W[4];
  return q_to_side(g, q);
}

// src/lf_flux.cu:22-78 (dissipation speed |sum_d max_LR(|u_d|+c) n_d|),
//   f = (F(L).n + F(R).n - alpha (Q_R - Q_L)) / 2,  F(Q).n = (rho q_n, rho q_n u + p n, ...,
//   rho q_n H),
// on the reconstructed conservative states: the mass fluxes are (rho u).n
// and the jumps are those of rho u and E (the reference re-forms rho u and
// E from the primitives, which equals them up to rounding).
template <class T>
__device__ __forceinline__ void lf_flux(const GasT<T>& g, const Side<T>& R, const Side<T>& L,
                                        T nx, T ny, T nz, T f[5]) {
  const T aR = g.gamma * R.p * R.ir;
  const T cR = aR * fast_rsq(aR);
  const T HR = aR * g.gam1_inv + R.k;
  const T aL = g.gamma * L.p * L.ir;
  const T cL = aL * fast_rsq(aL);
  const T HL = aL * g.gam1_inv + L.k;
  const T s0 = fmax(fabs(R.u) + cR, fabs(L.u) + cL);
  const T s1 = fmax(fabs(R.v) + cR, fabs(L.v) + cL);
  const T s2 = fmax(fabs(R.w) + cR, fabs(L.w) + cL);
  const T ha = T(0.5) * fabs(s0 * nx + s1 * ny + s2 * nz);
  const T hmL = T(0.5) * (L.ru * nx + L.rv * ny + L.rw * nz);  // rho q_n / 2
  const T hmR = T(0.5) * (R.ru * nx + R.rv * ny + R.rw * nz);
  const T hp = T(0.5) * (L.p + R.p);
  f[0] = hmL + hmR - ha * (R.r - L.r);
  f[1] = hmL * L.u + hmR * R.u + hp * nx - ha * (R.ru - L.ru);
  f[2] = hmL * L.v + hmR * R.v + hp * ny - ha * (R.rv - L.rv);
  f[3] = hmL * L.w + hmR * R.w + hp * nz - ha * (R.rw - L.rw);
  f[4] = hmL * HL + hmR * HR - ha * (R.E - L.E);
}

// conservative -> primitive (src/primitive_bc_kernel.cu:63-67)
template <class T>
__device__ __forceinline__ void q_to_prim(const GasT<T>& g, const T q[5], T& r, T& u, T& v, T& w,
                                          T& p) {
  r = q[0];
  const T ir = fast_rcp(r);
  u = q[1] * ir;
  v = q[2] * ir;
  w = q[3] * ir;
  p = g.gam1 * (q[4] - T(0.5) * r * (u * u + v * v + w * w));
}

// WENO5-JS on both sides of one face from the 6 projected stencil values
// w0..w5 (nodes m-2..m+3): left = weno5(w0..w4), right = weno5(w5..w1).
//
// Written on the first differences e_k = w_{k+1} - w_k, which the two sides
// share (the right side reads them reversed and negated), and relative to
// the centre value.  With s_k = IS_k^2 the JS weights are d_k / s_k,
// d = (0.1, 0.6, 0.3), and the candidates f_k / 6; scaling numerator and
// normaliser by s1 s2 s3 / 0.6 and taking the centre c out,
//   value = c + (p1 g1 + p2 g2 + p3 g3) / (p1 + 6 p2 + 3 p3),
//   p1 = s2 s3, p2 = s1 s3, p3 = s1 s2,
//   g1 = f1/6 - c = 5/6 E1 - 1/3 E0,  g2 = f2 - 6c = E1 + 2 E2,
//   g3 = f3/2 - 3c = 2 E2 - E3/2,
// with E0..E3 = b-a, c-b, d-c, e-d of the side's stencil a..e.  The
// smoothness indicators use IS1 = 13/12 (E1-E0)^2 + (0.5 (E1-E0) + E1)^2,
// IS2 = 13/12 (E2-E1)^2 + (0.5 (E1+E2))^2, IS3 = 13/12 (E3-E2)^2 +
// (0.5 (E3-E2) - E2)^2 (each + eps), the same polynomials as
// src/weno.cu:19-42 regrouped; the (13/12)(.)^2 + eps terms of the two
// middle sub-stencils are the same second differences on both sides.
template <class T>
__device__ __forceinline__ void weno5_side(T E0, T E1, T E2, T E3, T d1, T ta, T tb, T& num,
                                           T& den) {
  // d1 = E1 - E0; ta = 13/12 (E2-E1)^2 + eps; tb = 13/12 (E3-E2)^2 + eps
  const T eps = T(1.0e-6), k13 = T(13.0 / 12.0);
  const T h2 = T(0.5) * d1 + E1;
  const T h4 = T(0.5) * (E1 + E2);
  const T h6 = T(0.5) * (E3 - E2) - E2;
  const T IS1 = k13 * d1 * d1 + (h2 * h2 + eps);
  const T IS2 = ta + h4 * h4;
  const T IS3 = tb + h6 * h6;
  const T s1 = IS1 * IS1, s2 = IS2 * IS2, s3 = IS3 * IS3;
  const T p1 = s2 * s3, p2 = s1 * s3, p3 = s1 * s2;
  const T g1 = T(5.0 / 6.0) * E1 - T(1.0 / 3.0) * E0;
  const T g2 = E1 + T(2) * E2;
  const T g3 = T(2) * E2 - T(0.5) * E3;
  num = p1 * g1 + p2 * g2 + p3 * g3;
  den = p1 + T(6) * p2 + T(3) * p3;
}

template <class T>
__device__ __forceinline__ void weno5_pair(T w0, T w1, T w2, T w3, T w4, T w5, T& left,
                                           T& right) {
  const T k13 = T(13.0 / 12.0), eps = T(1.0e-6);
  const T e01 = w1 - w0, e12 = w2 - w1, e23 = w3 - w2, e34 = w4 - w3, e45 = w5 - w4;
  const T dc = e23 - e12;  // second difference centred at m
  const T dd = e34 - e23;  // centred at m+1
  const T tc = k13 * dc * dc + eps, td = k13 * dd * dd + eps;
  T nl, dl, nr, dr;
  weno5_side(e01, e12, e23, e34, e12 - e01, tc, td, nl, dl);
  weno5_side(-e45, -e34, -e23, -e12, e45 - e34, td, tc, nr, dr);
  if constexpr (sizeof(T) == sizeof(double)) {
    // one reciprocal for both sides: dl, dr lie in [1e-24, 1e120] for any
    // smoothness indicator below 1e30, so dl dr stays in range
    const T x = fast_rcp(dl * dr);
    left = w2 + nl * (dr * x);
    right = w3 + nr * (dl * x);
  } else {
    left = w2 + nl * fast_rcp(dl);
    right = w3 + nr * fast_rcp(dr);
  }
}

// Full face: q[s][c] holds nodes m-2..m+3 (face between m and m+1); B is
// the face's characteristic basis (face_flux: the Roe average of nodes m and
// m+1); (nx,ny,nz) the unnormalised face vector
// 0.5 (J grad xi |_m + J grad xi |_{m+1}).  Characteristic fields are
// reconstructed one at a time so only six projected values are live.
template <class T>
__device__ __forceinline__ void face_flux_basis(const GasT<T>& g, const T q[6][5],
                                                const Basis<T>& B, T nx, T ny, T nz, T f[5]) {
  // rows 0..2 of L q share A = K (E - u.m + q2/2 rho) and Bn = o (qn rho - n.m)
  T A[6], Bn[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    A[s] = B.K * (q[s][4] - (B.uu * q[s][1] + B.vv * q[s][2] + B.ww * q[s][3]) +
                  T(0.5) * B.q2 * q[s][0]);
    Bn[s] = B.o * (B.qn * q[s][0] - (B.nx * q[s][1] + B.ny * q[s][2] + B.nz * q[s][3]));
  }
  T WL[5], WR[5];
  weno5_pair(A[0] + Bn[0], A[1] + Bn[1], A[2] + Bn[2], A[3] + Bn[3], A[4] + Bn[4], A[5] + Bn[5],
             WL[0], WR[0]);
  weno5_pair(q[0][0] - T(2) * A[0], q[1][0] - T(2) * A[1], q[2][0] - T(2) * A[2],
             q[3][0] - T(2) * A[3], q[4][0] - T(2) * A[4], q[5][0] - T(2) * A[5], WL[1], WR[1]);
  weno5_pair(A[0] - Bn[0], A[1] - Bn[1], A[2] - Bn[2], A[3] - Bn[3], A[4] - Bn[4], A[5] - Bn[5],
             WL[2], WR[2]);
  T w[6];
#pragma unroll
  for (int s = 0; s < 6; s++)
    w[s] = -B.ql * q[s][0] + B.lx * q[s][1] + B.ly * q[s][2] + B.lz * q[s][3];
  weno5_pair(w[0], w[1], w[2], w[3], w[4], w[5], WL[3], WR[3]);
#pragma unroll
  for (int s = 0; s < 6; s++)
    w[s] = -B.qm * q[s][0] + B.mx * q[s][1] + B.my * q[s][2] + B.mz * q[s][3];
  weno5_pair(w[0], w[1], w[2], w[3], w[4], w[5], WL[4], WR[4]);
  const Side<T> L = char_to_side(g, B, WL), R = char_to_side(g, B, WR);
  lf_flux(g, R, L, nx, ny, nz, f);
}

template <class T>
__device__ __forceinline__ void face_flux(const GasT<T>& g, const T q[6][5], T nx, T ny, T nz,
                                          T f[5]) {
  Basis<T> B;
  roe_basis_q(g, q[2], q[3], nx, ny, nz, B);
  face_flux_basis(g, q, B, nx, ny, nz, f);
}

// Precision form of face_flux: the Roe basis, the characteristic projection
// and WENO5 run in T (float) on the deviations of the six stencil nodes from
// node m (WENO5-JS is translation invariant: weno5(x + c) = weno5(x) + c, and
// the projection is linear), so T's rounding acts on the local variation, not
// on the state itself.  The reconstructed states q_m + dq_L/R and the
// Lax-Friedrichs flux are formed in double, so face-flux differences carry no
// float cancellation error.
// The Roe-average inputs of a node in the precision form: rho, u, v, w and
// H of its double face state, rounded to T.
template <class T>
struct NodeT {
  T r, u, v, w, H;
};
template <class T>
__device__ __forceinline__ NodeT<T> node_t(const GasT<double>& g, const Side<double>& s) {
  return {(T)s.r, (T)s.u, (T)s.v, (T)s.w, (T)side_h(g, s)};
}

// The precision-form face from the stencil's deviations d[s] = q[s] - q[2]
// (in T), the state of node m (qm) and the Roe inputs of nodes m and m+1.
template <class T>
__device__ __forceinline__ void face_flux_dev(const GasT<T>& gt, const GasT<double>& g,
                                              const T d[6][5], const double qm[5],
                                              const NodeT<T>& nl, const NodeT<T>& nr, double nx,
                                              double ny, double nz, double f[5]) {
  Basis<T> B;
  roe_basis_h(gt, nl.r, nl.u, nl.v, nl.w, nl.H, nr.r, nr.u, nr.v, nr.w, nr.H, (T)nx, (T)ny,
              (T)nz, B);
  T A[6], Bn[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    A[s] = B.K * (d[s][4] - (B.uu * d[s][1] + B.vv * d[s][2] + B.ww * d[s][3]) +
                  T(0.5) * B.q2 * d[s][0]);
    Bn[s] = B.o * (B.qn * d[s][0] - (B.nx * d[s][1] + B.ny * d[s][2] + B.nz * d[s][3]));
  }
  T WL[5], WR[5];
  weno5_pair(A[0] + Bn[0], A[1] + Bn[1], A[2] + Bn[2], A[3] + Bn[3], A[4] + Bn[4], A[5] + Bn[5],
             WL[0], WR[0]);
  weno5_pair(d[0][0] - T(2) * A[0], d[1][0] - T(2) * A[1], d[2][0] - T(2) * A[2],
             d[3][0] - T(2) * A[3], d[4][0] - T(2) * A[4], d[5][0] - T(2) * A[5], WL[1], WR[1]);
  weno5_pair(A[0] - Bn[0], A[1] - Bn[1], A[2] - Bn[2], A[3] - Bn[3], A[4] - Bn[4], A[5] - Bn[5],
             WL[2], WR[2]);
  T w[6];
#pragma unroll
  for (int s = 0; s < 6; s++)
    w[s] = -B.ql * d[s][0] + B.lx * d[s][1] + B.ly * d[s][2] + B.lz * d[s][3];
  weno5_pair(w[0], w[1], w[2], w[3], w[4], w[5], WL[3], WR[3]);
#pragma unroll
  for (int s = 0; s < 6; s++)
    w[s] = -B.qm * d[s][0] + B.mx * d[s][1] + B.my * d[s][2] + B.mz * d[s][3];
  weno5_pair(w[0], w[1], w[2], w[3], w[4], w[5], WL[4], WR[4]);
  // dq = R W (deviation), then q = q_m + dq in double
  double qL[5], qR[5];
  {
    const T* Ws[2] = {WL, WR};
    double* out[2] = {qL, qR};
#pragma unroll
    for (int side = 0; side < 2; side++) {
      const T* W = Ws[side];
      const T sm = W[0] + W[2];
      const T dl = B.c * (W[2] - W[0]);
      const T rho = sm + W[1];
      const T m1 = B.uu * rho + B.nx * dl + B.lx * W[3] + B.mx * W[4];
      const T m2 = B.vv * rho + B.ny * dl + B.ly * W[3] + B.my * W[4];
      const T m3 = B.ww * rho + B.nz * dl + B.lz * W[3] + B.mz * W[4];
      const T E = B.H * sm + T(0.5) * B.q2 * W[1] + B.qn * dl + B.ql * W[3] + B.qm * W[4];
      out[side][0] = qm[0] + (double)rho;
      out[side][1] = qm[1] + (double)m1;
      out[side][2] = qm[2] + (double)m2;
      out[side][3] = qm[3] + (double)m3;
      out[side][4] = qm[4] + (double)E;
    }
  }
  lf_flux(g, q_to_side(g, qR), q_to_side(g, qL), nx, ny, nz, f);
}

template <class T>
__device__ __forceinline__ void face_flux_mixed(const GasT<T>& gt, const GasT<double>& g,
                                                const double q[6][5], const Side<double>& sl,
                                                const Side<double>& sr, double nx, double ny,
                                                double nz, double f[5]) {
  T d[6][5];
#pragma unroll
  for (int s = 0; s < 6; s++)
#pragma unroll
    for (int c = 0; c < 5; c++) d[s][c] = (T)(q[s][c] - q[2][c]);
  face_flux_dev(gt, g, d, q[2], node_t<T>(g, sl), node_t<T>(g, sr), nx, ny, nz, f);
}

// Deviations from node m of a 6-node stencil from its five consecutive
// differences e[k] = q[k+1] - q[k] (in T): a sweep carries the differences
// from face to face instead of forming five double differences per face.
template <class T>
__device__ __forceinline__ void dev_from_diffs(const T e[5][5], T d[6][5]) {
#pragma unroll
  for (int c = 0; c < 5; c++) {
    d[2][c] = T(0);
    d[1][c] = -e[1][c];
    d[0][c] = -(e[0][c] + e[1][c]);
    d[3][c] = e[2][c];
    d[4][c] = e[2][c] + e[3][c];
    d[5][c] = d[4][c] + e[4][c];
  }
}

// face flux of the chosen arithmetic form: F = double (plain), float (precision)
template <class F>
__device__ __forceinline__ void face_flux_form(const GasT<double>& g, const double q[6][5],
                                               double nx, double ny, double nz, double f[5]) {
  if constexpr (sizeof(F) == sizeof(double)) {
    face_flux<double>(g, q, nx, ny, nz, f);
  } else {
    const GasT<F> gt{(F)g.gamma, (F)g.gam1, (F)g.gam1_inv};
    face_flux_mixed<F>(gt, g, q, q_to_side(g, q[2]), q_to_side(g, q[3]), nx, ny, nz, f);
  }
}

}  // namespace cbns
