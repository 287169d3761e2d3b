// face_math.hpp -- per-face inviscid flux: Roe-averaged characteristic basis,
// characteristic WENO5-JS reconstruction of the conservative state on both
// sides, Lax-Friedrichs-type flux.  Method of src/rec_riemann_kernel.cu:44-89
// (calc_eigenvectors_roe, rec_weno5 x2, lf_flux), restructured for gfx950:
//   - the six stencil nodes are projected once (the two sides share four);
//   - L (left eigenvectors) is applied in closed form: rows 0..2 share two
//     dot products, rows 3..4 are the tangent projections;
//   - R (right eigenvectors) is applied in closed form;
//   - each WENO5 evaluation uses one division (weights multiplied through by
//     IS1^2 IS2^2 IS3^2) instead of seven;
//   - 1/x is formed once and multiplied where the reference divides by the
//     same x repeatedly.
// The result equals the reference formula up to rounding (the stage parity
// bar is 1e-12 relative; tests/test_gpu_parity.py).
#pragma once

#include <hip/hip_runtime.h>

namespace cbns {

template <class T>
struct GasT {
  T gamma, gam1, gam1_inv;
};

// WENO5-JS at the right face of the centre point c, src/weno.cu:13-44.
// Weights 0.1/IS1^2 : 0.6/IS2^2 : 0.3/IS3^2 are scaled by IS1^2 IS2^2 IS3^2.
template <class T>
__device__ __forceinline__ T weno5(T a, T b, T c, T d, T e) {
  const T eps = T(1.0e-6);
  const T d1 = a - T(2) * b + c;
  const T d2 = a - T(4) * b + T(3) * c;
  const T d3 = b - T(2) * c + d;
  const T d4 = b - d;
  const T d5 = c - T(2) * d + e;
  const T d6 = T(3) * c - T(4) * d + e;
  const T IS1 = T(13.0 / 12.0) * d1 * d1 + T(0.25) * d2 * d2 + eps;
  const T IS2 = T(13.0 / 12.0) * d3 * d3 + T(0.25) * d4 * d4 + eps;
  const T IS3 = T(13.0 / 12.0) * d5 * d5 + T(0.25) * d6 * d6 + eps;
  const T s1 = IS1 * IS1, s2 = IS2 * IS2, s3 = IS3 * IS3;
  const T w1 = T(0.1) * (s2 * s3);
  const T w2 = T(0.6) * (s1 * s3);
  const T w3 = T(0.3) * (s1 * s2);
  const T f1 = T(2) * a - T(7) * b + T(11) * c;
  const T f2 = -b + T(5) * c + T(2) * d;
  const T f3 = T(2) * c + T(5) * d - e;
  return (w1 * f1 + w2 * f2 + w3 * f3) / (T(6) * (w1 + w2 + w3));
}

// Roe-averaged characteristic basis of a face, src/calc_eigenvectors_roe.cu:28-150.
template <class T>
struct Basis {
  T nx, ny, nz, lx, ly, lz, mx, my, mz;  // orthonormal (n normal, l,m tangents)
  T uu, vv, ww, H, c, q2, qn, ql, qm;    // Roe state
  T K, o;                                // gam1/(2c^2), 0.5/c
};

template <class T>
__device__ __forceinline__ void roe_basis(const GasT<T>& g, T rL, T uL, T vL, T wL, T pL,
                                          T rR, T uR, T vR, T wR, T pR, T n_x, T n_y, T n_z,
                                          Basis<T>& B) {
  T l_x = T(0), l_y = n_z, l_z = -n_y;
  T m_x, m_y, m_z;
  T l2 = l_y * l_y + l_z * l_z;
  T labs = sqrt(l2);
  if (labs > T(1.0e-8)) {
    m_x = -n_y * n_y - n_z * n_z;
    m_y = n_x * n_y;
    m_z = n_x * n_z;
  } else {
    l_x = -n_z; l_y = T(0); l_z = n_x;
    m_x = n_x * n_y;
    m_y = -n_x * n_x - n_z * n_z;
    m_z = n_y * n_z;
    labs = sqrt(l_x * l_x + l_z * l_z);
  }
  const T inv_n = T(1) / sqrt(n_x * n_x + n_y * n_y + n_z * n_z);
  B.nx = n_x * inv_n; B.ny = n_y * inv_n; B.nz = n_z * inv_n;
  const T inv_l = T(1) / labs;
  B.lx = l_x * inv_l; B.ly = l_y * inv_l; B.lz = l_z * inv_l;
  const T inv_m = T(1) / sqrt(m_x * m_x + m_y * m_y + m_z * m_z);
  B.mx = m_x * inv_m; B.my = m_y * inv_m; B.mz = m_z * inv_m;

  // H = c^2/gam1 + |u|^2/2 with c^2 = gamma p / rho
  const T HL = g.gamma * pL / rL * g.gam1_inv + T(0.5) * (uL * uL + vL * vL + wL * wL);
  const T HR = g.gamma * pR / rR * g.gam1_inv + T(0.5) * (uR * uR + vR * vR + wR * wR);
  const T ratio = sqrt(rR / rL);
  const T inv1r = T(1) / (T(1) + ratio);
  B.uu = (uL + ratio * uR) * inv1r;
  B.vv = (vL + ratio * vR) * inv1r;
  B.ww = (wL + ratio * wR) * inv1r;
  B.H = (HL + ratio * HR) * inv1r;
  B.q2 = B.uu * B.uu + B.vv * B.vv + B.ww * B.ww;
  B.c = sqrt(g.gam1 * (B.H - T(0.5) * B.q2));
  B.qn = B.uu * B.nx + B.vv * B.ny + B.ww * B.nz;
  B.ql = B.uu * B.lx + B.vv * B.ly + B.ww * B.lz;
  B.qm = B.uu * B.mx + B.vv * B.my + B.ww * B.mz;
  const T inv_c = T(1) / B.c;
  B.o = T(0.5) * inv_c;
  B.K = T(0.5) * g.gam1 * inv_c * inv_c;
}

// W = L q for one node (q = rho, rho u, rho v, rho w, E).
// Rows 0 and 2 are A +/- Bn, row 1 is rho - 2A with
//   A = K (E - u.m + q2/2 rho),  Bn = o (qn rho - n.m).
template <class T>
__device__ __forceinline__ void char_project(const Basis<T>& B, const T q[5], T W[5]) {
  const T A = B.K * (q[4] - (B.uu * q[1] + B.vv * q[2] + B.ww * q[3]) + T(0.5) * B.q2 * q[0]);
  const T Bn = B.o * (B.qn * q[0] - (B.nx * q[1] + B.ny * q[2] + B.nz * q[3]));
  W[0] = A + Bn;
  W[1] = q[0] - T(2) * A;
  W[2] = A - Bn;
  W[3] = -B.ql * q[0] + B.lx * q[1] + B.ly * q[2] + B.lz * q[3];
  W[4] = -B.qm * q[0] + B.mx * q[1] + B.my * q[2] + B.mz * q[3];
}

// q = R W, then to primitives (src/weno.cu:106-118)
template <class T>
__device__ __forceinline__ void char_to_prim(const GasT<T>& g, const Basis<T>& B, const T W[5],
                                             T& r, T& u, T& v, T& w, T& p) {
  const T s = W[0] + W[2];
  const T dl = B.c * (W[2] - W[0]);
  const T rho = s + W[1];
  const T m1 = B.uu * rho + B.nx * dl + B.lx * W[3] + B.mx * W[4];
  const T m2 = B.vv * rho + B.ny * dl + B.ly * W[3] + B.my * W[4];
  const T m3 = B.ww * rho + B.nz * dl + B.lz * W[3] + B.mz * W[4];
  const T E = B.H * s + T(0.5) * B.q2 * W[1] + B.qn * dl + B.ql * W[3] + B.qm * W[4];
  const T ir = T(1) / rho;
  r = rho;
  u = m1 * ir;
  v = m2 * ir;
  w = m3 * ir;
  p = g.gam1 * (E - T(0.5) * rho * (u * u + v * v + w * w));
}

// src/lf_flux.cu:22-78 (dissipation speed |sum_d max_LR(|u_d|+c) n_d|)
template <class T>
__device__ __forceinline__ void lf_flux(const GasT<T>& g, T rR, T uR, T vR, T wR, T pR, T rL,
                                        T uL, T vL, T wL, T pL, T nx, T ny, T nz, T f[5]) {
  const T cR = sqrt(g.gamma * pR / rR);
  const T kR = T(0.5) * (uR * uR + vR * vR + wR * wR);
  const T HR = cR * cR * g.gam1_inv + kR;
  const T EtR = pR * g.gam1_inv + rR * kR;
  const T cL = sqrt(g.gamma * pL / rL);
  const T kL = T(0.5) * (uL * uL + vL * vL + wL * wL);
  const T HL = cL * cL * g.gam1_inv + kL;
  const T EtL = pL * g.gam1_inv + rL * kL;
  const T qR = uR * nx + vR * ny + wR * nz;
  const T qL = uL * nx + vL * ny + wL * nz;
  const T s0 = fmax(fabs(uR) + cR, fabs(uL) + cL);
  const T s1 = fmax(fabs(vR) + cR, fabs(vL) + cL);
  const T s2 = fmax(fabs(wR) + cR, fabs(wL) + cL);
  const T a = fabs(s0 * nx + s1 * ny + s2 * nz);
  const T mL = rL * qL, mR = rR * qR;
  f[0] = T(0.5) * (mL + mR - a * (rR - rL));
  f[1] = T(0.5) * (mL * uL + pL * nx + mR * uR + pR * nx - a * (rR * uR - rL * uL));
  f[2] = T(0.5) * (mL * vL + pL * ny + mR * vR + pR * ny - a * (rR * vR - rL * vL));
  f[3] = T(0.5) * (mL * wL + pL * nz + mR * wR + pR * nz - a * (rR * wR - rL * wL));
  f[4] = T(0.5) * (mL * HL + mR * HR - a * (EtR - EtL));
}

// conservative -> primitive (src/primitive_bc_kernel.cu:63-67)
template <class T>
__device__ __forceinline__ void q_to_prim(const GasT<T>& g, const T q[5], T& r, T& u, T& v, T& w,
                                          T& p) {
  r = q[0];
  const T ir = T(1) / r;
  u = q[1] * ir;
  v = q[2] * ir;
  w = q[3] * ir;
  p = g.gam1 * (q[4] - T(0.5) * r * (u * u + v * v + w * w));
}

// Full face: q[s][c] holds nodes m-2..m+3 (face between m and m+1); pl, pr
// are the primitives of nodes m and m+1; (nx,ny,nz) the unnormalised face
// vector 0.5 (J grad xi |_m + J grad xi |_{m+1}).
template <class T>
__device__ __forceinline__ void face_flux(const GasT<T>& g, const T q[6][5], const T pl[5],
                                          const T pr[5], T nx, T ny, T nz, T f[5]) {
  Basis<T> B;
  roe_basis(g, pl[0], pl[1], pl[2], pl[3], pl[4], pr[0], pr[1], pr[2], pr[3], pr[4], nx, ny, nz,
            B);
  T W[6][5];
#pragma unroll
  for (int s = 0; s < 6; s++) char_project(B, q[s], W[s]);
  T WL[5], WR[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    WL[m] = weno5(W[0][m], W[1][m], W[2][m], W[3][m], W[4][m]);
    WR[m] = weno5(W[5][m], W[4][m], W[3][m], W[2][m], W[1][m]);
  }
  T rL, uL, vL, wL, pL, rR, uR, vR, wR, pR;
  char_to_prim(g, B, WL, rL, uL, vL, wL, pL);
  char_to_prim(g, B, WR, rR, uR, vR, wR, pR);
  lf_flux(g, rR, uR, vR, wR, pR, rL, uL, vL, wL, pL, nx, ny, nz, f);
}

}  // namespace cbns
