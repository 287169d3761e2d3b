// learned.hpp -- device functions of the learned troubled-face indicator.
//
//   face_features   30 normalised characteristic stencil differences per face
//                   (float, on deviations from node m like face_flux_mixed):
//                   for each characteristic field c of W = L q (L = the
//                   Roe-averaged left eigenvectors, src/calc_eigenvectors_roe.cu)
//                   the six stencil values minus the face mean
//                   (w[2] + w[3]) / 2, over sqrt(sum_k (w[k+1] - w[k])^2 + 1e-6).
//   face_label      max |omega_k - d_k| over fields, sides and sub-stencils of
//                   the WENO5-JS weights (src/weno.cu:13-44), in double.
//   face_linear     the face flux with the linear weights d = (0.1, 0.6, 0.3):
//                   R (linear(L q)) = linear(q) because R L = I, so the
//                   reconstruction runs on the conservative state directly,
//                   with no characteristic basis.
//   mlp_logit_wave  30(32) -> 32 -> 32 -> 1 ReLU network for the 64 faces
//                   of one wavefront (one face per lane) on the bf16 matrix
//                   cores (v_mfma_f32_16x16x16_bf16, fp32 accumulate).
//
// MFMA operand layout (16x16x16, lane l, item v = 0..3):
//   A[i][k]: i = l % 16, k = 4 (l / 16) + v      (row-major weights)
//   B[k][j]: k = 4 (l / 16) + v, j = l % 16      (faces along j)
//   D[i][j]: i = 4 (l / 16) + v, j = l % 16
// With weights as A and faces as B, a layer's D tile is the next layer's B
// tile (hidden unit along k, face along j) without any data movement; only
// the input features pass through LDS (one 64-byte row per face).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "face_math.hpp"

namespace cbns {

constexpr int MLP_NFEAT = 30;   // features used
constexpr int MLP_IN = 32;      // padded input width (row stride)
constexpr int MLP_HID = 32;
constexpr int MLP_W1 = 0, MLP_B1 = 1024, MLP_W2 = 1056, MLP_B2 = 2080, MLP_W3 = 2112,
              MLP_B3 = 2144, MLP_NPARAM = 2145;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// packed bf16 operands: [tile t = 2 * row_tile + k_tile][lane][4]
struct MlpDev {
  const bf16x4* W1A;   // W1 as A
  const bf16x4* W2A;   // W2 as A
  const bf16x4* W2TA;  // W2^T as A (backward pass)
  const float* P;      // fp32 parameters (biases, output layer)
};

// round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ short f32_to_bf16(float x) {
  return __builtin_bit_cast(short, (__bf16)x);
}

// max(x, 0) on the bit pattern: negative floats (and -0) are negative
// integers, so one integer max is the ReLU
__device__ __forceinline__ float relu(float x) {
  return __int_as_float(max(__float_as_int(x), 0));
}

__device__ __forceinline__ f32x4 mfma_bf16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// characteristic values of the six stencil deviations d[s] = q[s] - q[2]
template <class T>
__device__ __forceinline__ void char_stencil(const Basis<T>& B, const T d[6][5], T w[5][6]) {
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const T A = B.K * (d[s][4] - (B.uu * d[s][1] + B.vv * d[s][2] + B.ww * d[s][3]) +
                       T(0.5) * B.q2 * d[s][0]);
    const T Bn = B.o * (B.qn * d[s][0] - (B.nx * d[s][1] + B.ny * d[s][2] + B.nz * d[s][3]));
    w[0][s] = A + Bn;
    w[1][s] = d[s][0] - T(2) * A;
    w[2][s] = A - Bn;
    w[3][s] = -B.ql * d[s][0] + B.lx * d[s][1] + B.ly * d[s][2] + B.lz * d[s][3];
    w[4][s] = -B.qm * d[s][0] + B.mx * d[s][1] + B.my * d[s][2] + B.mz * d[s][3];
  }
}

__device__ __forceinline__ void face_features(const GasT<float>& gt, const double q[6][5],
                                              const double pl[5], const double pr[5], double nx,
                                              double ny, double nz, float feat[MLP_IN]) {
  Basis<float> B;
  roe_basis(gt, (float)pl[0], (float)pl[1], (float)pl[2], (float)pl[3], (float)pl[4],
            (float)pr[0], (float)pr[1], (float)pr[2], (float)pr[3], (float)pr[4], (float)nx,
            (float)ny, (float)nz, B);
  float d[6][5];
#pragma unroll
  for (int s = 0; s < 6; s++)
#pragma unroll
    for (int c = 0; c < 5; c++) d[s][c] = (float)(q[s][c] - q[2][c]);
  float w[5][6];
  char_stencil(B, d, w);
#pragma unroll
  for (int c = 0; c < 5; c++) {
    float ss = 1.0e-6f;
#pragma unroll
    for (int k = 0; k < 5; k++) ss += (w[c][k + 1] - w[c][k]) * (w[c][k + 1] - w[c][k]);
    const float inv = fast_rsq(ss), mid = 0.5f * (w[c][2] + w[c][3]);
#pragma unroll
    for (int s = 0; s < 6; s++) feat[c * 6 + s] = (w[c][s] - mid) * inv;
  }
  feat[30] = feat[31] = 0.0f;
}

// max_k |omega_k - d_k| of one WENO5-JS evaluation on a..e
__device__ __forceinline__ double weight_dev(double a, double b, double c, double d, double e) {
  const double eps = 1.0e-6;
  const double d1 = a - 2.0 * b + c, d2 = a - 4.0 * b + 3.0 * c;
  const double d3 = b - 2.0 * c + d, d4 = b - d;
  const double d5 = c - 2.0 * d + e, d6 = 3.0 * c - 4.0 * d + e;
  const double IS1 = (13.0 / 12.0) * d1 * d1 + 0.25 * d2 * d2 + eps;
  const double IS2 = (13.0 / 12.0) * d3 * d3 + 0.25 * d4 * d4 + eps;
  const double IS3 = (13.0 / 12.0) * d5 * d5 + 0.25 * d6 * d6 + eps;
  const double a1 = 0.1 * fast_rcp(IS1 * IS1), a2 = 0.6 * fast_rcp(IS2 * IS2);
  const double a3 = 0.3 * fast_rcp(IS3 * IS3);
  const double is = fast_rcp(a1 + a2 + a3);
  return fmax(fabs(a1 * is - 0.1), fmax(fabs(a2 * is - 0.6), fabs(a3 * is - 0.3)));
}

__device__ __forceinline__ double face_label(const GasT<double>& g, const double q[6][5],
                                             const double pl[5], const double pr[5], double nx,
                                             double ny, double nz) {
  Basis<double> B;
  roe_basis(g, pl[0], pl[1], pl[2], pl[3], pl[4], pr[0], pr[1], pr[2], pr[3], pr[4], nx, ny, nz,
            B);
  double w[5][6];
  char_stencil(B, q, w);
  double m = 0.0;
#pragma unroll
  for (int c = 0; c < 5; c++) {
    m = fmax(m, weight_dev(w[c][0], w[c][1], w[c][2], w[c][3], w[c][4]));
    m = fmax(m, weight_dev(w[c][5], w[c][4], w[c][3], w[c][2], w[c][1]));
  }
  return m;
}

// Linear-weight face flux.  F = double: the plain arithmetic; F = float: the
// reconstruction on deviations from node m in float, the states and the flux
// in double (as face_flux_mixed).
//   left  = (2 q0 - 13 q1 + 47 q2 + 27 q3 - 3 q4) / 60
//   right = (2 q5 - 13 q4 + 47 q3 + 27 q2 - 3 q1) / 60
// one component: the left and right states of the face from its six
// stencil values v0..v5
template <class F>
__device__ __forceinline__ void linear_lr(double v0, double v1, double v2, double v3, double v4,
                                          double v5, double& L, double& R) {
  if constexpr (sizeof(F) == sizeof(double)) {
    L = (2.0 * v0 - 13.0 * v1 + 47.0 * v2 + 27.0 * v3 - 3.0 * v4) * (1.0 / 60.0);
    R = (2.0 * v5 - 13.0 * v4 + 47.0 * v3 + 27.0 * v2 - 3.0 * v1) * (1.0 / 60.0);
  } else {
    const F d0 = (F)(v0 - v2), d1 = (F)(v1 - v2), d3 = (F)(v3 - v2), d4 = (F)(v4 - v2);
    const F d5 = (F)(v5 - v2);
    L = v2 + (double)((F(2) * d0 - F(13) * d1 + F(27) * d3 - F(3) * d4) * F(1.0 / 60.0));
    R = v2 + (double)((F(2) * d5 - F(13) * d4 + F(47) * d3 - F(3) * d1) * F(1.0 / 60.0));
  }
}

template <class F>
__device__ __forceinline__ void face_linear(const GasT<double>& g, const double q[6][5], double nx,
                                            double ny, double nz, double f[5]) {
  double qL[5], qR[5];
#pragma unroll
  for (int c = 0; c < 5; c++)
    linear_lr<F>(q[0][c], q[1][c], q[2][c], q[3][c], q[4][c], q[5][c], qL[c], qR[c]);
  lf_flux(g, q_to_side(g, qR), q_to_side(g, qL), nx, ny, nz, f);
}

// Logit of this lane's face.  Every lane of the wavefront must call it
// (the matrix cores read all 64 lanes).  lds: 64 x 32 shorts for this wave.
__device__ __forceinline__ float mlp_logit_wave(const float x[MLP_IN], const MlpDev& w,
                                                short* lds) {
  const int l = threadIdx.x & 63;
  const int lq = l >> 4, lr = l & 15;
  {
    int* row = reinterpret_cast<int*>(lds + l * MLP_IN);
#pragma unroll
    for (int c = 0; c < MLP_IN / 2; c++)
      row[c] = (int)(uint16_t)f32_to_bf16(x[2 * c]) | ((int)f32_to_bf16(x[2 * c + 1]) << 16);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const bf16x4 a10 = w.W1A[0 * 64 + l], a11 = w.W1A[1 * 64 + l];
  const bf16x4 a12 = w.W1A[2 * 64 + l], a13 = w.W1A[3 * 64 + l];
  const bf16x4 a20 = w.W2A[0 * 64 + l], a21 = w.W2A[1 * 64 + l];
  const bf16x4 a22 = w.W2A[2 * 64 + l], a23 = w.W2A[3 * 64 + l];
  f32x4 c10, c11, c20, c21;  // bias rows 4 (l / 16) + v of the two 16-row tiles
  float w3[2][4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    c10[v] = w.P[MLP_B1 + 4 * lq + v];
    c11[v] = w.P[MLP_B1 + 16 + 4 * lq + v];
    c20[v] = w.P[MLP_B2 + 4 * lq + v];
    c21[v] = w.P[MLP_B2 + 16 + 4 * lq + v];
    w3[0][v] = w.P[MLP_W3 + 4 * lq + v];
    w3[1][v] = w.P[MLP_W3 + 16 + 4 * lq + v];
  }
  const float b3 = w.P[MLP_B3];
  float mine = 0.0f;
#pragma unroll
  for (int ft = 0; ft < 4; ft++) {
    const short* xr = lds + (16 * ft + lr) * MLP_IN + 4 * lq;
    const bf16x4 x0 = *reinterpret_cast<const bf16x4*>(xr);
    const bf16x4 x1 = *reinterpret_cast<const bf16x4*>(xr + 16);
    // the biases seed the accumulators (D = A B + C)
    f32x4 z0 = mfma_bf16(a10, x0, c10);
    z0 = mfma_bf16(a11, x1, z0);
    f32x4 z1 = mfma_bf16(a12, x0, c11);
    z1 = mfma_bf16(a13, x1, z1);
    bf16x4 h0, h1;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      h0[v] = f32_to_bf16(relu(z0[v]));
      h1[v] = f32_to_bf16(relu(z1[v]));
    }
    f32x4 y0 = mfma_bf16(a20, h0, c20);
    y0 = mfma_bf16(a21, h1, y0);
    f32x4 y1 = mfma_bf16(a22, h0, c21);
    y1 = mfma_bf16(a23, h1, y1);
    float part = 0.0f;
#pragma unroll
    for (int v = 0; v < 4; v++)
      part += w3[0][v] * relu(y0[v]) + w3[1][v] * relu(y1[v]);
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    if (lq == ft) mine = part + b3;
  }
  return mine;
}

}  // namespace cbns
