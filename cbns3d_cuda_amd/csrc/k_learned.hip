// k_learned.hip -- the learned troubled-face indicator on the GPU.
//
//   k_face_samples<AX>  training data the plain solver emits: per face of
//                       axis AX, the 32 indicator features (float) and the
//                       label max |omega_k - d_k| of its FP64 WENO5 weights
//   k_indicator<AX>     inference for every face the flux kernels may read
//                       (their face range plus one on each side for the
//                       one-cell dilation): features, MLP on the matrix
//                       cores, one flag byte per face
//   k_mlp_forward       logits of a feature batch
//   k_mlp_grad          weighted logistic loss and its gradient: forward and
//                       backward on the matrix cores; weight gradients are
//                       16x16x16 products over faces after an LDS transpose
//   k_mlp_adam          Adam update of the fp32 parameters
//   k_mlp_pack          fp32 parameters -> bf16 operand tiles
#include <cmath>
#include <cstring>
#include <vector>

#include "cbns_internal.hpp"
#include "learned.hpp"

namespace cbns {

namespace {

template <int AX>
__device__ __forceinline__ int axis_stride(const Consts& c) {
  return AX == 0 ? 1 : (AX == 1 ? c.PX : c.PXY);
}

// face (m, m+1) along AX at node a: stencil, node primitives, face vector
template <int AX>
__device__ __forceinline__ void load_face(const Consts& c, const double* __restrict__ Q,
                                          const double* __restrict__ met, int a, double q[6][5],
                                          double pl[5], double pr[5], double n[3]) {
  const long S = c.S;
  const int st = axis_stride<AX>(c);
#pragma unroll
  for (int s = 0; s < 6; s++)
#pragma unroll
    for (int v = 0; v < 5; v++) q[s][v] = Q[v * S + a + (s - 2) * st];
  const GasT<double> g{c.gamma, c.gam1, c.gam1_inv};
  q_to_prim(g, q[2], pl[0], pl[1], pl[2], pl[3], pl[4]);
  q_to_prim(g, q[3], pr[0], pr[1], pr[2], pr[3], pr[4]);
  const double Ja = met[JAC * S + a], Jb = met[JAC * S + a + st];
#pragma unroll
  for (int d = 0; d < 3; d++)
    n[d] = 0.5 * (met[(3 * AX + d) * S + a] * Ja + met[(3 * AX + d) * S + a + st] * Jb);
}

template <int AX>
__global__ void __launch_bounds__(256) k_face_samples(Consts c, Box box,
                                                      const double* __restrict__ Q,
                                                      const double* __restrict__ met,
                                                      float* __restrict__ feat,
                                                      float* __restrict__ label) {
  int i, j, k;
  if (!box_node3(box, i, j, k)) return;
  const long smp = (long)(i - box.lo[0]) +
                   (long)box.n[0] * ((long)(j - box.lo[1]) + (long)box.n[1] * (k - box.lo[2]));
  double q[6][5], pl[5], pr[5], n[3];
  load_face<AX>(c, Q, met, idx(c, i, j, k), q, pl, pr, n);
  const GasT<float> gt{(float)c.gamma, (float)c.gam1, (float)c.gam1_inv};
  float x[MLP_IN];
  face_features(gt, q, pl, pr, n[0], n[1], n[2], x);
  float4* out = reinterpret_cast<float4*>(feat + smp * MLP_IN);
#pragma unroll
  for (int v = 0; v < MLP_IN / 4; v++)
    out[v] = make_float4(x[4 * v], x[4 * v + 1], x[4 * v + 2], x[4 * v + 3]);
  const GasT<double> g{c.gamma, c.gam1, c.gam1_inv};
  label[smp] = (float)face_label(g, q, pl, pr, n[0], n[1], n[2]);
}

// Every thread of the block reaches the MLP (out-of-box lanes with zero
// features), so the matrix cores see full wavefronts.
template <int AX>
__global__ void __launch_bounds__(256) k_indicator(Consts c, Box box, const double* __restrict__ Q,
                                                   const double* __restrict__ met, MlpDev w,
                                                   float thresh, uint8_t* __restrict__ flags,
                                                   unsigned* __restrict__ count) {
  __shared__ short lds[4][64 * MLP_IN];
  int i = 0, j = 0, k = 0;
  const bool valid = box_node3(box, i, j, k);
  float x[MLP_IN];
  int a = 0;
  if (valid) {
    a = idx(c, i, j, k);
    double q[6][5], pl[5], pr[5], n[3];
    load_face<AX>(c, Q, met, a, q, pl, pr, n);
    const GasT<float> gt{(float)c.gamma, (float)c.gam1, (float)c.gam1_inv};
    face_features(gt, q, pl, pr, n[0], n[1], n[2], x);
  } else {
#pragma unroll
    for (int v = 0; v < MLP_IN; v++) x[v] = 0.0f;
  }
  const float s = mlp_logit_wave(x, w, lds[threadIdx.y]);
  const bool flag = valid && s > thresh;
  if (valid) flags[a] = flag ? 1 : 0;
  if (count) {
    // count faces of the flux kernels' range, not the one-face halo
    const int pa = AX == 0 ? i : (AX == 1 ? j : k);
    const bool inner = pa > box.lo[AX] && pa < box.lo[AX] + box.n[AX] - 1;
    const unsigned long long bal = __ballot(flag && inner);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(count, (unsigned)__popcll(bal));
  }
}

__global__ void __launch_bounds__(256) k_mlp_forward(MlpDev w, const float* __restrict__ X,
                                                     long n, float* __restrict__ logit) {
  __shared__ short lds[4][64 * MLP_IN];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const long smp = ((long)blockIdx.x * 4 + wv) * 64 + l;
  float x[MLP_IN];
  if (smp < n) {
    const float4* row = reinterpret_cast<const float4*>(X + smp * MLP_IN);
#pragma unroll
    for (int v = 0; v < MLP_IN / 4; v++) {
      const float4 t = row[v];
      x[4 * v] = t.x; x[4 * v + 1] = t.y; x[4 * v + 2] = t.z; x[4 * v + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int v = 0; v < MLP_IN; v++) x[v] = 0.0f;
  }
  const float s = mlp_logit_wave(x, w, lds[wv]);
  if (smp < n) logit[smp] = s;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x4 to_bf16x4(float4 t) {
  bf16x4 r;
  r[0] = f32_to_bf16(t.x); r[1] = f32_to_bf16(t.y);
  r[2] = f32_to_bf16(t.z); r[3] = f32_to_bf16(t.w);
  return r;
}

__device__ __forceinline__ float softplus(float s) {
  return s > 0.0f ? s + log1pf(expf(-s)) : log1pf(expf(s));
}

// One wavefront takes 64 samples per iteration (four 16-face tiles).
// grad (MLP_NPARAM floats) and loss_sum are accumulated atomically.
__global__ void __launch_bounds__(256) k_mlp_grad(MlpDev w, const float* __restrict__ X,
                                                  const float* __restrict__ Y, long n,
                                                  float w_pos, float inv_n, int iters,
                                                  float* __restrict__ grad,
                                                  float* __restrict__ loss_sum) {
  __shared__ short tr[4][4 * 32 * 16];  // per wave: dz2, h1, dz1, x as [row][16 faces]
  __shared__ float gacc[MLP_NPARAM + 1];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, lq = l >> 4, lr = l & 15;
  for (int t = threadIdx.x; t <= MLP_NPARAM; t += blockDim.x) gacc[t] = 0.0f;
  __syncthreads();
  bf16x4 a1[4], a2[4], a2t[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    a1[t] = w.W1A[t * 64 + l];
    a2[t] = w.W2A[t * 64 + l];
    a2t[t] = w.W2TA[t * 64 + l];
  }
  f32x4 cb1[2], cb2[2];
  float w3[2][4];
#pragma unroll
  for (int ht = 0; ht < 2; ht++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = 16 * ht + 4 * lq + v;
      cb1[ht][v] = w.P[MLP_B1 + r];
      cb2[ht][v] = w.P[MLP_B2 + r];
      w3[ht][v] = w.P[MLP_W3 + r];
    }
  const float b3 = w.P[MLP_B3];
  const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  f32x4 gW1[2][2] = {{zero, zero}, {zero, zero}}, gW2[2][2] = {{zero, zero}, {zero, zero}};
  float gb1[2][4] = {}, gb2[2][4] = {}, gw3[2][4] = {};
  float gb3 = 0.0f, lsum = 0.0f;
  short* T = tr[wv];
  for (int it = 0; it < iters; it++) {
    const long chunk = ((long)blockIdx.x * 4 + wv) + (long)it * gridDim.x * 4;
    for (int ft = 0; ft < 4; ft++) {
      const long smp = chunk * 64 + 16 * ft + lr;
      const bool valid = smp < n;
      bf16x4 x[2];
      if (valid) {
        x[0] = to_bf16x4(*reinterpret_cast<const float4*>(X + smp * MLP_IN + 4 * lq));
        x[1] = to_bf16x4(*reinterpret_cast<const float4*>(X + smp * MLP_IN + 16 + 4 * lq));
      } else {
        x[0] = bf16x4{0, 0, 0, 0};
        x[1] = bf16x4{0, 0, 0, 0};
      }
      // forward
      f32x4 z[2], y[2];
      bf16x4 h[2];
#pragma unroll
      for (int ht = 0; ht < 2; ht++) {
        z[ht] = mfma_bf16(a1[2 * ht], x[0], cb1[ht]);  // biases seed the accumulators
        z[ht] = mfma_bf16(a1[2 * ht + 1], x[1], z[ht]);
#pragma unroll
        for (int v = 0; v < 4; v++) h[ht][v] = f32_to_bf16(relu(z[ht][v]));
      }
#pragma unroll
      for (int ht = 0; ht < 2; ht++) {
        y[ht] = mfma_bf16(a2[2 * ht], h[0], cb2[ht]);
        y[ht] = mfma_bf16(a2[2 * ht + 1], h[1], y[ht]);
      }
      float part = 0.0f;
#pragma unroll
      for (int ht = 0; ht < 2; ht++)
#pragma unroll
        for (int v = 0; v < 4; v++) part += w3[ht][v] * relu(y[ht][v]);
      part += __shfl_xor(part, 16);
      part += __shfl_xor(part, 32);
      const float s = part + b3;
      // loss and output-layer gradient
      const float yv = valid ? Y[smp] : 0.0f;
      const float sg = 1.0f / (1.0f + expf(-s));
      const float ds = valid ? (yv * w_pos * (sg - 1.0f) + (1.0f - yv) * sg) * inv_n : 0.0f;
      if (valid && lq == 0) {
        lsum += yv * w_pos * softplus(-s) + (1.0f - yv) * softplus(s);
        gb3 += ds;
      }
      bf16x4 dz2b[2];
#pragma unroll
      for (int ht = 0; ht < 2; ht++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const float act = y[ht][v] > 0.0f;
          gw3[ht][v] += ds * relu(y[ht][v]);
          const float dz2 = act ? ds * w3[ht][v] : 0.0f;
          gb2[ht][v] += dz2;
          dz2b[ht][v] = f32_to_bf16(dz2);
        }
      // dh1 = W2^T dz2, dz1 = dh1 (z1 > 0)
      bf16x4 dz1b[2];
#pragma unroll
      for (int it2 = 0; it2 < 2; it2++) {
        f32x4 d = mfma_bf16(a2t[2 * it2], dz2b[0], zero);
        d = mfma_bf16(a2t[2 * it2 + 1], dz2b[1], d);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const float dz1 = z[it2][v] > 0.0f ? d[v] : 0.0f;
          gb1[it2][v] += dz1;
          dz1b[it2][v] = f32_to_bf16(dz1);
        }
      }
      // transposes through LDS: [row][face]
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = 16 * t + 4 * lq + v;
          T[0 * 512 + r * 16 + lr] = dz2b[t][v];
          T[1 * 512 + r * 16 + lr] = h[t][v];
          T[2 * 512 + r * 16 + lr] = dz1b[t][v];
          T[3 * 512 + r * 16 + lr] = x[t][v];
        }
      wave_sync_lds();
      // gW2[h][i] += sum_f dz2[h][f] h1[i][f];  gW1[i][k] += sum_f dz1[i][f] x[k][f]
#pragma unroll
      for (int r = 0; r < 2; r++)
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
          const bf16x4 A2 = *reinterpret_cast<const bf16x4*>(T + 0 * 512 + (16 * r + lr) * 16 + 4 * lq);
          const bf16x4 B2 = *reinterpret_cast<const bf16x4*>(T + 1 * 512 + (16 * cc + lr) * 16 + 4 * lq);
          gW2[r][cc] = mfma_bf16(A2, B2, gW2[r][cc]);
          const bf16x4 A1 = *reinterpret_cast<const bf16x4*>(T + 2 * 512 + (16 * r + lr) * 16 + 4 * lq);
          const bf16x4 B1 = *reinterpret_cast<const bf16x4*>(T + 3 * 512 + (16 * cc + lr) * 16 + 4 * lq);
          gW1[r][cc] = mfma_bf16(A1, B1, gW1[r][cc]);
        }
      wave_sync_lds();
    }
  }
  // block reduction in LDS, then one global atomic per parameter
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int cc = 0; cc < 2; cc++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = 16 * r + 4 * lq + v, col = 16 * cc + lr;
        atomicAdd(&gacc[MLP_W2 + row * MLP_HID + col], gW2[r][cc][v]);
        atomicAdd(&gacc[MLP_W1 + row * MLP_IN + col], gW1[r][cc][v]);
      }
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = 16 * t + 4 * lq + v;
      atomicAdd(&gacc[MLP_B1 + row], gb1[t][v]);
      atomicAdd(&gacc[MLP_B2 + row], gb2[t][v]);
      atomicAdd(&gacc[MLP_W3 + row], gw3[t][v]);
    }
  if (lq == 0) {
    atomicAdd(&gacc[MLP_B3], gb3);
    atomicAdd(&gacc[MLP_NPARAM], lsum);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < MLP_NPARAM; t += blockDim.x) atomicAdd(&grad[t], gacc[t]);
  if (threadIdx.x == 0) atomicAdd(loss_sum, gacc[MLP_NPARAM]);
}

// Adam on the fp32 parameters; W1's two padding input columns stay zero.
__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= MLP_NPARAM) return;
  if (t < MLP_B1 && (t % MLP_IN) >= MLP_NFEAT) return;
  const float g = grad[t];
  const float mt = b1 * m[t] + (1.0f - b1) * g;
  const float vt = b2 * v[t] + (1.0f - b2) * g * g;
  m[t] = mt;
  v[t] = vt;
  P[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// A-operand tiles t = 2 * row_tile + k_tile: lane l item e holds
// M[16 row_tile + l % 16][16 k_tile + 4 (l / 16) + e].
__global__ void k_mlp_pack(const float* __restrict__ P, bf16x4* __restrict__ W1A,
                           bf16x4* __restrict__ W2A, bf16x4* __restrict__ W2TA) {
  const int t = threadIdx.x;  // 256 threads: 4 tiles x 64 lanes
  const int tile = t >> 6, l = t & 63, rt = tile >> 1, kt = tile & 1;
  bf16x4 a, b, c;
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int row = 16 * rt + (l & 15), col = 16 * kt + 4 * (l >> 4) + e;
    a[e] = f32_to_bf16(P[MLP_W1 + row * MLP_IN + col]);
    b[e] = f32_to_bf16(P[MLP_W2 + row * MLP_HID + col]);
    c[e] = f32_to_bf16(P[MLP_W2 + col * MLP_HID + row]);
  }
  W1A[t] = a;
  W2A[t] = b;
  W2TA[t] = c;
}

__global__ void k_gather_rows(const float* __restrict__ src, const int64_t* __restrict__ idx,
                              int64_t n, int row, float* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * row) return;
  const int64_t r = t / row, c = t - r * row;
  dst[t] = src[idx[r] * row + c];
}

__global__ void k_zero_counts(unsigned* cnt) {
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0u;
}

}  // namespace

// ---- block side -----------------------------------------------------------

// faces of axis `a` the flux kernels compute: -1..N-1 (periodic) or 0..N-2
long face_count(const cbns_block* b, int axis) { return face_box(b->c, axis, 0).count(); }

int launch_face_samples(cbns_block* b, int axis, const double* Q, float* feat, float* label) {
  const Box box = face_box(b->c, axis, 0);
  if (box.count() == 0) return CBNS_OK;
  switch (axis) {
    case 0: k_face_samples<0><<<box_grid(box), box_block(), 0, b->stream>>>(b->c, box, Q, b->met, feat, label); break;
    case 1: k_face_samples<1><<<box_grid(box), box_block(), 0, b->stream>>>(b->c, box, Q, b->met, feat, label); break;
    default: k_face_samples<2><<<box_grid(box), box_block(), 0, b->stream>>>(b->c, box, Q, b->met, feat, label); break;
  }
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

int launch_indicator(cbns_block* b, const double* Q) {
  hipStream_t st = b->stream;
  const MlpDev w{reinterpret_cast<const bf16x4*>(b->ind_pack),
                 reinterpret_cast<const bf16x4*>(b->ind_pack) + 256, nullptr, b->ind_P};
  unsigned* cnt = b->ind_count_on ? b->ind_count : nullptr;
  if (cnt) {
    k_zero_counts<<<1, 64, 0, st>>>(cnt);
    CBNS_HIP(b, hipGetLastError());
  }
  ProfScope ps(b, CBNS_K_INDICATOR, st);
  for (int a = 0; a < 3; a++) {
    const Box box = face_box(b->c, a, 1);
    if (box.count() == 0) continue;
    uint8_t* fl = b->flags + (long)a * b->c.S;
    switch (a) {
      case 0: k_indicator<0><<<box_grid(box), box_block(), 0, st>>>(b->c, box, Q, b->met, w, b->ind_thresh, fl, cnt ? cnt + 0 : nullptr); break;
      case 1: k_indicator<1><<<box_grid(box), box_block(), 0, st>>>(b->c, box, Q, b->met, w, b->ind_thresh, fl, cnt ? cnt + 1 : nullptr); break;
      default: k_indicator<2><<<box_grid(box), box_block(), 0, st>>>(b->c, box, Q, b->met, w, b->ind_thresh, fl, cnt ? cnt + 2 : nullptr); break;
    }
    CBNS_HIP(b, hipGetLastError());
  }
  b->flags_valid = true;
  return CBNS_OK;
}

int upload_indicator(cbns_block* b, const float* params) {
  CBNS_HIP(b, hipMemcpyAsync(b->ind_P, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice,
                             b->stream));
  bf16x4* pk = reinterpret_cast<bf16x4*>(b->ind_pack);
  k_mlp_pack<<<1, 256, 0, b->stream>>>(b->ind_P, pk, pk + 256, pk + 512);
  CBNS_HIP(b, hipGetLastError());
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  return CBNS_OK;
}

}  // namespace cbns

// ---- MLP object (training) -----------------------------------------------

struct cbns_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float* P = nullptr;     // MLP_NPARAM parameters
  float* grad = nullptr;  // MLP_NPARAM gradient + 1 loss slot
  float* m = nullptr;
  float* v = nullptr;
  cbns::bf16x4* pack = nullptr;  // 3 x 256 operand tiles
  int t = 0;                     // Adam step count
  std::string err;
};

namespace {

int mlp_fail(cbns_mlp* m, hipError_t e, const char* what) {
  m->err = std::string(what) + ": " + hipGetErrorString(e);
  return CBNS_ERR_HIP;
}

#define MLP_HIP(m, call)                                  \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return mlp_fail((m), e_, #call); \
  } while (0)

cbns::MlpDev mlp_dev(const cbns_mlp* m) {
  return cbns::MlpDev{m->pack, m->pack + 256, m->pack + 512, m->P};
}

int mlp_repack(cbns_mlp* m) {
  cbns::k_mlp_pack<<<1, 256, 0, m->stream>>>(m->P, m->pack, m->pack + 256, m->pack + 512);
  MLP_HIP(m, hipGetLastError());
  return CBNS_OK;
}

}  // namespace

extern "C" {

int cbns_mlp_create(cbns_mlp** out) {
  if (!out) return CBNS_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CBNS_ERR_NO_DEVICE;
  cbns_mlp* m = new cbns_mlp();
  auto bad = [&](hipError_t e) {
    if (e == hipSuccess) return false;
    cbns_mlp_destroy(m);
    return true;
  };
  if (bad(hipGetDevice(&m->device)) || bad(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) ||
      bad(hipMalloc(&m->P, cbns::MLP_NPARAM * sizeof(float))) ||
      bad(hipMalloc(&m->grad, (cbns::MLP_NPARAM + 1) * sizeof(float))) ||
      bad(hipMalloc(&m->m, cbns::MLP_NPARAM * sizeof(float))) ||
      bad(hipMalloc(&m->v, cbns::MLP_NPARAM * sizeof(float))) ||
      bad(hipMalloc(&m->pack, 768 * sizeof(cbns::bf16x4))) ||
      bad(hipMemset(m->P, 0, cbns::MLP_NPARAM * sizeof(float))) ||
      bad(hipMemset(m->m, 0, cbns::MLP_NPARAM * sizeof(float))) ||
      bad(hipMemset(m->v, 0, cbns::MLP_NPARAM * sizeof(float))))
    return CBNS_ERR_HIP;
  if (mlp_repack(m) || hipStreamSynchronize(m->stream) != hipSuccess) {
    cbns_mlp_destroy(m);
    return CBNS_ERR_HIP;
  }
  *out = m;
  return CBNS_OK;
}

void cbns_mlp_destroy(cbns_mlp* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : {(void*)m->P, (void*)m->grad, (void*)m->m, (void*)m->v, (void*)m->pack})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char* cbns_mlp_last_error(const cbns_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

int cbns_mlp_set_params(cbns_mlp* m, const float* params) {
  if (!m || !params) return CBNS_ERR_ARG;
  std::vector<float> p(params, params + cbns::MLP_NPARAM);
  for (int r = 0; r < cbns::MLP_HID; r++)
    for (int c = cbns::MLP_NFEAT; c < cbns::MLP_IN; c++) p[cbns::MLP_W1 + r * cbns::MLP_IN + c] = 0.0f;
  MLP_HIP(m, hipMemcpyAsync(m->P, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(m, hipMemsetAsync(m->m, 0, cbns::MLP_NPARAM * sizeof(float), m->stream));
  MLP_HIP(m, hipMemsetAsync(m->v, 0, cbns::MLP_NPARAM * sizeof(float), m->stream));
  m->t = 0;
  if (int rc = mlp_repack(m)) return rc;
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return CBNS_OK;
}

int cbns_mlp_get_params(cbns_mlp* m, float* params) {
  if (!m || !params) return CBNS_ERR_ARG;
  MLP_HIP(m, hipMemcpyAsync(params, m->P, cbns::MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return CBNS_OK;
}

int cbns_mlp_forward(cbns_mlp* m, const float* feat, int64_t n, float* logit) {
  if (!m || n < 0 || (n > 0 && (!feat || !logit))) return CBNS_ERR_ARG;
  if (n == 0) return CBNS_OK;
  const long blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffL) return CBNS_ERR_ARG;
  cbns::k_mlp_forward<<<(unsigned)blocks, 256, 0, m->stream>>>(mlp_dev(m), feat, n, logit);
  MLP_HIP(m, hipGetLastError());
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return CBNS_OK;
}

int cbns_mlp_loss_grad(cbns_mlp* m, const float* feat, const float* target, int64_t n, float w_pos,
                       double* loss, float* grad) {
  if (!m || n <= 0 || !feat || !target) return CBNS_ERR_ARG;
  const long chunks = (n + 63) / 64;
  const long waves = std::min<long>(chunks, 4L * 1024);
  const unsigned blocks = (unsigned)((waves + 3) / 4);
  const int iters = (int)((chunks + 4L * blocks - 1) / (4L * blocks));
  MLP_HIP(m, hipMemsetAsync(m->grad, 0, (cbns::MLP_NPARAM + 1) * sizeof(float), m->stream));
  cbns::k_mlp_grad<<<blocks, 256, 0, m->stream>>>(mlp_dev(m), feat, target, n, w_pos,
                                                  1.0f / (float)n, iters, m->grad,
                                                  m->grad + cbns::MLP_NPARAM);
  MLP_HIP(m, hipGetLastError());
  float lsum = 0.0f;
  MLP_HIP(m, hipMemcpyAsync(&lsum, m->grad + cbns::MLP_NPARAM, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (grad)
    MLP_HIP(m, hipMemcpyAsync(grad, m->grad, cbns::MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  if (loss) *loss = (double)lsum / (double)n;
  return CBNS_OK;
}

int cbns_mlp_adam_step(cbns_mlp* m, float lr, float beta1, float beta2, float eps) {
  if (!m || !(lr > 0.0f)) return CBNS_ERR_ARG;
  m->t++;
  const float c1 = (float)(1.0 - std::pow((double)beta1, m->t));
  const float c2 = (float)(1.0 - std::pow((double)beta2, m->t));
  cbns::k_mlp_adam<<<(cbns::MLP_NPARAM + 255) / 256, 256, 0, m->stream>>>(m->P, m->grad, m->m, m->v,
                                                                          lr, beta1, beta2, eps, c1, c2);
  MLP_HIP(m, hipGetLastError());
  if (int rc = mlp_repack(m)) return rc;
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return CBNS_OK;
}

}  // extern "C"

extern "C" {

int cbns_dev_alloc(int64_t bytes, void** d) {
  if (!d || bytes <= 0) return CBNS_ERR_ARG;
  *d = nullptr;
  return hipMalloc(d, (size_t)bytes) == hipSuccess ? CBNS_OK : CBNS_ERR_HIP;
}

int cbns_dev_free(void* d) {
  if (!d) return CBNS_OK;
  return hipFree(d) == hipSuccess ? CBNS_OK : CBNS_ERR_HIP;
}

int cbns_dev_upload(void* d, const void* host, int64_t bytes) {
  if (!d || !host || bytes < 0) return CBNS_ERR_ARG;
  return hipMemcpy(d, host, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? CBNS_OK
                                                                                 : CBNS_ERR_HIP;
}

int cbns_dev_download(void* host, const void* d, int64_t bytes) {
  if (!d || !host || bytes < 0) return CBNS_ERR_ARG;
  return hipMemcpy(host, d, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? CBNS_OK
                                                                                 : CBNS_ERR_HIP;
}

int cbns_gather_rows(const float* src, const int64_t* idx, int64_t n, int32_t row, float* dst) {
  if (n < 0 || row <= 0 || (n > 0 && (!src || !idx || !dst))) return CBNS_ERR_ARG;
  if (n == 0) return CBNS_OK;
  const int64_t total = n * row;
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffL) return CBNS_ERR_ARG;
  cbns::k_gather_rows<<<(unsigned)blocks, 256>>>(src, idx, n, row, dst);
  if (hipGetLastError() != hipSuccess) return CBNS_ERR_HIP;
  return hipDeviceSynchronize() == hipSuccess ? CBNS_OK : CBNS_ERR_HIP;
}

}  // extern "C"
