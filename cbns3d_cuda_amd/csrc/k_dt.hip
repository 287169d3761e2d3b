// k_dt.hip -- global CFL time step: per-node spectral-radius sum
// (src/time_step_kernel.cu:43-94) reduced to its maximum with wavefront
// shuffles and one LDS pass per workgroup, then a single-workgroup pass over
// the partial maxima (the reference uses thrust::max_element,
// src/calc_dt.cu:50-52).  max is exact, so the result does not depend on the
// reduction order.  Any NaN makes the result NaN (the reference checks only
// the maximum, src/calc_dt.cu:56-58).
#include "cbns_internal.hpp"

namespace cbns {

namespace {

constexpr int TPB = 256;

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline double block_max(double v, double* lds) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  if (wid == 0) {
    v = lane < (TPB / 64) ? lds[lane] : -INFINITY;
    v = wave_max(v);
  }
  return v;
}

__global__ void __launch_bounds__(TPB) k_time_step(Consts c, const double* __restrict__ prim,
                                                   const double* __restrict__ met,
                                                   double* __restrict__ partial) {
  __shared__ double lds[TPB / 64];
  const long N = (long)c.IM * c.JM * c.KM;
  const long S = c.S;
  double lam = -INFINITY;
  int nan_seen = 0;
  for (long t = (long)blockIdx.x * TPB + threadIdx.x; t < N; t += (long)gridDim.x * TPB) {
    const int i = (int)(t % c.IM);
    const int j = (int)((t / c.IM) % c.JM);
    const int k = (int)(t / ((long)c.IM * c.JM));
    const int o = idx(c, i, j, k);
    const double rr = prim[o], uu = prim[S + o], vv = prim[2 * S + o], ww = prim[3 * S + o];
    const double pp = prim[4 * S + o];
    const double cs = sqrt(c.gamma * pp / rr);
    double factor_v = 0.0;
    if (c.viscous) {
      const double T_l = prim[5 * S + o];
      const double mu = (1.0 + c.C_T_inf) / (T_l + c.C_T_inf) * sqrt(T_l * T_l * T_l);
      factor_v = 2.0 * mu * c.C_dt_v * sqrt(uu * uu + vv * vv + ww * ww) / cs / rr;
    }
    double sum = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double n_x = met[(3 * d + 0) * S + o], n_y = met[(3 * d + 1) * S + o],
                   n_z = met[(3 * d + 2) * S + o];
      const double qn = uu * n_x + vv * n_y + ww * n_z;
      const double n_abs = sqrt(n_x * n_x + n_y * n_y + n_z * n_z);
      double tt = fabs(qn) + cs * n_abs;
      if (c.viscous) tt += n_abs * n_abs * factor_v;
      sum = d == 0 ? tt : sum + tt;
    }
    if (isnan(sum)) nan_seen = 1;
    else lam = fmax(lam, sum);
  }
  lam = block_max(lam, lds);
  __syncthreads();
  const int any_nan = __syncthreads_or(nan_seen);
  if (threadIdx.x == 0) partial[blockIdx.x] = any_nan ? NAN : lam;
}

__global__ void __launch_bounds__(TPB) k_dt_final(const double* __restrict__ partial, int n,
                                                  double CFL, double* __restrict__ out) {
  __shared__ double lds[TPB / 64];
  double lam = -INFINITY;
  int nan_seen = 0;
  for (int t = threadIdx.x; t < n; t += TPB) {
    const double v = partial[t];
    if (isnan(v)) nan_seen = 1;
    else lam = fmax(lam, v);
  }
  lam = block_max(lam, lds);
  __syncthreads();
  const int any_nan = __syncthreads_or(nan_seen);
  if (threadIdx.x == 0) {
    out[0] = any_nan ? NAN : CFL / lam;
    out[1] = any_nan ? NAN : lam;
  }
}

}  // namespace

int launch_calc_dt(cbns_block* b, double* dt_host) {
  const Consts& c = b->c;
  const long N = (long)c.IM * c.JM * c.KM;
  int nb = (int)((N + TPB - 1) / TPB);
  if (nb > b->red_n - 4) nb = b->red_n - 4;
  k_time_step<<<nb, TPB, 0, b->stream>>>(c, b->prim, b->met, b->red + 4);
  CBNS_HIP(b, hipGetLastError());
  k_dt_final<<<1, TPB, 0, b->stream>>>(b->red + 4, nb, c.CFL, b->red);
  CBNS_HIP(b, hipGetLastError());
  CBNS_HIP(b, hipMemcpyAsync(b->h_scalar, b->red, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  *dt_host = b->h_scalar[0];
  return CBNS_OK;
}

}  // namespace cbns
