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
constexpr int DT_KCH = 32;  // nodes along k per thread of the time-step pass

dim3 time_step_grid(const Consts& c) {
  return dim3((unsigned)(c.IM + 63) / 64, (unsigned)(c.JM + 3) / 4,
              (unsigned)(c.KM + DT_KCH - 1) / DT_KCH);
}

// Per-node spectral-radius sums and their maximum per workgroup: 64 x 4
// threads on (i, j), each walking DT_KCH nodes along k (node addresses
// without integer division; loads coalesced along i).
__global__ void __launch_bounds__(TPB) k_time_step(Consts c, const double* __restrict__ prim,
                                                   const double* __restrict__ met,
                                                   double* __restrict__ partial) {
  __shared__ double lds[TPB / 64];
  const int i = (int)blockIdx.x * 64 + (int)threadIdx.x;
  const int j = (int)blockIdx.y * 4 + (int)threadIdx.y;
  const int k0 = (int)blockIdx.z * DT_KCH, k1 = min(k0 + DT_KCH, c.KM);
  double lam = -INFINITY;
  int nan_seen = 0;
  if (i < c.IM && j < c.JM) {
    for (int k = k0; k < k1; k++) {
      const double sum = node_lambda(c, prim, met, idx(c, i, j, k));
      if (isnan(sum)) nan_seen = 1;
      else lam = fmax(lam, sum);
    }
  }
  const int tid = (int)threadIdx.x + 64 * (int)threadIdx.y;
  lam = block_max(lam, lds, tid, TPB);
  __syncthreads();
  const int any_nan = __syncthreads_or(nan_seen);
  if (tid == 0)
    partial[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = any_nan ? NAN : lam;
}

// out[0] = CFL / max, limited to cap (the rest of the run, t_end - t_cur;
// src/Block3d.cpp:241-246), out[1] = max; both NaN if any partial is NaN
__global__ void __launch_bounds__(TPB) k_dt_final(const double* __restrict__ partial, int n,
                                                  double CFL, double cap,
                                                  double* __restrict__ out) {
  __shared__ double lds[TPB / 64];
  double lam = -INFINITY;
  int nan_seen = 0;
  for (int t = threadIdx.x; t < n; t += TPB) {
    const double v = partial[t];
    if (isnan(v)) nan_seen = 1;
    else lam = fmax(lam, v);
  }
  lam = block_max(lam, lds, threadIdx.x, TPB);
  __syncthreads();
  const int any_nan = __syncthreads_or(nan_seen);
  if (threadIdx.x == 0) {
    const double dt = CFL / lam;
    out[0] = any_nan ? NAN : (cap < dt ? cap : dt);
    out[1] = any_nan ? NAN : lam;
  }
}

// Sutherland viscosity (unscaled) over the viscous node range [-2, N+1]^3,
// as calc_viscous_terms leaves it (src/viscous_terms_kernel.cu:61-66)
__global__ void __launch_bounds__(256) k_sutherland(Consts c, Box box,
                                                    const double* __restrict__ prim,
                                                    double* __restrict__ mu) {
  int i, j, k;
  if (!box_node3(box, i, j, k)) return;
  const int o = idx(c, i, j, k);
  const double T_l = prim[P_T * (long)c.S + o], C = c.C_T_inf;
  mu[o] = (1.0 + C) / (T_l + C) * sqrt(T_l * T_l * T_l);
}

}  // namespace

int launch_sutherland(cbns_block* b) {
  const Consts& c = b->c;
  Box box;
  for (int a = 0; a < 3; a++) {
    box.lo[a] = -2;
    box.n[a] = axis_n(c, a) + 4;
  }
  k_sutherland<<<box_grid(box), box_block(), 0, b->stream>>>(c, box, b->prim, b->mu);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

long time_step_blocks(const Consts& c) {
  const dim3 g = time_step_grid(c);
  return (long)g.x * g.y * g.z;
}

// The time step on the device: out[0] = CFL / max (limited to cap), out[1]
// = max; NaN if any node's sum is NaN.
int launch_dt_device(cbns_block* b, double cap, double* out) {
  const Consts& c = b->c;
  const int nb = (int)time_step_blocks(c);
  if (nb > b->red_n - 4) return fail(b, CBNS_ERR_STATE, "time-step partials exceed their buffer");
  k_time_step<<<time_step_grid(c), dim3(64, 4), 0, b->stream>>>(c, b->prim, b->met, b->red + 4);
  CBNS_HIP(b, hipGetLastError());
  k_dt_final<<<1, TPB, 0, b->stream>>>(b->red + 4, nb, c.CFL, cap, out);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

int launch_calc_dt(cbns_block* b, double* dt_host) {
  int rc = launch_dt_device(b, INFINITY, b->red);
  if (rc) return rc;
  CBNS_HIP(b, hipMemcpyAsync(b->h_scalar, b->red, 2 * sizeof(double), hipMemcpyDeviceToHost,
                             b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  *dt_host = b->h_scalar[0];
  return CBNS_OK;
}

// The time step from per-workgroup maxima another pass wrote (the viscous
// flux pass of a step's first stage, k_viscous.hip): dt, limited to cap,
// into out[0] on the device, where the stage kernels read it.
int launch_dt_from_partials(cbns_block* b, const double* partial, int n, double cap,
                            double* out) {
  k_dt_final<<<1, TPB, 0, b->stream>>>(partial, n, b->c.CFL, cap, out);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

}  // namespace cbns
