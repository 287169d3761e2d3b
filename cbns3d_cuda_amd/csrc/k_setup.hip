// k_setup.hip -- once-per-run kernels: grid metrics and Jacobian (FP64),
// the node-array gradient (test launcher), freestream initial condition and
// the Q <-> primitive conversions over the whole ghosted array.
#include "cbns_internal.hpp"

namespace cbns {

namespace {

// One-axis derivative of src/gradient_kernel.cu:40-131.  f(o) is the value
// at offset o along the axis from position p of an n-node line.
template <class F>
__device__ inline double grad1(int p, int n, F f) {
  double r = 0.0;
  if (0 == p) r = -25.0 / 12.0 * f(0) + 4.0 * f(1) - 3.0 * f(2) + (4.0 / 3.0) * f(3) - 0.25 * f(4);
  if (1 == p) r = -5.0 / 6.0 * f(0) - 0.25 * f(-1) + 1.5 * f(1) - 0.5 * f(2) + (1.0 / 12.0) * f(3);
  if ((1 < p) && (n - 2 > p))
    r = (1.0 / 12.0) * f(-2) - 2.0 / 3.0 * f(-1) + (2.0 / 3.0) * f(1) - 1.0 / 12.0 * f(2);
  if (n - 2 == p) r = (5.0 / 6.0) * f(0) - 1.0 / 12.0 * f(-3) + 0.5 * f(-2) - 1.5 * f(-1) + 0.25 * f(1);
  if (n - 1 == p) r = (25.0 / 12.0) * f(0) + 0.25 * f(-4) - 4.0 / 3.0 * f(-3) + 3.0 * f(-2) - 4.0 * f(-1);
  return r;
}

struct Period {
  double v[9];
};

// src/calc_metrics.cu:50-91 + src/metrics_kernel.cu:52-70, one thread per node
__global__ void k_metrics(Consts c, const double* __restrict__ x, const double* __restrict__ y,
                          const double* __restrict__ z, Period per, double* __restrict__ met,
                          int* __restrict__ bad) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long NN = (long)c.IM * c.JM * c.KM;
  if (t >= NN) return;
  const int i = (int)(t % c.IM);
  const int j = (int)((t / c.IM) % c.JM);
  const int k = (int)(t / ((long)c.IM * c.JM));
  const int pos[3] = {i, j, k};
  const int n[3] = {c.IM, c.JM, c.KM};
  const long s[3] = {1, c.IM, (long)c.IM * c.JM};
  const double* X[3] = {x, y, z};
  double d[3][3];  // d[coord][axis]
  for (int cc = 0; cc < 3; cc++)
    for (int a = 0; a < 3; a++) {
      const double* f = X[cc] + t;
      const int p = pos[a];
      if (periodic(c, a)) {
        double fv[5];
        for (int o = -2; o <= 2; o++) {
          int q = p + o, wrap = 0;
          while (q < 0) { q += n[a]; wrap--; }
          while (q >= n[a]) { q -= n[a]; wrap++; }
          fv[o + 2] = f[(long)(q - p) * s[a]] + wrap * per.v[3 * a + cc];
        }
        d[cc][a] = (1.0 / 12.0) * fv[0] - 2.0 / 3.0 * fv[1] + (2.0 / 3.0) * fv[3] - 1.0 / 12.0 * fv[4];
      } else {
        d[cc][a] = grad1(p, n[a], [&](int o) { return f[(long)o * s[a]]; });
      }
    }
  const double x_xi = d[0][0], x_eta = d[0][1], x_zeta = d[0][2];
  const double y_xi = d[1][0], y_eta = d[1][1], y_zeta = d[1][2];
  const double z_xi = d[2][0], z_eta = d[2][1], z_zeta = d[2][2];
  const double Jac_a = -x_eta * y_xi * z_zeta + x_eta * y_zeta * z_xi + x_xi * y_eta * z_zeta -
                       x_xi * y_zeta * z_eta - x_zeta * y_eta * z_xi + x_zeta * y_xi * z_eta;
  const int o = idx(c, i, j, k);
  const long S = c.S;
  met[JAC * S + o] = Jac_a;
  met[XI_X * S + o] = (y_eta * z_zeta - y_zeta * z_eta) / Jac_a;
  met[XI_Y * S + o] = (-x_eta * z_zeta + x_zeta * z_eta) / Jac_a;
  met[XI_Z * S + o] = (x_eta * y_zeta - x_zeta * y_eta) / Jac_a;
  met[ETA_X * S + o] = (-y_xi * z_zeta + y_zeta * z_xi) / Jac_a;
  met[ETA_Y * S + o] = (x_xi * z_zeta - x_zeta * z_xi) / Jac_a;
  met[ETA_Z * S + o] = (-x_xi * y_zeta + x_zeta * y_xi) / Jac_a;
  met[ZETA_X * S + o] = (-y_eta * z_xi + y_xi * z_eta) / Jac_a;
  met[ZETA_Y * S + o] = (x_eta * z_xi - x_xi * z_eta) / Jac_a;
  met[ZETA_Z * S + o] = (-x_eta * y_xi + x_xi * y_eta) / Jac_a;
  if (!(Jac_a > 0.0)) atomicOr(bad, 1);
}

__device__ inline int ghost_src(const Consts& c, int a, int m) {
  const int n = axis_n(c, a);
  if (periodic(c, a)) return ((m % n) + n) % n;
  return m < 0 ? 0 : (m >= n ? n - 1 : m);
}

// ghost extension of the metrics: nearest boundary node on reference axes
// (src/metrics_kernel.cu:72-560: faces, edges and corners), wrap on periodic
// axes.  One thread per ghosted storage node; interior nodes return.
__global__ void k_metrics_ghost(Consts c, double* __restrict__ met) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.S) return;
  const int i = (int)(t % c.PX) - G;
  const int j = (int)((t / c.PX) % c.PY) - G;
  const int k = (int)(t / c.PXY) - G;
  if (i >= 0 && i < c.IM && j >= 0 && j < c.JM && k >= 0 && k < c.KM) return;
  const int s = idx(c, ghost_src(c, 0, i), ghost_src(c, 1, j), ghost_src(c, 2, k));
  for (int q = 0; q < 10; q++) met[(long)q * c.S + t] = met[(long)q * c.S + s];
}

// src/gradient_kernel.cu:13-133 on node arrays (the calc_gradient launcher)
__global__ void k_gradient_nodes(int IM, int JM, int KM, const double* __restrict__ f,
                                 double* __restrict__ fx, double* __restrict__ fy,
                                 double* __restrict__ fz) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long NN = (long)IM * JM * KM;
  if (t >= NN) return;
  const int i = (int)(t % IM);
  const int j = (int)((t / IM) % JM);
  const int k = (int)(t / ((long)IM * JM));
  const double* fp = f + t;
  fx[t] = grad1(i, IM, [&](int o) { return fp[o]; });
  fy[t] = grad1(j, JM, [&](int o) { return fp[(long)o * IM]; });
  fz[t] = grad1(k, KM, [&](int o) { return fp[(long)o * IM * JM]; });
}

// src/initial_condition_kernel.cu:21-45 over the ghosted array
__global__ void k_freestream(Consts c, double aoa, double* __restrict__ prim) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.S) return;
  const long S = c.S;
  prim[P_RHO * S + t] = 1.0;
  prim[P_U * S + t] = cos(aoa);
  prim[P_V * S + t] = sin(aoa);
  prim[P_W * S + t] = 0.0;
  prim[P_P * S + t] = c.p_inf;
  prim[P_T * S + t] = 1.0;
}

// src/conservative_kernel.cu:21-46 over the ghosted array
__global__ void k_conservative(Consts c, const double* __restrict__ prim, double* __restrict__ Q) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.S) return;
  const long S = c.S;
  const double rr = prim[t], uu = prim[S + t], vv = prim[2 * S + t], ww = prim[3 * S + t];
  const double pp = prim[4 * S + t];
  Q[t] = rr;
  Q[S + t] = rr * uu;
  Q[2 * S + t] = rr * vv;
  Q[3 * S + t] = rr * ww;
  Q[4 * S + t] = pp / c.gam1 + 0.5 * rr * (uu * uu + vv * vv + ww * ww);
}

// src/primitive_kernel.cu:21-49 over the ghosted array
__global__ void k_primitive(Consts c, const double* __restrict__ Q, double* __restrict__ prim) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= c.S) return;
  const long S = c.S;
  const double rr = Q[t];
  const double uu = Q[S + t] / rr;
  const double vv = Q[2 * S + t] / rr;
  const double ww = Q[3 * S + t] / rr;
  const double pp = c.gam1 * (Q[4 * S + t] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  prim[t] = rr;
  prim[S + t] = uu;
  prim[2 * S + t] = vv;
  prim[3 * S + t] = ww;
  prim[4 * S + t] = pp;
  prim[5 * S + t] = c.gM2 * pp / rr;
}

__global__ void k_fill(double* __restrict__ p, long n, double v) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

inline unsigned nblk(long n, int th) { return (unsigned)((n + th - 1) / th); }

}  // namespace

int launch_metrics(cbns_block* b, const double* dx, const double* dy, const double* dz,
                   const double* period, int* bad_jac) {
  const Consts& c = b->c;
  Period per{};
  if (period)
    for (int q = 0; q < 9; q++) per.v[q] = period[q];
  int* dbad = nullptr;
  CBNS_HIP(b, hipMallocAsync((void**)&dbad, sizeof(int), b->stream));
  const int zero = 0;
  CBNS_HIP(b, hipMemcpyAsync(dbad, &zero, sizeof(int), hipMemcpyHostToDevice, b->stream));
  const long NN = (long)c.IM * c.JM * c.KM;
  k_metrics<<<nblk(NN, 256), 256, 0, b->stream>>>(c, dx, dy, dz, per, b->met, dbad);
  CBNS_HIP(b, hipGetLastError());
  k_metrics_ghost<<<nblk(c.S, 256), 256, 0, b->stream>>>(c, b->met);
  CBNS_HIP(b, hipGetLastError());
  int hbad = 0;
  CBNS_HIP(b, hipMemcpyAsync(&hbad, dbad, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  CBNS_HIP(b, hipFreeAsync(dbad, b->stream));
  *bad_jac = hbad;
  return CBNS_OK;
}

int launch_gradient_nodes(hipStream_t s, int IM, int JM, int KM, const double* f, double* fx,
                          double* fy, double* fz) {
  const long NN = (long)IM * JM * KM;
  k_gradient_nodes<<<nblk(NN, 256), 256, 0, s>>>(IM, JM, KM, f, fx, fy, fz);
  return hipGetLastError() == hipSuccess ? CBNS_OK : CBNS_ERR_HIP;
}

int launch_freestream(cbns_block* b) {
  const double aoa = b->p.angle_attack * 3.14159265358979323846 / 180.0;
  k_freestream<<<nblk(b->c.S, 256), 256, 0, b->stream>>>(b->c, aoa, b->prim);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

int launch_conservative_all(cbns_block* b, double* Q) {
  k_conservative<<<nblk(b->c.S, 256), 256, 0, b->stream>>>(b->c, b->prim, Q);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

int launch_primitive_all(cbns_block* b, const double* Q) {
  k_primitive<<<nblk(b->c.S, 256), 256, 0, b->stream>>>(b->c, Q, b->prim);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

int launch_fill(hipStream_t s, double* p, long n, double v) {
  k_fill<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
  return hipGetLastError() == hipSuccess ? CBNS_OK : CBNS_ERR_HIP;
}

}  // namespace cbns
