// cbns_api.hip -- extern "C" entry points of libcbns (include/cbns.h).
// Host orchestration of the block3d_cuda launcher API (include/*.h of the
// reference) on one HIP stream per block; status codes instead of exit().
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "cbns_internal.hpp"

namespace cbns {

int fail(cbns_block* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return code;
}

int hip_fail(cbns_block* b, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(b, CBNS_ERR_HIP, m);
}

void prof_begin(cbns_block* b, int id, std::pair<hipEvent_t, hipEvent_t>& ev, hipStream_t s) {
  if (!b->prof.on) return;
  if (b->prof.pool.empty()) {
    hipEvent_t a = nullptr, c = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&c) != hipSuccess) return;
    b->prof.pool.emplace_back(a, c);
  }
  ev = b->prof.pool.back();
  b->prof.pool.pop_back();
  (void)hipEventRecord(ev.first, s);
  (void)id;
}

void prof_end(cbns_block* b, int id, std::pair<hipEvent_t, hipEvent_t>& ev, hipStream_t s) {
  if (!b->prof.on || !ev.first) return;
  (void)hipEventRecord(ev.second, s);
  b->prof.live[id].push_back(ev);
}

ProfScope::ProfScope(cbns_block* b_, int id_, hipStream_t s_)
    : b(b_), id(id_), s(s_ ? s_ : b_->stream) {
  prof_begin(b, id, ev, s);
}

namespace {

void derive(cbns_block* b) {
  // src/Block3d.cpp:99-113
  const cbns_params& p = b->p;
  Consts& c = b->c;
  c.IM = p.IM; c.JM = p.JM; c.KM = p.KM;
  c.PX = p.IM + 2 * G; c.PY = p.JM + 2 * G; c.PZ = p.KM + 2 * G;
  c.PXY = c.PX * c.PY;
  c.S = c.PXY * c.PZ;
  for (int a = 0; a < 3; a++) c.bc[a] = p.bc[a];
  c.viscous = p.viscous; c.visc_centered = p.visc_centered;
  c.i_begin = p.i_begin; c.i_end = p.i_end;
  c.CFL = p.CFL;
  c.gamma = p.gamma;
  c.Mach2 = p.Mach * p.Mach;
  c.Re_inv = 1.0 / p.Re;
  c.Pr_inv = 1.0 / p.Pr;
  c.gam1 = p.gamma - 1.0;
  c.gM2 = p.gamma * c.Mach2;
  c.gam1_inv = 1.0 / c.gam1;
  c.p_inf = 1.0 / c.gM2;
  c.C_T_inf = 110.5 / p.T_inf;
  c.C_dt_v = std::max(4.0 / 3.0, p.gamma * c.Pr_inv) * c.Re_inv;
  c.qcoef = -c.Pr_inv * c.gam1_inv / c.Mach2;
}

int dmalloc(cbns_block* b, double** p, long n) {
  CBNS_HIP(b, hipMalloc((void**)p, sizeof(double) * (size_t)n));
  return CBNS_OK;
}

double* field_ptr(cbns_block* b, int f) {
  const long S = b->c.S;
  if (f >= CBNS_F_Q && f < CBNS_F_Q + 5) return b->Q + (f - CBNS_F_Q) * S;
  if (f >= CBNS_F_QP && f < CBNS_F_QP + 5) return b->qp + (f - CBNS_F_QP) * S;
  if (f >= CBNS_F_PRIM && f < CBNS_F_PRIM + 6) return b->prim + (f - CBNS_F_PRIM) * S;
  if (f >= CBNS_F_MET && f < CBNS_F_MET + 10) return b->met + (f - CBNS_F_MET) * S;
  if (f == CBNS_F_MU) return b->mu;
  if (f >= CBNS_F_FLUX && f < CBNS_F_FLUX + 15) return b->flux ? b->flux + (f - CBNS_F_FLUX) * S : nullptr;
  if (f >= CBNS_F_DFV && f < CBNS_F_DFV + 4) return b->dfv ? b->dfv + (f - CBNS_F_DFV) * S : nullptr;
  return nullptr;
}

// copy between a 2-ghost-layer host array (the reference's IM_G x JM_G x KM_G
// layout, src/Block3d.cpp:325-335) and one device field
void ref_to_internal(const Consts& c, const double* src, long src_stride, long src_off,
                     std::vector<double>& dst) {
  const int NG = CBNS_NG_REF;
  const long RX = c.IM + 2 * NG, RY = c.JM + 2 * NG, RZ = c.KM + 2 * NG;
  for (long k = 0; k < RZ; k++)
    for (long j = 0; j < RY; j++)
      for (long i = 0; i < RX; i++) {
        const long r = i + RX * (j + RY * k);
        dst[idx(c, (int)i - NG, (int)j - NG, (int)k - NG)] = src[r * src_stride + src_off];
      }
}
void internal_to_ref(const Consts& c, const std::vector<double>& src, double* dst,
                     long dst_stride, long dst_off) {
  const int NG = CBNS_NG_REF;
  const long RX = c.IM + 2 * NG, RY = c.JM + 2 * NG, RZ = c.KM + 2 * NG;
  for (long k = 0; k < RZ; k++)
    for (long j = 0; j < RY; j++)
      for (long i = 0; i < RX; i++) {
        const long r = i + RX * (j + RY * k);
        dst[r * dst_stride + dst_off] = src[idx(c, (int)i - NG, (int)j - NG, (int)k - NG)];
      }
}

// lazily allocated scratch (zero-filled like every array): the face fluxes
// and dfv exist only for the reference-decomposed calls, the viscous-flux
// scratch for any viscous stage
int lazy_alloc(cbns_block* b, double** p, long n) {
  if (*p) return CBNS_OK;
  int rc = dmalloc(b, p, n);
  if (rc) return rc;
  if (launch_fill(b->stream, *p, n, 0.0)) return fail(b, CBNS_ERR_HIP, "fill failed");
  return CBNS_OK;
}

int ensure_stage_scratch(cbns_block* b) {
  const long S = b->c.S;
  int rc;
  if (fused_stage_uses_face_flux() && (rc = lazy_alloc(b, &b->flux, 15 * S))) return rc;
  if (b->c.viscous && (rc = lazy_alloc(b, &b->vtmp, 12 * S))) return rc;
  if (b->c.viscous && (rc = lazy_alloc(b, &b->lam_part, visc_flux_blocks(b->c)))) return rc;
  return CBNS_OK;
}

int ensure_unfused(cbns_block* b) {
  const long S = b->c.S;
  int rc;
  if ((rc = lazy_alloc(b, &b->flux, 15 * S))) return rc;
  if (b->c.viscous) {
    if ((rc = lazy_alloc(b, &b->dfv, 4 * S))) return rc;
    if ((rc = lazy_alloc(b, &b->vtmp, 12 * S))) return rc;
  }
  return CBNS_OK;
}

int check_block(cbns_block* b) {
  if (!b) return CBNS_ERR_ARG;
  if (!b->have_mesh) return fail(b, CBNS_ERR_STATE, "no mesh: call cbns_set_mesh first");
  CBNS_HIP(b, hipSetDevice(b->device));
  return CBNS_OK;
}

double* buf_ptr(cbns_block* b, int buf) { return buf == CBNS_BUF_Q ? b->Q : b->qp; }

}  // namespace
}  // namespace cbns

using namespace cbns;

extern "C" {

int cbns_create(const cbns_params* p, cbns_block** out) {
  if (!p || !out) return CBNS_ERR_ARG;
  *out = nullptr;
  if (p->IM < 5 || p->JM < 5 || p->KM < 5) return CBNS_ERR_ARG;
  for (int a = 0; a < 3; a++)
    if (p->bc[a] != CBNS_BC_REF && p->bc[a] != CBNS_BC_PERIODIC) return CBNS_ERR_ARG;
  if (!(p->gamma > 1.0) || !(p->Mach > 0.0) || !(p->CFL > 0.0)) return CBNS_ERR_ARG;
  if (p->viscous && (!(p->Re > 0.0) || !(p->Pr > 0.0) || !(p->T_inf > 0.0))) return CBNS_ERR_ARG;
  const long S = (long)(p->IM + 2 * G) * (p->JM + 2 * G) * (p->KM + 2 * G);
  if (S >= (1L << 31) - 1) return CBNS_ERR_ARG;  // 32-bit node offsets within a field
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CBNS_ERR_NO_DEVICE;
  cbns_block* b = new (std::nothrow) cbns_block();
  if (!b) return CBNS_ERR_ARG;
  b->p = *p;
  derive(b);
  if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
    delete b;
    return CBNS_ERR_HIP;
  }
  int rc = 0;
  if ((rc = dmalloc(b, &b->met, 10 * S)) || (rc = dmalloc(b, &b->Q, 5 * S)) ||
      (rc = dmalloc(b, &b->qp, 5 * S)) || (rc = dmalloc(b, &b->qspare, 5 * S)) ||
      (rc = dmalloc(b, &b->prim, 6 * S)) || (rc = dmalloc(b, &b->mu, S)) ||
      (rc = dmalloc(b, &b->rhs, 5 * S))) {
    cbns_destroy(b);
    return rc;
  }
  b->red_n = (int)(4 + std::max<long>(4096, time_step_blocks(b->c)));
  if ((rc = dmalloc(b, &b->red, b->red_n))) {
    cbns_destroy(b);
    return rc;
  }
  if (hipHostMalloc((void**)&b->h_scalar, 4 * sizeof(double), hipHostMallocDefault) !=
      hipSuccess) {
    cbns_destroy(b);
    return CBNS_ERR_HIP;
  }
  // defined contents everywhere (ghost corners a boundary rule never writes)
  double* zs[] = {b->met, b->Q, b->qp, b->qspare, b->prim, b->mu, b->rhs};
  const long zn[] = {10 * S, 5 * S, 5 * S, 5 * S, 6 * S, S, 5 * S};
  for (int q = 0; q < 7; q++)
    if (launch_fill(b->stream, zs[q], zn[q], 0.0)) {
      cbns_destroy(b);
      return CBNS_ERR_HIP;
    }
  if (hipStreamSynchronize(b->stream) != hipSuccess) {
    cbns_destroy(b);
    return CBNS_ERR_HIP;
  }
  *out = b;
  return CBNS_OK;
}

void cbns_destroy(cbns_block* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  double* ps[] = {b->met, b->Q, b->qp, b->qspare, b->prim, b->mu, b->rhs, b->flux, b->dfv, b->vtmp,
                  b->red, b->lam_part};
  for (double* p : ps)
    if (p) (void)hipFree(p);
  for (void* p : {(void*)b->flags, (void*)b->ind_P, b->ind_pack, (void*)b->ind_count,
                  (void*)b->tslot, (void*)b->tlist, (void*)b->tflux, (void*)b->tcount})
    if (p) (void)hipFree(p);
  if (b->h_scalar) (void)hipHostFree(b->h_scalar);
  for (auto& v : b->prof.live)
    for (auto& ev : v) b->prof.pool.push_back(ev);
  for (auto& ev : b->prof.pool) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* cbns_last_error(const cbns_block* b) { return b ? b->err.c_str() : "null block"; }

int64_t cbns_field_size(const cbns_block* b) { return b ? (int64_t)b->c.S : 0; }

int cbns_set_mesh(cbns_block* b, const double* x, const double* y, const double* z,
                  const double* period) {
  if (!b || !x || !y || !z) return CBNS_ERR_ARG;
  for (int a = 0; a < 3; a++)
    if (b->c.bc[a] == CBNS_BC_PERIODIC && !period)
      return fail(b, CBNS_ERR_ARG, "periodic axis needs the period vectors");
  CBNS_HIP(b, hipSetDevice(b->device));
  const long NN = (long)b->c.IM * b->c.JM * b->c.KM;
  double* d = nullptr;
  int rc;
  if ((rc = dmalloc(b, &d, 3 * NN))) return rc;
  const size_t nb = sizeof(double) * (size_t)NN;
  hipError_t e = hipMemcpyAsync(d, x, nb, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d + NN, y, nb, hipMemcpyHostToDevice, b->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d + 2 * NN, z, nb, hipMemcpyHostToDevice, b->stream);
  int bad = 0;
  rc = e == hipSuccess ? launch_metrics(b, d, d + NN, d + 2 * NN, period, &bad)
                       : hip_fail(b, e, "coordinate upload");
  (void)hipStreamSynchronize(b->stream);
  (void)hipFree(d);
  if (rc) return rc;
  b->have_mesh = true;
  if (bad) return fail(b, CBNS_ERR_MESH, "non-positive Jacobian in the mesh");
  return CBNS_OK;
}

int cbns_get_metrics(cbns_block* b, double* xi_x, double* xi_y, double* xi_z, double* eta_x,
                     double* eta_y, double* eta_z, double* zeta_x, double* zeta_y, double* zeta_z,
                     double* Jac) {
  int rc = check_block(b);
  if (rc) return rc;
  double* outs[10] = {xi_x, xi_y, xi_z, eta_x, eta_y, eta_z, zeta_x, zeta_y, zeta_z, Jac};
  const Consts& c = b->c;
  std::vector<double> h((size_t)c.S);
  for (int q = 0; q < 10; q++) {
    if (!outs[q]) return fail(b, CBNS_ERR_ARG, "null metric output");
    CBNS_HIP(b, hipMemcpyAsync(h.data(), b->met + (long)q * c.S, sizeof(double) * c.S,
                               hipMemcpyDeviceToHost, b->stream));
    CBNS_HIP(b, hipStreamSynchronize(b->stream));
    for (int k = 0; k < c.KM; k++)
      for (int j = 0; j < c.JM; j++)
        for (int i = 0; i < c.IM; i++)
          outs[q][i + (long)c.IM * (j + (long)c.JM * k)] = h[idx(c, i, j, k)];
  }
  return CBNS_OK;
}

int cbns_calc_gradient(int32_t IM, int32_t JM, int32_t KM, const double* f, double* fx,
                       double* fy, double* fz) {
  if (IM < 5 || JM < 5 || KM < 5 || !f || !fx || !fy || !fz) return CBNS_ERR_ARG;
  const long NN = (long)IM * JM * KM;
  double* d = nullptr;
  if (hipMalloc((void**)&d, sizeof(double) * 4 * NN) != hipSuccess) return CBNS_ERR_HIP;
  const size_t nb = sizeof(double) * NN;
  int rc = CBNS_OK;
  if (hipMemcpy(d, f, nb, hipMemcpyHostToDevice) != hipSuccess) rc = CBNS_ERR_HIP;
  if (!rc) rc = launch_gradient_nodes(nullptr, IM, JM, KM, d, d + NN, d + 2 * NN, d + 3 * NN);
  if (!rc && (hipMemcpy(fx, d + NN, nb, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(fy, d + 2 * NN, nb, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(fz, d + 3 * NN, nb, hipMemcpyDeviceToHost) != hipSuccess))
    rc = CBNS_ERR_HIP;
  (void)hipFree(d);
  return rc;
}

int cbns_initial_condition(cbns_block* b) {
  if (!b) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  return launch_freestream(b);
}

int cbns_calc_conservative(cbns_block* b) {
  if (!b) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  int rc = launch_conservative_all(b, b->Q);
  if (rc) return rc;
  const size_t nb = sizeof(double) * 5 * (size_t)b->c.S;
  CBNS_HIP(b, hipMemcpyAsync(b->qp, b->Q, nb, hipMemcpyDeviceToDevice, b->stream));
  CBNS_HIP(b, hipMemcpyAsync(b->qspare, b->Q, nb, hipMemcpyDeviceToDevice, b->stream));
  return CBNS_OK;
}

int cbns_set_conservative(cbns_block* b, const double* Q) {
  if (!b || !Q) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  const Consts& c = b->c;
  std::vector<double> h((size_t)c.S, 0.0);
  for (int n = 0; n < 5; n++) {
    // keep the device's outer ghost layers (beyond the reference's two)
    CBNS_HIP(b, hipMemcpyAsync(h.data(), b->Q + (long)n * c.S, sizeof(double) * c.S,
                               hipMemcpyDeviceToHost, b->stream));
    CBNS_HIP(b, hipStreamSynchronize(b->stream));
    ref_to_internal(c, Q, 5, n, h);
    CBNS_HIP(b, hipMemcpyAsync(b->Q + (long)n * c.S, h.data(), sizeof(double) * c.S,
                               hipMemcpyHostToDevice, b->stream));
    CBNS_HIP(b, hipStreamSynchronize(b->stream));
  }
  const size_t nb = sizeof(double) * 5 * (size_t)c.S;
  CBNS_HIP(b, hipMemcpyAsync(b->qp, b->Q, nb, hipMemcpyDeviceToDevice, b->stream));
  CBNS_HIP(b, hipMemcpyAsync(b->qspare, b->Q, nb, hipMemcpyDeviceToDevice, b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  return CBNS_OK;
}

int cbns_get_conservative(cbns_block* b, double* Q) {
  if (!b || !Q) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  const Consts& c = b->c;
  std::vector<double> h((size_t)c.S);
  for (int n = 0; n < 5; n++) {
    CBNS_HIP(b, hipMemcpyAsync(h.data(), b->Q + (long)n * c.S, sizeof(double) * c.S,
                               hipMemcpyDeviceToHost, b->stream));
    CBNS_HIP(b, hipStreamSynchronize(b->stream));
    internal_to_ref(c, h, Q, 5, n);
  }
  return CBNS_OK;
}

int cbns_calc_primitive(cbns_block* b) {
  if (!b) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  return launch_primitive_all(b, b->Q);
}

int cbns_get_primitive(cbns_block* b, double* rho, double* u, double* v, double* w, double* p) {
  if (!b || !rho || !u || !v || !w || !p) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  const Consts& c = b->c;
  double* outs[5] = {rho, u, v, w, p};
  std::vector<double> h((size_t)c.S);
  for (int n = 0; n < 5; n++) {
    CBNS_HIP(b, hipMemcpyAsync(h.data(), b->prim + (long)n * c.S, sizeof(double) * c.S,
                               hipMemcpyDeviceToHost, b->stream));
    CBNS_HIP(b, hipStreamSynchronize(b->stream));
    internal_to_ref(c, h, outs[n], 1, 0);
  }
  return CBNS_OK;
}

int cbns_set_bc_calc_primitive(cbns_block* b, int32_t buf) {
  int rc = check_block(b);
  if (rc) return rc;
  if (buf != CBNS_BUF_Q && buf != CBNS_BUF_QP) return fail(b, CBNS_ERR_ARG, "bad buffer id");
  return launch_bc_primitive(b, buf_ptr(b, buf));
}

int cbns_calc_dt(cbns_block* b, double* dt) {
  int rc = check_block(b);
  if (rc) return rc;
  if (!dt) return fail(b, CBNS_ERR_ARG, "null dt");
  {
    ProfScope ps(b, CBNS_K_DT);
    if ((rc = launch_calc_dt(b, dt))) return rc;
  }
  if (std::isnan(*dt)) return fail(b, CBNS_ERR_NAN, "An error occurred while calculating the time step.");
  return CBNS_OK;
}

int cbns_calc_numerical_flux(cbns_block* b, int32_t buf) {
  int rc = check_block(b);
  if (rc) return rc;
  if (buf != CBNS_BUF_Q && buf != CBNS_BUF_QP) return fail(b, CBNS_ERR_ARG, "bad buffer id");
  if ((rc = ensure_unfused(b))) return rc;
  return launch_numerical_flux(b, buf_ptr(b, buf));
}

int cbns_calc_viscous_terms(cbns_block* b) {
  // Of the reference call's outputs (gradients, mu, tau, q;
  // src/calc_viscous_terms.cu:16-88) only mu has a field here: the
  // gradients, stress and heat flux are formed again, with their divergence,
  // by cbns_calc_viscous_flux_contribution and never stored.
  int rc = check_block(b);
  if (rc) return rc;
  if (!b->c.viscous) return CBNS_OK;
  if ((rc = ensure_unfused(b))) return rc;
  return launch_sutherland(b);
}

int cbns_calc_viscous_flux_contribution(cbns_block* b) {
  int rc = check_block(b);
  if (rc) return rc;
  if (!b->c.viscous) return CBNS_OK;
  if ((rc = ensure_unfused(b))) return rc;
  return launch_viscous_unfused(b);
}

int cbns_update_rk3(cbns_block* b, double dt, int32_t stage) {
  int rc = check_block(b);
  if (rc) return rc;
  if (stage < 1 || stage > 3) return fail(b, CBNS_ERR_ARG, "stage must be 1, 2 or 3");
  if ((rc = ensure_unfused(b))) return rc;
  return launch_update_rk3(b, dt, stage);
}

int cbns_rk3_stage(cbns_block* b, double dt, int32_t stage) {
  int rc = check_block(b);
  if (rc) return rc;
  if (stage < 1 || stage > 3) return fail(b, CBNS_ERR_ARG, "stage must be 1, 2 or 3");
  if ((rc = ensure_stage_scratch(b))) return rc;
  // stage 1: Q -> qp; stage 2: qp -> qspare, then qspare becomes qp;
  // stage 3: qp -> Q.  Boundary nodes no rule writes hold identical values
  // in all three buffers (set/calc_conservative copy Q into each).
  const double* qin = stage == 1 ? b->Q : b->qp;
  double* qout = stage == 1 ? b->qp : (stage == 2 ? b->qspare : b->Q);
  if ((rc = launch_fused_stage(b, dt, stage, qin, qout))) return rc;
  if (stage == 2) std::swap(b->qp, b->qspare);
  ProfScope ps(b, CBNS_K_BC);
  return launch_bc_only(b, stage == 3 ? b->Q : b->qp, stage_prim_mask(b, stage));
}

int cbns_rk3_step(cbns_block* b, double t_end, double* t_cur, double* dt_out) {
  int rc = check_block(b);
  if (rc) return rc;
  if (!t_cur) return fail(b, CBNS_ERR_ARG, "null t_cur");
  if ((rc = ensure_stage_scratch(b))) return rc;
  // The step's time step (limited to the rest of the run) is formed on the
  // device, where the stages read it: on viscous blocks by the first stage's
  // viscous flux pass, which reads every node's primitives and metrics
  // anyway, otherwise by the time-step pass of cbns_calc_dt (the same
  // per-node term and maximum either way).  dt reaches the host once, after
  // the step.  A NaN step leaves the state untouched and is reported.
  b->dt_cap = t_end - *t_cur;
  if (!b->c.viscous) {
    ProfScope ps(b, CBNS_K_DT);
    if ((rc = launch_dt_device(b, b->dt_cap, b->red + 2))) return rc;
  }
  b->dt_fold = true;
  b->step_internal = true;
  for (int s = 1; s <= 3 && !rc; s++) rc = cbns_rk3_stage(b, 0.0, s);
  b->dt_fold = false;
  b->step_internal = false;
  if (rc) return rc;
  CBNS_HIP(b, hipMemcpyAsync(b->h_scalar + 2, b->red + 2, 2 * sizeof(double),
                             hipMemcpyDeviceToHost, b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  const double dt = b->h_scalar[2];
  if (std::isnan(dt)) return fail(b, CBNS_ERR_NAN, "An error occurred while calculating the time step.");
  *t_cur += dt;
  if (dt_out) *dt_out = dt;
  return CBNS_OK;
}

int cbns_set_form(cbns_block* b, int32_t form) {
  if (!b) return CBNS_ERR_ARG;
  if (form < 0 || form > (CBNS_FORM_FP32 | CBNS_FORM_HYBRID)) return fail(b, CBNS_ERR_ARG, "bad form");
  if ((form & CBNS_FORM_HYBRID) && !b->have_indicator)
    return fail(b, CBNS_ERR_STATE, "hybrid form needs cbns_set_indicator first");
  b->form = form;
  b->flags_valid = false;
  return CBNS_OK;
}

int cbns_face_count(cbns_block* b, int32_t axis, int64_t* n) {
  if (!b || !n || axis < 0 || axis > 2) return CBNS_ERR_ARG;
  *n = face_count(b, axis);
  return CBNS_OK;
}

int cbns_face_samples(cbns_block* b, int32_t axis, int32_t buf, float* d_feat, float* d_label) {
  int rc = check_block(b);
  if (rc) return rc;
  if (axis < 0 || axis > 2 || (buf != CBNS_BUF_Q && buf != CBNS_BUF_QP) || !d_feat || !d_label)
    return fail(b, CBNS_ERR_ARG, "bad face_samples argument");
  if ((rc = launch_face_samples(b, axis, buf_ptr(b, buf), d_feat, d_label))) return rc;
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  return CBNS_OK;
}

int cbns_set_indicator(cbns_block* b, const float* params, float threshold, int32_t options) {
  if (!b || !params || !std::isfinite(threshold) || options < 0 || options > 15) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  if (!b->flags) {
    CBNS_HIP(b, hipMalloc((void**)&b->flags, 3 * (size_t)b->c.S));
    CBNS_HIP(b, hipMemsetAsync(b->flags, 0, 3 * (size_t)b->c.S, b->stream));
  }
  // troubled-face lists: a third of each axis' faces, or all of them
  long cap[3], total = 0;
  for (int a = 0; a < 3; a++) {
    const long nf = face_count(b, a);
    cap[a] = (options & 8) ? nf : (nf + 2) / 3;
    total += cap[a];
  }
  if (!b->tslot) {
    CBNS_HIP(b, hipMalloc((void**)&b->tslot, 3 * sizeof(int) * (size_t)b->c.S));
    CBNS_HIP(b, hipMemsetAsync(b->tslot, 0, 3 * sizeof(int) * (size_t)b->c.S, b->stream));
    CBNS_HIP(b, hipMalloc((void**)&b->tcount, 4 * sizeof(unsigned)));
  }
  if (total > b->ttotal) {
    if (b->tlist) CBNS_HIP(b, hipFree(b->tlist));
    if (b->tflux) CBNS_HIP(b, hipFree(b->tflux));
    b->tlist = nullptr;
    b->tflux = nullptr;
    CBNS_HIP(b, hipMalloc((void**)&b->tlist, sizeof(int) * (size_t)total));
    CBNS_HIP(b, hipMalloc((void**)&b->tflux, 5 * sizeof(double) * (size_t)total));
    b->ttotal = total;
  }
  // slots are positions in tflux, whose field stride stays ttotal
  for (int a = 0; a < 3; a++) {
    b->tcap[a] = cap[a];
    b->toff[a] = a == 0 ? 0 : b->toff[a - 1] + cap[a - 1];
  }
  if (!b->ind_P) CBNS_HIP(b, hipMalloc((void**)&b->ind_P, 2176 * sizeof(float)));
  if (!b->ind_pack) CBNS_HIP(b, hipMalloc(&b->ind_pack, 768 * 8));
  if (!b->ind_count) {
    CBNS_HIP(b, hipMalloc((void**)&b->ind_count, 4 * sizeof(unsigned)));
    CBNS_HIP(b, hipMemsetAsync(b->ind_count, 0, 4 * sizeof(unsigned), b->stream));
  }
  std::vector<float> p(params, params + CBNS_MLP_NPARAM);
  for (int r = 0; r < 32; r++) p[r * 32 + 30] = p[r * 32 + 31] = 0.0f;  // padding inputs
  int rc = upload_indicator(b, p.data());
  if (rc) return rc;
  b->ind_thresh = threshold;
  b->ind_per_stage = (options & 1) != 0;
  b->ind_count_on = (options & 2) != 0;
  b->ind_every = (options & 4) ? std::max(1, (int)std::floor(1.0 / b->p.CFL)) : 1;
  b->ind_age = 0;
  b->have_indicator = true;
  b->flags_valid = false;
  return CBNS_OK;
}

int cbns_get_flags(cbns_block* b, uint8_t* host) {
  if (!b || !host) return CBNS_ERR_ARG;
  if (!b->flags) return fail(b, CBNS_ERR_STATE, "no indicator loaded");
  CBNS_HIP(b, hipSetDevice(b->device));
  CBNS_HIP(b, hipMemcpyAsync(host, b->flags, 3 * (size_t)b->c.S, hipMemcpyDeviceToHost, b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  return CBNS_OK;
}

int cbns_flag_count(cbns_block* b, int64_t counts[3]) {
  if (!b || !counts) return CBNS_ERR_ARG;
  if (!b->ind_count) return fail(b, CBNS_ERR_STATE, "no indicator loaded");
  unsigned h[4] = {0, 0, 0, 0};
  CBNS_HIP(b, hipMemcpyAsync(h, b->ind_count, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  for (int a = 0; a < 3; a++) counts[a] = h[a];
  return CBNS_OK;
}

int cbns_get_field(cbns_block* b, int32_t f, double* host) {
  if (!b || !host) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  if (f >= CBNS_F_FLUX && ensure_unfused(b)) return CBNS_ERR_HIP;
  double* d = field_ptr(b, f);
  if (!d) return fail(b, CBNS_ERR_ARG, "field not available");
  CBNS_HIP(b, hipMemcpyAsync(host, d, sizeof(double) * b->c.S, hipMemcpyDeviceToHost, b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  return CBNS_OK;
}

int cbns_set_field(cbns_block* b, int32_t f, const double* host) {
  if (!b || !host) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipSetDevice(b->device));
  if (f >= CBNS_F_FLUX && ensure_unfused(b)) return CBNS_ERR_HIP;
  double* d = field_ptr(b, f);
  if (!d) return fail(b, CBNS_ERR_ARG, "field not available");
  CBNS_HIP(b, hipMemcpyAsync(d, host, sizeof(double) * b->c.S, hipMemcpyHostToDevice, b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  if (f >= CBNS_F_QP && f < CBNS_F_QP + 5) {
    // the ping-pong partner of Q_p mirrors it (boundary nodes no rule writes)
    CBNS_HIP(b, hipMemcpyAsync(b->qspare + (long)(f - CBNS_F_QP) * b->c.S, d,
                               sizeof(double) * b->c.S, hipMemcpyDeviceToDevice, b->stream));
    CBNS_HIP(b, hipStreamSynchronize(b->stream));
  }
  return CBNS_OK;
}

int cbns_set_profiling(cbns_block* b, int32_t on) {
  if (!b) return CBNS_ERR_ARG;
  b->prof.on = on != 0;
  return CBNS_OK;
}

int cbns_kernel_time(cbns_block* b, int32_t k, int32_t reset, double* ms, int64_t* launches) {
  if (!b || k < 0 || k >= CBNS_K_COUNT) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  double tot = 0.0;
  for (auto& ev : b->prof.live[k]) {
    float t = 0.0f;
    CBNS_HIP(b, hipEventElapsedTime(&t, ev.first, ev.second));
    tot += t;
  }
  if (ms) *ms = tot;
  if (launches) *launches = (int64_t)b->prof.live[k].size();
  if (reset) {
    for (auto& ev : b->prof.live[k]) b->prof.pool.push_back(ev);
    b->prof.live[k].clear();
  }
  return CBNS_OK;
}

int cbns_synchronize(cbns_block* b) {
  if (!b) return CBNS_ERR_ARG;
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  return CBNS_OK;
}

}  // extern "C"
