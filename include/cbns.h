/*
 * cbns.h -- C ABI of the MI355X-native RK3 step of a curvilinear single-block
 * compressible Navier-Stokes solver (drop-in for the block3d_cuda launcher
 * API of chenbei102/cbns3d_cuda, the include/ headers).
 *
 * Conventions
 *  - Every entry point returns 0 (CBNS_OK) or a CBNS_ERR_* code; nothing
 *    exits the process or throws across this boundary.  cbns_last_error()
 *    gives the message of the last failure on a block.
 *  - One block = one device allocation set + one HIP stream.  Calls on a
 *    block are stream-ordered; entry points that hand data to the host
 *    (cbns_get_*, cbns_calc_dt, cbns_rk3_step) synchronise the stream.
 *    A block is thread-compatible, not thread-safe.
 *  - Host arrays at the edge use the reference's layouts:
 *      node arrays      IM*JM*KM doubles, i fastest (x, y, z, metrics, f)
 *      ghost arrays     (IM+4)*(JM+4)*(KM+4) doubles, 2 ghost layers
 *      conservative Q   5*(IM+4)*(JM+4)*(KM+4) doubles, component fastest
 *  - Internally every field is a structure-of-arrays FP64 array with
 *    CBNS_G ghost layers; cbns_get_field/cbns_set_field expose that layout.
 */
#ifndef CBNS_H
#define CBNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBNS_G 4          /* ghost layers in device storage          */
#define CBNS_NG_REF 2     /* ghost layers of the reference's arrays  */

enum {
  CBNS_OK = 0,
  CBNS_ERR_ARG = 1,       /* bad argument or size                     */
  CBNS_ERR_HIP = 2,       /* HIP runtime error                        */
  CBNS_ERR_NAN = 3,       /* time step is NaN (non-physical state)    */
  CBNS_ERR_MESH = 4,      /* non-positive Jacobian                    */
  CBNS_ERR_STATE = 5,     /* call order (e.g. no mesh yet)            */
  CBNS_ERR_NO_DEVICE = 6, /* no HIP device visible                    */
  CBNS_ERR_SIZE = 7       /* input beyond a fixed capacity            */
};

/* boundary kind per computational axis */
enum { CBNS_BC_REF = 0,        /* the reference's RAE2822 C-grid rule for the axis */
       CBNS_BC_PERIODIC = 1 }; /* periodic wrap                                     */

/* stage buffer selector: the reference's data.Q / data.Q_p */
enum { CBNS_BUF_Q = 0, CBNS_BUF_QP = 1 };

/* arithmetic form of the fused stage: FP64 or FP32, optionally with the
 * learned hybrid (CBNS_FORM_HYBRID bit; needs cbns_set_indicator first) */
enum { CBNS_FORM_FP64 = 0,     /* plain FP64                                 */
       CBNS_FORM_FP32 = 1,     /* FP32 face/viscous arithmetic, FP64 state   */
       CBNS_FORM_HYBRID = 2 }; /* flagged faces WENO5, the rest linear       */

/* field ids for cbns_get_field / cbns_set_field (ghosted SoA, CBNS_G) */
enum {
  CBNS_F_Q = 0,        /* 0..4   conservative state Q                */
  CBNS_F_QP = 5,       /* 5..9   stage state Q_p                     */
  CBNS_F_PRIM = 10,    /* 10..15 rho,u,v,w,p,T                       */
  CBNS_F_MET = 16,     /* 16..25 xi_x..zeta_z, Jac                   */
  CBNS_F_MU = 26,      /* 26     Sutherland viscosity (unscaled)     */
  CBNS_F_FLUX = 27,    /* 27..41 face fluxes Ep,Fp,Gp (5 each)        */
  CBNS_F_DFV = 42,     /* 42..45 viscous divergence                  */
  CBNS_F_COUNT = 46
};

/* Runtime parameters: the reference's input.txt / .ini keys
 * (src/Block3d.cpp:44-92, python/rae2822.ini:1-26) plus the switches that
 * replace its compile-time IS_INVISCID and its fixed boundary set. */
typedef struct cbns_params {
  int32_t IM, JM, KM;
  int32_t i_begin, i_end;     /* wall extent on eta=0 (reference axes)      */
  int32_t viscous;            /* 0: IS_INVISCID build, 1: viscous          */
  int32_t visc_centered;      /* 0: reference transverse offset of the
                                 viscous divergence, 1: centred            */
  int32_t bc[3];              /* CBNS_BC_* per axis                        */
  double CFL, gamma, Mach, Re, Pr, Pr_t, T_inf, angle_attack;
} cbns_params;

typedef struct cbns_block cbns_block;

/* copy_block_info + allocate_mem: include/copy_block_info.h:9,
 * include/allocate_mem.h:9 */
int cbns_create(const cbns_params* params, cbns_block** out);
/* free_mem: include/free_mem.h:9 */
void cbns_destroy(cbns_block* blk);
const char* cbns_last_error(const cbns_block* blk);

/* calc_metrics: include/calc_metrics.h:9-10.  x,y,z node arrays.  period
 * (9 doubles, may be NULL when no axis is periodic): period[3*a+c] is the
 * shift of coordinate c across one period of axis a. */
int cbns_set_mesh(cbns_block* blk, const double* x, const double* y, const double* z,
                  const double* period);
/* get_metrics: include/get_metrics.h:9-13 (node arrays) */
int cbns_get_metrics(cbns_block* blk, double* xi_x, double* xi_y, double* xi_z,
                     double* eta_x, double* eta_y, double* eta_z,
                     double* zeta_x, double* zeta_y, double* zeta_z, double* Jac);
/* calc_gradient: include/calc_gradient.h:9-10 (node arrays IM*JM*KM) */
int cbns_calc_gradient(int32_t IM, int32_t JM, int32_t KM, const double* f,
                       double* fx, double* fy, double* fz);
/* No counterpart in the reference: cbns_calc_gradient for callers that
 * cannot allocate device memory (HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and results (bit for bit); the
 * device staging is a fixed array of the library, so IM*JM*KM may not
 * exceed CBNS_NOALLOC_MAX_NODES (else CBNS_ERR_SIZE, before any device
 * work).  Calls that overlap on one GPU share that array and race. */
#define CBNS_NOALLOC_MAX_NODES (1L << 20)
int cbns_calc_gradient_noalloc(int32_t IM, int32_t JM, int32_t KM, const double* f,
                               double* fx, double* fy, double* fz);

/* initial_condition: include/initial_condition.h:9 (freestream primitives) */
int cbns_initial_condition(cbns_block* blk);
/* calc_conservative: include/calc_conservative.h:9-10 (Q from primitives,
 * copied into Q_p) */
int cbns_calc_conservative(cbns_block* blk);
/* set_conservative / get_conservative: include/set_conservative.h:9-10,
 * include/get_conservative.h:9-10 (AoS, 2 ghost layers) */
int cbns_set_conservative(cbns_block* blk, const double* Q);
int cbns_get_conservative(cbns_block* blk, double* Q);
/* calc_primitive: include/calc_primitive.h:9-10 (Q -> rho,u,v,w,p,T) */
int cbns_calc_primitive(cbns_block* blk);
/* get_primitive: include/get_primitive.h:9-10 (2 ghost layers) */
int cbns_get_primitive(cbns_block* blk, double* rho, double* u, double* v,
                       double* w, double* p);

/* set_bc_calc_primitive: include/set_bc_calc_primitive.h:9-10 */
int cbns_set_bc_calc_primitive(cbns_block* blk, int32_t buf);
/* calc_dt: include/calc_dt.h:9 */
int cbns_calc_dt(cbns_block* blk, double* dt);
/* calc_numerical_flux: include/calc_numerical_flux.h:9-10 */
int cbns_calc_numerical_flux(cbns_block* blk, int32_t buf);
/* calc_viscous_terms: include/calc_viscous_terms.h:9 */
int cbns_calc_viscous_terms(cbns_block* blk);
/* calc_viscous_flux_contribution: include/calc_viscous_flux_contribution.h:9 */
int cbns_calc_viscous_flux_contribution(cbns_block* blk);
/* update_rk3: include/time_integration.h:9-10 */
int cbns_update_rk3(cbns_block* blk, double dt, int32_t stage);

/* Fused stage: calc_numerical_flux + calc_viscous_flux_contribution +
 * update_rk3 + set_bc_calc_primitive (+ calc_viscous_terms) of one stage of
 * src/Block3d.cpp:249-287, run as the MI355X sweep kernels. */
int cbns_rk3_stage(cbns_block* blk, double dt, int32_t stage);
/* One pass of the time loop src/Block3d.cpp:239-287: dt (clamped to t_end),
 * t_cur += dt, three fused stages.  dt is formed on the device (bit-equal to
 * cbns_calc_dt) and read there by the stages; a NaN dt returns
 * CBNS_ERR_NAN with the state unchanged. */
int cbns_rk3_step(cbns_block* blk, double t_end, double* t_cur, double* dt_out);
/* arithmetic form of cbns_rk3_stage / cbns_rk3_step (CBNS_FORM_*) */
int cbns_set_form(cbns_block* blk, int32_t form);

/* raw ghosted SoA field access (tests, tools) */
int cbns_get_field(cbns_block* blk, int32_t field, double* host);
int cbns_set_field(cbns_block* blk, int32_t field, const double* host);
int cbns_synchronize(cbns_block* blk);
/* doubles per ghosted field, (IM+2G)*(JM+2G)*(KM+2G) */
int64_t cbns_field_size(const cbns_block* blk);

/* Device timing of the kernels of cbns_rk3_stage / cbns_rk3_step (no
 * counterpart in the reference, which times only the whole loop with
 * std::chrono, src/Block3d.cpp:235-299).  While profiling is on, each launch
 * is bracketed by HIP events on the block's stream; cbns_kernel_time returns
 * the summed milliseconds and launch count of one kernel id since the last
 * reset and reset=1 clears it. */
enum {
  CBNS_K_XFACES = 0,     /* xi-face fluxes                          */
  CBNS_K_MARCH_ETA = 1,  /* eta sweep (+ Fv's divergence)           */
  CBNS_K_MARCH_ZETA = 2, /* zeta + xi sweep (+ Gv's, Ev's divergence) + RK stage combination */
  CBNS_K_VISC_FLUX = 3,  /* viscous fluxes Ev,Fv,Gv at the nodes     */
  CBNS_K_VISC_DIV = 4,   /* viscous divergence pass (reference-decomposed calls; the fused stage has none) */
  CBNS_K_BC = 5,         /* boundary / ghost fill                   */
  CBNS_K_DT = 6,         /* CFL time step reduction                 */
  CBNS_K_INDICATOR = 7,  /* learned troubled-face flags (hybrid)    */
  CBNS_K_WENO_LIST = 8,  /* WENO fluxes of the troubled faces (hybrid) */
  CBNS_K_COUNT = 9
};
int cbns_set_profiling(cbns_block* blk, int32_t on);
int cbns_kernel_time(cbns_block* blk, int32_t kernel, int32_t reset, double* ms, int64_t* launches);

/* ---- learned troubled-face indicator (no counterpart in the reference) ----
 * A 30(32)->32->32->1 ReLU network reads, per face, the 30 normalised
 * characteristic stencil differences and predicts where the WENO5 nonlinear
 * weights of src/weno.cu:13-44 leave the linear ones.  Parameters: one float
 * array of CBNS_MLP_NPARAM: W1[32][32] (hidden x input; input columns 30, 31
 * are padding and stay zero), b1[32], W2[32][32], b2[32], w3[32], b3.
 * Arguments named d_* are DEVICE pointers on the block's / network's device;
 * the caller makes their contents ready (synchronises its own stream) before
 * the call; every call below returns after its device work is complete. */
#define CBNS_MLP_NFEAT 32    /* floats per feature row (30 used)  */
#define CBNS_MLP_NPARAM 2145 /* floats of the parameter array      */

/* faces of one axis the flux kernels compute (the sample count of
 * cbns_face_samples): (faces along axis) x (updated nodes of the other two) */
int cbns_face_count(cbns_block* blk, int32_t axis, int64_t* n);
/* Training samples from the plain FP64 solver state of buffer `buf`
 * (CBNS_BUF_*): d_feat[n][32] features, d_label[n] = max |omega_k - d_k|
 * over characteristic fields, sides and sub-stencils; sample order i fastest
 * over the face box. */
int cbns_face_samples(cbns_block* blk, int32_t axis, int32_t buf, float* d_feat, float* d_label);
/* Load the indicator used by CBNS_FORM_HYBRID: a face is troubled when its
 * logit exceeds `threshold`; troubled faces dilated by one cell along their
 * axis run characteristic WENO5, the others the linear 5th-order upwind
 * reconstruction.  options bit 0: evaluate the indicator at every stage
 * (default: once per step, on the stage-1 input); bit 1: count troubled faces
 * (cbns_flag_count); bit 2: keep the flags for floor(1/CFL) steps (a wave
 * moves at most CFL cells per step along each axis, so within that many steps
 * it stays inside the one-cell dilation).  The troubled faces of each axis
 * are listed when the flags change and their WENO fluxes formed in one pass
 * per stage; lists hold a third of an axis' faces, and a step whose
 * troubled faces exceed that runs WENO at every face (the plain form's
 * marches; the indicator then runs only every 8 steps, until the lists fit
 * again) -- bit 3: lists that hold every face (never that fallback). */
int cbns_set_indicator(cbns_block* blk, const float* params, float threshold, int32_t options);
/* troubled faces (before dilation) per axis found by the last indicator pass
 * among the faces the flux kernels compute (cbns_face_count); needs options
 * bit 1 */
int cbns_flag_count(cbns_block* blk, int64_t counts[3]);
/* the troubled-face bytes of the last indicator pass: 3 ghosted fields of
 * uint8 (axis-major; face m of an axis, between nodes m and m+1, at node m),
 * 3 * cbns_field_size bytes */
int cbns_get_flags(cbns_block* blk, uint8_t* host);

/* The network alone, for training on the GPU (forward, backward and Adam in
 * HIP kernels on the bf16 matrix cores with fp32 accumulation). */
typedef struct cbns_mlp cbns_mlp;
int  cbns_mlp_create(cbns_mlp** out);
void cbns_mlp_destroy(cbns_mlp* mlp);
const char* cbns_mlp_last_error(const cbns_mlp* mlp);
/* host arrays of CBNS_MLP_NPARAM floats; set also clears the Adam moments */
int cbns_mlp_set_params(cbns_mlp* mlp, const float* params);
int cbns_mlp_get_params(cbns_mlp* mlp, float* params);
/* d_logit[n] = network(d_feat[n][32]) */
int cbns_mlp_forward(cbns_mlp* mlp, const float* d_feat, int64_t n, float* d_logit);
/* weighted logistic loss mean_i [t_i w_pos softplus(-s_i) + (1-t_i) softplus(s_i)]
 * over targets d_target[n] in {0,1}; its gradient is kept for
 * cbns_mlp_adam_step and, when `grad` (host, CBNS_MLP_NPARAM) is given, copied
 * out. */
int cbns_mlp_loss_grad(cbns_mlp* mlp, const float* d_feat, const float* d_target, int64_t n,
                       float w_pos, double* loss, float* grad);
/* one Adam step on the last gradient */
int cbns_mlp_adam_step(cbns_mlp* mlp, float lr, float beta1, float beta2, float eps);

/* Device buffers of this library's HIP runtime, for callers that have no
 * HIP runtime of their own (or one from another build, as a Python process
 * with a framework that bundles its own: device pointers must not cross
 * runtimes).  All calls are synchronous on the current device. */
int cbns_dev_alloc(int64_t bytes, void** d);
int cbns_dev_free(void* d);
int cbns_dev_upload(void* d, const void* host, int64_t bytes);
int cbns_dev_download(void* host, const void* d, int64_t bytes);
/* d_dst[r][0..row) = d_src[d_idx[r]][0..row) for r < n (float rows) */
int cbns_gather_rows(const float* d_src, const int64_t* d_idx, int64_t n, int32_t row, float* d_dst);

#ifdef __cplusplus
}
#endif
#endif
