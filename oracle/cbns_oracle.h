/*
 * cbns_oracle.h -- CPU oracle for one explicit SSP-RK3 step of the curvilinear
 * single-block compressible Navier-Stokes solver (chenbei102/cbns3d_cuda).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (cbns3d_cuda_amd/) never links or calls it.
 *
 * Every function restates, in plain C and in the reference's operation order,
 * a reference device function or kernel; each cites the file:line it follows.
 * Parity pins (what checks each function) are listed in the header comment of
 * cbns_oracle.c.
 *
 * Storage layout (shared with the GPU library so the tests can hand the same
 * arrays to both): every field is a ghosted SoA array of
 *   (IM+2G) x (JM+2G) x (KM+2G) doubles, i fastest, G = ORC_G = 4,
 * node (i,j,k) (i in [-G, IM+G)) at (i+G) + PX*((j+G) + PY*(k+G)).
 * Multi-component arrays (Q, metrics, ...) are field-major: comp c at c*S.
 */
#ifndef CBNS_ORACLE_H
#define CBNS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_G 4

/* Boundary kind per axis.  REF = the reference's own RAE2822 C-grid rule for
 * that axis (src/primitive_bc_kernel.cu:84-492); PERIODIC = wrap. */
enum { ORC_BC_REF = 0, ORC_BC_PERIODIC = 1 };

/* Same field order as cbns_params in include/cbns.h. */
typedef struct orc_params {
  int IM, JM, KM;
  int i_begin, i_end;
  int viscous;          /* 0: inviscid build (IS_INVISCID), 1: viscous      */
  int visc_centered;    /* 0: reference transverse offset, 1: centred      */
  int bc[3];            /* per axis: ORC_BC_REF or ORC_BC_PERIODIC         */
  double CFL, gamma, Mach, Re, Pr, Pr_t, T_inf, angle_attack;
} orc_params;

typedef struct orc_grid {
  int IM, JM, KM;
  long PX, PY, PZ, S;   /* ghosted extents and field size                    */
  int i_begin, i_end;
  int viscous, visc_centered;
  int bc[3];
  /* derived constants, src/Block3d.cpp:99-113 */
  double CFL, gamma, Mach, Re, Pr, Pr_t, T_inf, angle_attack;
  double Mach2, gM2, Re_inv, Pr_inv, Pr_t_inv, gam1, gam1_inv, gPr, p_inf;
  double C_T_inf, C_dt_v;
} orc_grid;

/* metric component slots in the 10-field metric array */
enum { ORC_XI_X = 0, ORC_XI_Y, ORC_XI_Z, ORC_ETA_X, ORC_ETA_Y, ORC_ETA_Z,
       ORC_ZETA_X, ORC_ZETA_Y, ORC_ZETA_Z, ORC_JAC };
/* primitive slots in the 6-field primitive array */
enum { ORC_RHO = 0, ORC_U, ORC_V, ORC_W, ORC_P, ORC_T };

int  orc_grid_init(orc_grid* g, const orc_params* p);
long orc_idx(const orc_grid* g, int i, int j, int k);

/* ---- per-point functions ------------------------------------------------ */
double orc_weno5(double um2, double um1, double u0, double up1, double up2);
void orc_eigenvectors_roe(const orc_grid* g, const double pl[5], const double pr[5],
                          double nx, double ny, double nz, double R[25], double L[25]);
void orc_rec_weno5(const orc_grid* g, const double f[25], const double R[25],
                   const double L[25], double prim[5]);
void orc_lf_flux(const orc_grid* g, const double pr[5], const double pl[5],
                 double nx, double ny, double nz, double flux[5]);
void orc_face_flux(const orc_grid* g, const double q[30], const double pl[5],
                   const double pr[5], double nx, double ny, double nz, double flux[5]);

/* ---- grid drivers ------------------------------------------------------- */
void orc_gradient_nodes(int IM, int JM, int KM, const double* f,
                        double* fx, double* fy, double* fz);
int  orc_calc_metrics(const orc_grid* g, const double* x, const double* y,
                      const double* z, const double period[9], double* met);
void orc_bc_primitive(const orc_grid* g, double* Q, double* prim, const double* met);
void orc_numerical_flux(const orc_grid* g, const double* Q, const double* prim,
                        const double* met, double* flux);
void orc_viscous(const orc_grid* g, const double* prim, const double* met,
                 double* mu, double* dfv);
void orc_update_rk3(const orc_grid* g, double dt, int stage, const double* met,
                    const double* flux, const double* dfv, double* Q, double* Qp);
double orc_calc_dt(const orc_grid* g, const double* prim, const double* mu,
                   const double* met);
int  orc_rk3_stage(const orc_grid* g, double dt, int stage, double* Q, double* Qp,
                   double* prim, const double* met, double* mu);
int  orc_rk3_step(const orc_grid* g, double t_end, double* t_cur, double* dt_out,
                  double* Q, double* Qp, double* prim, const double* met, double* mu);
void orc_conservative(const orc_grid* g, const double* prim, double* Q);
int  orc_threads(void);

/* ---- learned troubled-face indicator (orc_learned.c) -------------------- */
enum {
  ORC_MLP_NFEAT = 30, ORC_MLP_IN = 32, ORC_MLP_HID = 32,
  ORC_MLP_W1 = 0, ORC_MLP_B1 = 1024, ORC_MLP_W2 = 1056, ORC_MLP_B2 = 2080,
  ORC_MLP_W3 = 2112, ORC_MLP_B3 = 2144, ORC_MLP_NPARAM = 2145
};
float  orc_bf16_round(float x);
void   orc_face_features(const orc_grid* g, const double q[30], const double pl[5],
                         const double pr[5], double nx, double ny, double nz, float feat[32]);
void   orc_weno5_weights(double a, double b, double c, double d, double e, double om[3]);
double orc_face_label(const orc_grid* g, const double q[30], const double pl[5],
                      const double pr[5], double nx, double ny, double nz);
void   orc_mlp_forward(const float* params, const float* X, long n, int bf16, float* logits);
double orc_mlp_loss_grad(const float* params, const float* X, const float* y, long n,
                         double w_pos, int bf16, double* grad);
void   orc_face_flux_linear(const orc_grid* g, const double q[30], double nx, double ny,
                            double nz, double flux[5]);
void   orc_numerical_flux_hybrid(const orc_grid* g, const double* Q, const double* prim,
                                 const double* met, const unsigned char* flags, double* flux);
int    orc_rk3_stage_hybrid(const orc_grid* g, double dt, int stage, double* Q, double* Qp,
                            double* prim, const double* met, double* mu,
                            const unsigned char* flags);
void   orc_adam(float* params, const double* grad, double* m, double* v, long np, double lr,
                double b1, double b2, double eps, int t);

#ifdef __cplusplus
}
#endif
#endif
