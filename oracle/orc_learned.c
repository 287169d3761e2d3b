/*
 * orc_learned.c -- CPU oracle for the learned troubled-face indicator
 * (TEST INFRASTRUCTURE ONLY: loaded by the tests, never by the product).
 *
 * The reference has no learned form; what is restated here is the method the
 * GPU library implements on top of the reference's face reconstruction:
 *
 *   label     max over the 5 characteristic fields, both reconstruction sides
 *             and the 3 sub-stencils of |omega_k - d_k|, where omega are the
 *             WENO5-JS nonlinear weights of src/weno.cu:13-44 and d = (0.1,
 *             0.6, 0.3) the linear ones, on W = L q with L the Roe-averaged
 *             left eigenvectors of src/calc_eigenvectors_roe.cu:13-152.
 *   features  per field c, the six stencil values w_c[s] (nodes m-2..m+3)
 *             minus the face mean (w_c[2] + w_c[3]) / 2, divided by
 *             sqrt(sum_k (w_c[k+1] - w_c[k])^2 + 1e-6): 30 numbers, padded
 *             to 32 with zeros.  The 1e-6 is WENO5-JS's own epsilon, below
 *             which the nonlinear weights are the linear ones.
 *   MLP       30(32) -> 32 -> 32 -> 1, ReLU, logit output.  Parameters in
 *             one float array: W1[32][32] (hidden x input), b1[32],
 *             W2[32][32], b2[32], w3[32], b3.  In bf16 mode the oracle
 *             rounds to bfloat16 (round to nearest even) at exactly the
 *             points where the GPU kernels feed the matrix cores: x, W1, h1,
 *             W2 in the forward pass; dz2, dz1 and W2 in the backward pass.
 *   loss      weighted logistic loss, mean over the batch:
 *             y w_pos softplus(-s) + (1 - y) softplus(s).
 *
 * Pins: bf16 rounding against torch's float->bfloat16 conversion; features
 * against a direct numpy evaluation from orc_eigenvectors_roe; labels against
 * constant, linear and step stencils (closed-form weights); gradients in
 * double mode against central finite differences of the loss.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "cbns_oracle.h"

float orc_bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

static void char_values(const orc_grid* g, const double q[30], const double pl[5],
                        const double pr[5], double nx, double ny, double nz, double w[5][6]) {
  double R[25], L[25];
  orc_eigenvectors_roe(g, pl, pr, nx, ny, nz, R, L);
  for (int c = 0; c < 5; c++)
    for (int s = 0; s < 6; s++) {
      double acc = 0.0;
      for (int k = 0; k < 5; k++) acc += L[c * 5 + k] * q[s * 5 + k];
      w[c][s] = acc;
    }
}

void orc_face_features(const orc_grid* g, const double q[30], const double pl[5],
                       const double pr[5], double nx, double ny, double nz, float feat[32]) {
  double w[5][6];
  char_values(g, q, pl, pr, nx, ny, nz, w);
  for (int c = 0; c < 5; c++) {
    double ss = 1.0e-6;
    for (int k = 0; k < 5; k++) ss += (w[c][k + 1] - w[c][k]) * (w[c][k + 1] - w[c][k]);
    const double inv = 1.0 / sqrt(ss), mid = 0.5 * (w[c][2] + w[c][3]);
    for (int s = 0; s < 6; s++) feat[c * 6 + s] = (float)((w[c][s] - mid) * inv);
  }
  feat[30] = feat[31] = 0.0f;
}

/* nonlinear weights of src/weno.cu:13-44 for the stencil a..e */
void orc_weno5_weights(double a, double b, double c, double d, double e, double om[3]) {
  const double eps = 1.0e-6;
  const double d1 = a - 2.0 * b + c, d2 = a - 4.0 * b + 3.0 * c;
  const double d3 = b - 2.0 * c + d, d4 = b - d;
  const double d5 = c - 2.0 * d + e, d6 = 3.0 * c - 4.0 * d + e;
  const double IS1 = (13.0 / 12.0) * d1 * d1 + 0.25 * d2 * d2 + eps;
  const double IS2 = (13.0 / 12.0) * d3 * d3 + 0.25 * d4 * d4 + eps;
  const double IS3 = (13.0 / 12.0) * d5 * d5 + 0.25 * d6 * d6 + eps;
  const double w1 = 0.1 / (IS1 * IS1), w2 = 0.6 / (IS2 * IS2), w3 = 0.3 / (IS3 * IS3);
  const double sum = w1 + w2 + w3;
  om[0] = w1 / sum; om[1] = w2 / sum; om[2] = w3 / sum;
}

double orc_face_label(const orc_grid* g, const double q[30], const double pl[5],
                      const double pr[5], double nx, double ny, double nz) {
  static const double dlin[3] = {0.1, 0.6, 0.3};
  double w[5][6];
  char_values(g, q, pl, pr, nx, ny, nz, w);
  double m = 0.0;
  for (int c = 0; c < 5; c++) {
    double om[3];
    orc_weno5_weights(w[c][0], w[c][1], w[c][2], w[c][3], w[c][4], om);
    for (int k = 0; k < 3; k++) m = fmax(m, fabs(om[k] - dlin[k]));
    orc_weno5_weights(w[c][5], w[c][4], w[c][3], w[c][2], w[c][1], om);
    for (int k = 0; k < 3; k++) m = fmax(m, fabs(om[k] - dlin[k]));
  }
  return m;
}

/* ---- MLP ---------------------------------------------------------------- */
enum { NI = ORC_MLP_IN, NH = ORC_MLP_HID };

static double rnd(double v, int bf16) { return bf16 ? (double)orc_bf16_round((float)v) : v; }

typedef struct {
  double x[NI], z1[NH], h1[NH], z2[NH], h2[NH], s;
} fwd_t;

static void forward(const float* P, const float* x, int bf16, fwd_t* f) {
  const float *W1 = P + ORC_MLP_W1, *b1 = P + ORC_MLP_B1, *W2 = P + ORC_MLP_W2;
  const float *b2 = P + ORC_MLP_B2, *w3 = P + ORC_MLP_W3;
  for (int i = 0; i < NI; i++) f->x[i] = rnd(x[i], bf16);
  for (int h = 0; h < NH; h++) {
    double acc = 0.0;
    for (int i = 0; i < NI; i++) acc += rnd(W1[h * NI + i], bf16) * f->x[i];
    f->z1[h] = acc + b1[h];
    f->h1[h] = rnd(f->z1[h] > 0.0 ? f->z1[h] : 0.0, bf16);
  }
  for (int h = 0; h < NH; h++) {
    double acc = 0.0;
    for (int i = 0; i < NH; i++) acc += rnd(W2[h * NH + i], bf16) * f->h1[i];
    f->z2[h] = acc + b2[h];
    f->h2[h] = f->z2[h] > 0.0 ? f->z2[h] : 0.0;
  }
  double s = P[ORC_MLP_B3];
  for (int h = 0; h < NH; h++) s += (double)w3[h] * f->h2[h];
  f->s = s;
}

void orc_mlp_forward(const float* params, const float* X, long n, int bf16, float* logits) {
#pragma omp parallel for schedule(static)
  for (long r = 0; r < n; r++) {
    fwd_t f;
    forward(params, X + r * NI, bf16, &f);
    logits[r] = (float)f.s;
  }
}

static double softplus(double s) { return s > 0.0 ? s + log1p(exp(-s)) : log1p(exp(s)); }
static double sigmoid(double s) { return 1.0 / (1.0 + exp(-s)); }

double orc_mlp_loss_grad(const float* params, const float* X, const float* y, long n,
                         double w_pos, int bf16, double* grad) {
  const float* W2 = params + ORC_MLP_W2;
  const float* w3 = params + ORC_MLP_W3;
  double loss = 0.0;
  if (grad) memset(grad, 0, sizeof(double) * ORC_MLP_NPARAM);
  for (long r = 0; r < n; r++) {
    fwd_t f;
    forward(params, X + r * NI, bf16, &f);
    const double t = y[r];
    loss += t * w_pos * softplus(-f.s) + (1.0 - t) * softplus(f.s);
    if (!grad) continue;
    const double ds = (t * w_pos * (sigmoid(f.s) - 1.0) + (1.0 - t) * sigmoid(f.s)) / (double)n;
    grad[ORC_MLP_B3] += ds;
    double dz2[NH], dz2r[NH], dz1[NH];
    for (int h = 0; h < NH; h++) {
      grad[ORC_MLP_W3 + h] += ds * f.h2[h];
      dz2[h] = f.z2[h] > 0.0 ? ds * (double)w3[h] : 0.0;
      dz2r[h] = rnd(dz2[h], bf16);
      grad[ORC_MLP_B2 + h] += dz2[h];
      for (int i = 0; i < NH; i++) grad[ORC_MLP_W2 + h * NH + i] += dz2r[h] * f.h1[i];
    }
    for (int i = 0; i < NH; i++) {
      double acc = 0.0;
      for (int h = 0; h < NH; h++) acc += rnd(W2[h * NH + i], bf16) * dz2r[h];
      dz1[i] = f.z1[i] > 0.0 ? acc : 0.0;
      grad[ORC_MLP_B1 + i] += dz1[i];
      const double d = rnd(dz1[i], bf16);
      for (int k = 0; k < ORC_MLP_NFEAT; k++) grad[ORC_MLP_W1 + i * NI + k] += d * f.x[k];
    }
  }
  return loss / (double)n;
}

void orc_adam(float* params, const double* grad, double* m, double* v, long np, double lr,
              double b1, double b2, double eps, int t) {
  const double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
  for (long i = 0; i < np; i++) {
    m[i] = b1 * m[i] + (1.0 - b1) * grad[i];
    v[i] = b2 * v[i] + (1.0 - b2) * grad[i] * grad[i];
    params[i] = (float)(params[i] - lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps));
  }
}
