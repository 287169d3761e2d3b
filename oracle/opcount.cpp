// opcount.cpp -- counts the floating-point operations of one face of the
// reference's inviscid flux (normal average + calc_eigenvectors_roe +
// rec_weno5 x2 + lf_flux, src/rec_riemann_kernel.cu:44-89) by running the
// oracle's per-point functions over an operation-counting scalar.
// TEST INFRASTRUCTURE ONLY: the count is the algorithmic work per face that
// the benchmark's roofline divides by the measured kernel time.
// Convention: each IEEE + - * / sqrt is one operation; fabs and comparisons
// are counted separately and not included in "flops".
#include <cmath>
#include <cstdio>

#include "cbns_oracle.h"

namespace {
long n_add, n_mul, n_div, n_sqrt, n_abs, n_cmp;

struct Cnt {
  double v;
  Cnt() : v(0) {}
  Cnt(double x) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { n_add++; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { n_add++; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { n_mul++; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { n_div++; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt& operator+=(Cnt& a, Cnt b) { a = a + b; return a; }
inline Cnt& operator-=(Cnt& a, Cnt b) { a = a - b; return a; }
inline Cnt& operator*=(Cnt& a, Cnt b) { a = a * b; return a; }
inline Cnt& operator/=(Cnt& a, Cnt b) { a = a / b; return a; }
inline bool operator>(Cnt a, Cnt b) { n_cmp++; return a.v > b.v; }
inline Cnt sqrt(Cnt a) { n_sqrt++; return Cnt(std::sqrt(a.v)); }
inline Cnt fabs(Cnt a) { n_abs++; return Cnt(std::fabs(a.v)); }
}  // namespace

#define ORC_REAL Cnt
#define ORC_FN(name) cnt_##name
#define NEQ 5
#include "orc_point.inc"

static void count_face(const orc_grid* g, double nx, double ny, double nz, const char* tag) {
  n_add = n_mul = n_div = n_sqrt = n_abs = n_cmp = 0;
  Cnt q[30], pl[5], pr[5], f[5];
  for (int s = 0; s < 6; s++) {
    q[s * 5 + 0] = 1.0 + 0.01 * s; q[s * 5 + 1] = 0.3; q[s * 5 + 2] = -0.2;
    q[s * 5 + 3] = 0.1; q[s * 5 + 4] = 2.5 + 0.02 * s;
  }
  const double p0[5] = {1.0, 0.3, -0.2, 0.1, 0.9}, p1[5] = {1.01, 0.29, -0.21, 0.1, 0.91};
  for (int c = 0; c < 5; c++) { pl[c] = p0[c]; pr[c] = p1[c]; }
  // face vector 0.5 (J grad xi|_m + J grad xi|_m+1), src/rec_riemann_kernel.cu:44-49
  Cnt Ja = 1.0, Jb = 1.1, ax = nx, bx = nx, ay = ny, by = ny, az = nz, bz = nz;
  Cnt n_x = Cnt(0.5) * (ax * Ja + bx * Jb);
  Cnt n_y = Cnt(0.5) * (ay * Ja + by * Jb);
  Cnt n_z = Cnt(0.5) * (az * Ja + bz * Jb);
  cnt_face_flux(g, q, pl, pr, n_x, n_y, n_z, f);
  const long flops = n_add + n_mul + n_div + n_sqrt;
  std::printf("{\"branch\": \"%s\", \"add_sub\": %ld, \"mul\": %ld, \"div\": %ld, \"sqrt\": %ld, "
              "\"abs\": %ld, \"cmp\": %ld, \"flops\": %ld}\n",
              tag, n_add, n_mul, n_div, n_sqrt, n_abs, n_cmp, flops);
}

int main() {
  orc_params p{};
  p.IM = p.JM = p.KM = 8;
  p.gamma = 1.4; p.Mach = 0.5; p.Re = 1600; p.Pr = 0.72; p.Pr_t = 0.9; p.T_inf = 273.15; p.CFL = 0.5;
  orc_grid g;
  orc_grid_init(&g, &p);
  count_face(&g, 0.3, 0.5, -0.2, "general normal");
  count_face(&g, 1.0, 0.0, 0.0, "normal along x");
  return 0;
}
