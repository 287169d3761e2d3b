// k_fused.hip -- one RK3 stage as sweep kernels (cbns_rk3_stage):
//
//   k_march<1>   lanes along i, each thread marches a chunk of nodes along
//                eta holding the 6-node stencil and the previous face flux in
//                registers: each node is loaded once per sweep and each face
//                is computed once.  One of the two marches (xi_axis: the
//                zeta march) also computes, at every node, the xi-face on its
//                right; the left one comes from the lane below by a wavefront
//                shuffle.  In the zeta march the xi-stencils come from the
//                neighbouring lanes' windows and a wavefront updates 58
//                nodes; in the eta march (the other placement) a wavefront
//                updates 63 nodes, its lane 0 only supplies that face, and the
//                xi-stencils are loaded (march_width).  Writes rhs = dF, or
//                dE + dF with the xi-faces here.
//                (FUSE_XI=0: a separate k_xfaces pass writes the xi-face
//                fluxes to HBM and the march reads them back.)
//   [viscous]    k_visc_fluxes (k_viscous.hip) writes the viscous node
//                fluxes before the marches; each march subtracts their
//                4th-order face values from its face fluxes (visc_face), so
//                the divergence needs no pass of its own
//   k_march<2>   the same march along zeta; adds dG, divides by J and
//                applies the Shu-Osher stage combination, writing the stage
//                output and its primitives (the boundary pass then only
//                fills boundaries and ghosts).
//
// Node primitives for the Roe average are recomputed from Q (the
// primitive array holds exactly f(Q) at every node a face reads:
// boundary rules write Q first and derive the primitives from it).
#include "cbns_internal.hpp"
#include "face_math.hpp"
#include "learned.hpp"

namespace cbns {

namespace {

constexpr int MARCH_X = 64;  // lanes along i
constexpr int MARCH_Y = 4;   // lines per workgroup along the other axis
#ifndef ZETA_ROWS
#define ZETA_ROWS 4  // the zeta march's lines along j per workgroup
#endif
// workgroup size of a march (the eta march: MARCH_Y wavefronts on one line)
template <int AX>
constexpr int march_rows() { return AX == 2 ? ZETA_ROWS : MARCH_Y; }
#ifndef MARCH_CHUNK
#define MARCH_CHUNK 32
#endif
constexpr int CHUNK = MARCH_CHUNK;  // nodes per thread along the march axis
#ifndef WPE_X
#define WPE_X 3  // waves per SIMD the x-face kernel is compiled for
#endif
#ifndef WPE_M
#define WPE_M 2  // waves per SIMD the march kernels are compiled for
#endif
#ifndef WPE_H
#define WPE_H 2  // the same for the hybrid marches (linear faces only; three
                 // waves per SIMD measured 2-4 % slower, four spilled)
#endif
#ifndef FUSE_XI
#define FUSE_XI 1  // xi-faces inside a march (no xi-face pass through HBM)
#endif
// The march that also forms the xi-faces: the zeta march, whose RK-update
// traffic then overlaps the extra face work (the eta march alone is bound
// by its face arithmetic, the zeta march by its update traffic).  Measured
// against the eta march: +3 to +5 % of the FP64 step, +9 % of the FP32 step.
// Both placements are supported (xi_axis 1 or 2).
template <class F>
constexpr int xi_axis() { return 2; }
// nodes a march wavefront updates along i (see k_march)
template <int AX, class F>
constexpr int march_width() {
  return !(FUSE_XI && AX == xi_axis<F>()) ? MARCH_X : (AX == 2 ? MARCH_X - 6 : MARCH_X - 1);
}

template <class T>
__device__ __forceinline__ void load_q(const double* __restrict__ Q, long S, int o, T q[5]) {
#pragma unroll
  for (int n = 0; n < 5; n++) q[n] = (T)Q[n * S + o];
}

// The characteristic-WENO fluxes of a hybrid stage's troubled faces
// (k_weno_list): face m of the axis at slot[m], field n of slot s at
// flux[n * stride + s].
struct TroubledFaces {
  const int* slot;
  const double* flux;
  long stride;
};

// Viscous terms in face form.  The reference's viscous contribution is the
// 4th-order central difference (1/12, -2/3, 0, 2/3, -1/12) of each node flux
// along its axis (src/viscous_contribution_kernel.cu:44-67); that difference
// at node m equals h(m+1/2) - h(m-1/2) with the face value
// h(m+1/2) = 7/12 (v[m] + v[m+1]) - 1/12 (v[m-1] + v[m+2]), so the marches
// subtract h from the inviscid face fluxes they already difference: the eta
// march h of Fv, the zeta march h of Gv and, at its xi-faces, h of Ev.  The
// differences are taken on the reference's lines (offset by -1 in both
// transverse indices; 0 for the centred divergence).
template <class T>
__device__ __forceinline__ T visc_face(T vm1, T v0, T v1, T v2) {
  return T(7.0 / 12.0) * (v0 + v1) - T(1.0 / 12.0) * (vm1 + v2);
}

// Face flux of the stage's form at face `face` (node offset).  HYB: the
// linear-weight reconstruction, replaced by the face's characteristic-WENO
// flux from tf where the face is troubled (its flag or a neighbour's along
// the axis: one-cell dilation); the marches hold no WENO arithmetic, so they
// run at the occupancy of the linear faces.  Slots read for faces outside
// the listed range (lanes whose faces no node uses) are clamped into tflux.
template <class F, bool HYB>
__device__ __forceinline__ void stage_face(const GasT<double>& g, const double q[6][5],
                                           double nx, double ny, double nz, bool troubled,
                                           const TroubledFaces& tf, int face, double f[5]) {
  if (!HYB) {
    face_flux_form<F>(g, q, nx, ny, nz, f);
    return;
  }
  face_linear<F>(g, q, nx, ny, nz, f);
  if (troubled) {
    const long sl = min(max(tf.slot[face], 0), (int)(tf.stride - 1));
#pragma unroll
    for (int n = 0; n < 5; n++) f[n] = tf.flux[n * tf.stride + sl];
  }
}

template <class F, bool HYB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE_X, WPE_X))) k_xfaces(Consts c, Box box, const double* __restrict__ Q,
                                                const double* __restrict__ met,
                                                const uint8_t* __restrict__ flags,
                                                TroubledFaces tf,
                                                double* __restrict__ fx) {
  int i, j, k;
  if (!box_node3(box, i, j, k)) return;
  const long S = c.S;
  const int a = idx(c, i, j, k), b = a + 1;
  const double Ja = met[JAC * S + a], Jb = met[JAC * S + b];
  const double nx = 0.5 * (met[XI_X * S + a] * Ja + met[XI_X * S + b] * Jb);
  const double ny = 0.5 * (met[XI_Y * S + a] * Ja + met[XI_Y * S + b] * Jb);
  const double nz = 0.5 * (met[XI_Z * S + a] * Ja + met[XI_Z * S + b] * Jb);
  double q[6][5], f[5];
#pragma unroll
  for (int s = 0; s < 6; s++) load_q(Q, S, a + (s - 2), q[s]);
  const GasT<double> g{c.gamma, c.gam1, c.gam1_inv};
  const bool troubled = HYB && (flags[a - 1] | flags[a] | flags[a + 1]);
  stage_face<F, HYB>(g, q, nx, ny, nz, troubled, tf, a, f);
#pragma unroll
  for (int n = 0; n < 5; n++) fx[n * S + a] = f[n];
}

struct StageArgs {
  double dt;
  const double* dtp;    // dt on the device (the folded time step) or null: use dt
  int stage;
  double* prim;         // primitives of the stage output (AX == 2)
  const double* Qn;     // level-n state (stages 2, 3)
  const double* Qin;    // stage input (stencil source)
  double* Qout;         // stage output
  const double* fx;     // xi-face fluxes (AX == 1)
  double* rhs;          // dE + dF (+ viscous)
  const void* vf;       // viscous node fluxes Ev, Fv, Gv (12 fields of the face type) or null
  const uint8_t* flags; // troubled faces of the march axis (HYB)
  const uint8_t* flags_x;  // troubled xi-faces (HYB, fused xi-faces)
  int pmask;            // AX == 2: primitives to write (bit n: field n of rho, u, v, w, p, T)
  TroubledFaces tf;     // WENO fluxes of the march axis' troubled faces (HYB)
  TroubledFaces tf_x;   // and of the troubled xi-faces (HYB, fused xi-faces)
};

// AX = 1: march along eta, rhs = dE + dF.  AX = 2: march along zeta, finish.
template <int AX, class F, bool HYB>
__global__ void __launch_bounds__(MARCH_X * march_rows<AX>())
__attribute__((amdgpu_waves_per_eu(HYB ? WPE_H : WPE_M, HYB ? WPE_H : WPE_M)))
k_march(Consts c, StageArgs sa,
                                                            const double* __restrict__ met) {
  constexpr bool XI = AX == xi_axis<F>() && FUSE_XI;
  // the update's time step; a NaN step (non-physical state) leaves the
  // stage output unwritten, and the caller reports it
  const double dt = AX == 2 && sa.dtp ? *sa.dtp : sa.dt;
  if (AX == 2 && !(dt == dt)) return;
  const int ax_t = AX == 1 ? 2 : 1;  // transverse axis other than i
  const int lane = threadIdx.x;
  // XI in the eta march: a wavefront updates 63 nodes, its lane 0 sits one
  // node to the left and loads its xi-stencils.  XI in the zeta march (XSH):
  // the xi-stencils come from the neighbouring lanes' march windows by
  // wavefront shuffles, so a wavefront holds 3 nodes on either side of the
  // 58 it updates (the zeta march is bound by its memory instructions, the
  // eta march by arithmetic).
  constexpr bool XSH = XI && AX == 2;
  constexpr int XL = XSH ? 3 : (XI ? 1 : 0);  // lanes left of the first updated node
  // eta march: the workgroup's wavefronts take adjacent tiles of one line
  // along i, so each step along eta reads whole rows of the fields (-9 % for
  // the FP64 eta march, -12 % FP32).  zeta march: wavefronts on adjacent
  // lines (its 58-node tiles would leave idle wavefronts on a line)
  constexpr bool ROWS = AX == 1;
  const int i = upd_lo(c, 0) - XL +
                (ROWS ? blockIdx.x * MARCH_Y + threadIdx.y : blockIdx.x) * march_width<AX, F>() +
                lane;
  const int t = upd_lo(c, ax_t) + (ROWS ? blockIdx.y : blockIdx.y * march_rows<AX>() + threadIdx.y);
  if (i > upd_hi(c, 0) + (XSH ? 3 : 0) || t > upd_hi(c, ax_t)) return;
  const bool own = lane >= XL && lane < XL + march_width<AX, F>() && i <= upd_hi(c, 0);
  const int mlo = upd_lo(c, AX), mhi = upd_hi(c, AX);
  const int m0 = mlo + blockIdx.z * CHUNK;
  if (m0 > mhi) return;
  const int m1 = min(m0 + CHUNK, mhi + 1);
  const long S = c.S;
  const int st = AX == 1 ? c.PX : c.PXY;
  const int base = AX == 1 ? idx(c, i, 0, t) : idx(c, i, t, 0);  // node m at base + m*st
  const double* __restrict__ Q = sa.Qin;
  const double* __restrict__ nX = met + (3 * AX + 0) * S;
  const double* __restrict__ nY = met + (3 * AX + 1) * S;
  const double* __restrict__ nZ = met + (3 * AX + 2) * S;
  const double* __restrict__ J = met + JAC * S;
  const GasT<double> g{c.gamma, c.gam1, c.gam1_inv};

  double q[6][5];  // nodes m-2 .. m+3 for the face m+1/2
#pragma unroll
  for (int s = 0; s < 6; s++) load_q(Q, S, base + (m0 - 3 + s) * st, q[s]);
  // J grad xi_AX at the face's two nodes
  int o = base + (m0 - 1) * st;
  double ja = J[o], jb = J[o + st];
  double ax = nX[o] * ja, ay = nY[o] * ja, az = nZ[o] * ja;
  double bx = nX[o + st] * jb, by = nY[o + st] * jb, bz = nZ[o + st] * jb;
  double fprev[5];
  // troubled flags of faces m-1, m, m+1 (face m between nodes m and m+1)
  const uint8_t* __restrict__ fl = sa.flags;
  uint8_t t0 = 0, t1 = 0, t2 = 0;
  if (HYB) {
    t0 = fl[base + (m0 - 2) * st];
    t1 = fl[base + (m0 - 1) * st];
    t2 = fl[base + m0 * st];
  }
  // Precision form (float faces, not hybrid): the sweep carries the float
  // differences e[k] = q[k+1] - q[k] of the window's nodes and the Roe
  // inputs of nodes m (nl) and m+1 (nr), each formed once per node, instead
  // of five double differences and two double face states per face.
  constexpr bool DC = sizeof(F) != sizeof(double) && !HYB;
  const GasT<F> gt{(F)c.gamma, (F)c.gam1, (F)c.gam1_inv};
  F e[5][5];
  NodeT<F> nl{}, nr{};
  if constexpr (DC) {
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int n = 0; n < 5; n++) e[k][n] = (F)(q[k + 1][n] - q[k][n]);
    nl = node_t<F>(g, q_to_side(g, q[2]));
    nr = node_t<F>(g, q_to_side(g, q[3]));
    F d[6][5];
    dev_from_diffs(e, d);
    face_flux_dev<F>(gt, g, d, q[2], nl, nr, 0.5 * (ax + bx), 0.5 * (ay + by), 0.5 * (az + bz),
                     fprev);
  } else {
    stage_face<F, HYB>(g, q, 0.5 * (ax + bx), 0.5 * (ay + by), 0.5 * (az + bz), t0 | t1 | t2,
                       sa.tf, base + (m0 - 1) * st, fprev);
  }
  // viscous node flux of the march axis (Fv or Gv) on its offset line,
  // nodes m-1, m, m+1 of the face m+1/2 (visc_face); Ev of the xi-faces on
  // the zeta march's offset line
  const bool VIS = sa.vf != nullptr;
  const int vsh = c.visc_centered ? 0 : -1;
  const F* __restrict__ ve = static_cast<const F*>(sa.vf);
  const F* __restrict__ va = VIS ? ve + (long)(4 * AX) * S : nullptr;
  const int vbase = AX == 1 ? idx(c, i + vsh, 0, t + vsh) : idx(c, i + vsh, t + vsh, 0);
  const int ebase = idx(c, i, t + vsh, vsh);  // XSH: Ev of node m at ebase + m*st
  // The window (nodes m-1, m, m+1) is held in registers, or, in the FP64
  // zeta march, whose registers the xi-faces take, in a per-thread LDS ring
  // (slot vr holds node m-1, the next two slots the others; no thread reads
  // another's entries, so no barrier).
  constexpr bool VRING = AX == 2 && XSH && sizeof(F) == sizeof(double) && !HYB;
  constexpr int RR = VRING ? march_rows<AX>() : 1;
  __shared__ F vring[VRING ? 3 : 1][4][RR][VRING ? MARCH_X : 1];
  const int ty = VRING ? threadIdx.y : 0, tl = VRING ? lane : 0;
  F vw[3][4];
  int vr = 0;
  if (VIS) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      F w[4];
#pragma unroll
      for (int s = 0; s < 4; s++) w[s] = va[e * S + vbase + (m0 - 2 + s) * st];
      fprev[e + 1] -= (double)visc_face(w[0], w[1], w[2], w[3]);
#pragma unroll
      for (int s = 0; s < 3; s++) {
        if constexpr (VRING) vring[s][e][ty][tl] = w[s + 1];
        else vw[s][e] = w[s + 1];
      }
    }
  }
  // eta march: the next node's state and the next face's metrics are
  // fetched one iteration ahead (nodes up to mhi + 4 lie in the ghosted
  // array): -1.6 % for the FP64 eta march; the hybrid's eta march, bound by
  // memory latency at its register count, ran 2 % slower with it
  constexpr bool PF = AX == 1 && !HYB;
  double qpf[5] = {0, 0, 0, 0, 0}, jpf = 0.0, xpf = 0.0, ypf = 0.0, zpf = 0.0;
  if constexpr (PF) {
    load_q(Q, S, base + (m0 + 3) * st, qpf);
    const int on = base + (m0 + 1) * st;
    jpf = J[on]; xpf = nX[on]; ypf = nY[on]; zpf = nZ[on];
  }
  for (int m = m0; m < m1; m++) {
    // shift the window by one node
#pragma unroll
    for (int s = 0; s < 5; s++)
#pragma unroll
      for (int n = 0; n < 5; n++) q[s][n] = q[s + 1][n];
    o = base + m * st;
    if constexpr (PF) {
#pragma unroll
      for (int n = 0; n < 5; n++) q[5][n] = qpf[n];
      load_q(Q, S, o + 4 * st, qpf);
    } else {
      load_q(Q, S, o + 3 * st, q[5]);
    }
    F vn[4], ev[4];  // viscous flux of node m+2 along the axis; Ev of node m (XSH)
    if (VIS) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        vn[e] = va[e * S + vbase + (m + 2) * st];
        if constexpr (XSH) ev[e] = ve[e * S + ebase + m * st];
      }
    }
    if constexpr (DC) {
#pragma unroll
      for (int n = 0; n < 5; n++) {
#pragma unroll
        for (int k = 0; k < 4; k++) e[k][n] = e[k + 1][n];
        e[4][n] = (F)(q[5][n] - q[4][n]);
      }
      nl = nr;
      nr = node_t<F>(g, q_to_side(g, q[3]));
    }
    const double jo = jb;  // J at node m, loaded as the previous face's right node
    // the zeta sweep's update operands are fetched before the face work so
    // their latency overlaps it; with FP64 faces and the xi-faces in this
    // march they are fetched where they are used (hoisted they would push
    // the kernel past 256 VGPRs)
    constexpr bool HOIST = AX == 2 && !(XI && sizeof(F) == sizeof(double)) && !HYB;
    double rin[5], qnv[5];
    if (HOIST) {
#pragma unroll
      for (int n = 0; n < 5; n++) {
        rin[n] = sa.rhs[n * S + o];
        qnv[n] = sa.stage > 1 ? sa.Qn[n * S + o] : 0.0;
      }
    }
    ax = bx; ay = by; az = bz;
    if constexpr (PF) {
      jb = jpf;
      bx = xpf * jb; by = ypf * jb; bz = zpf * jb;
      const int on = o + 2 * st;
      jpf = J[on]; xpf = nX[on]; ypf = nY[on]; zpf = nZ[on];
    } else {
      jb = J[o + st];
      bx = nX[o + st] * jb; by = nY[o + st] * jb; bz = nZ[o + st] * jb;
    }
    double f[5], fxr[5];
    if (XI) {
      // the xi-face between node (i, m) and (i+1, m): stencil i-2 .. i+3
      double qx[6][5], nxx, nxy, nxz;
      bool tx = false;
      const double* __restrict__ xm = met + XI_X * S;
      if constexpr (XSH) {
        // every lane takes part in the shuffles
        if constexpr (HYB) {
          // linear xi-face from the neighbouring lanes' values, one
          // component at a time (six values live, not the whole stencil)
#pragma unroll
          for (int n = 0; n < 5; n++) {
            const double v = q[2][n];
            linear_lr<F>(__shfl(v, lane - 2), __shfl(v, lane - 1), v, __shfl(v, lane + 1),
                         __shfl(v, lane + 2), __shfl(v, lane + 3), qx[0][n], qx[1][n]);
          }
        } else if constexpr (!DC) {
#pragma unroll
          for (int s = 0; s < 6; s++)
#pragma unroll
            for (int n = 0; n < 5; n++)
              qx[s][n] = s == 2 ? q[2][n] : __shfl(q[2][n], lane + s - 2);
        }
        const double jx = xm[o] * jo, jy = xm[S + o] * jo, jz = xm[2 * S + o] * jo;
        nxx = 0.5 * (jx + __shfl(jx, lane + 1));
        nxy = 0.5 * (jy + __shfl(jy, lane + 1));
        nxz = 0.5 * (jz + __shfl(jz, lane + 1));
        if (HYB) {
          const int fo = sa.flags_x[o];
          tx = (__shfl(fo, lane - 1) | fo | __shfl(fo, lane + 1)) != 0;
        }
      } else {
#pragma unroll
        for (int s = 0; s < 6; s++) {
          if (s == 2) {
#pragma unroll
            for (int n = 0; n < 5; n++) qx[2][n] = q[2][n];
          } else {
            load_q(Q, S, o + (s - 2), qx[s]);
          }
        }
        const double Ja = jo, Jb = J[o + 1];
        nxx = 0.5 * (xm[o] * Ja + xm[o + 1] * Jb);
        nxy = 0.5 * (xm[S + o] * Ja + xm[S + o + 1] * Jb);
        nxz = 0.5 * (xm[2 * S + o] * Ja + xm[2 * S + o + 1] * Jb);
        tx = HYB && (sa.flags_x[o - 1] | sa.flags_x[o] | sa.flags_x[o + 1]);
      }
      if constexpr (XSH && DC) {
        // the float difference to lane+1's node, the differences of lanes
        // -2..+2 and lane+1's Roe inputs, by shuffles
        F ex[5], ec[5][5];
#pragma unroll
        for (int n = 0; n < 5; n++) ex[n] = (F)(__shfl(q[2][n], lane + 1) - q[2][n]);
#pragma unroll
        for (int n = 0; n < 5; n++) {
          ec[0][n] = __shfl(ex[n], lane - 2);
          ec[1][n] = __shfl(ex[n], lane - 1);
          ec[2][n] = ex[n];
          ec[3][n] = __shfl(ex[n], lane + 1);
          ec[4][n] = __shfl(ex[n], lane + 2);
        }
        const NodeT<F> xr{__shfl(nl.r, lane + 1), __shfl(nl.u, lane + 1), __shfl(nl.v, lane + 1),
                          __shfl(nl.w, lane + 1), __shfl(nl.H, lane + 1)};
        F dx[6][5];
        dev_from_diffs(ec, dx);
        face_flux_dev<F>(gt, g, dx, q[2], nl, xr, nxx, nxy, nxz, fxr);
      } else if constexpr (XSH && HYB) {
        // qx[0], qx[1]: the reconstructed left and right states
        lf_flux(g, q_to_side(g, qx[1]), q_to_side(g, qx[0]), nxx, nxy, nxz, fxr);
        if (tx) {
          const long sl = min(max(sa.tf_x.slot[o], 0), (int)(sa.tf_x.stride - 1));
#pragma unroll
          for (int n = 0; n < 5; n++) fxr[n] = sa.tf_x.flux[n * sa.tf_x.stride + sl];
        }
      } else {
        stage_face<F, HYB>(g, qx, nxx, nxy, nxz, tx, sa.tf_x, o, fxr);
      }
      if constexpr (XSH) {
        if (VIS) {
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const F v = ev[e];
            fxr[e + 1] -= (double)visc_face(__shfl(v, lane - 1), v, __shfl(v, lane + 1),
                                            __shfl(v, lane + 2));
          }
        }
      }
    }
    if (HYB) {
      t0 = t1;
      t1 = t2;
      t2 = fl[o + st];
    }
    if constexpr (DC) {
      F d[6][5];
      dev_from_diffs(e, d);
      face_flux_dev<F>(gt, g, d, q[2], nl, nr, 0.5 * (ax + bx), 0.5 * (ay + by), 0.5 * (az + bz),
                       f);
    } else {
      stage_face<F, HYB>(g, q, 0.5 * (ax + bx), 0.5 * (ay + by), 0.5 * (az + bz), t0 | t1 | t2,
                         sa.tf, o, f);
    }
    if (VIS) {
      const int s0 = vr, s1 = vr == 2 ? 0 : vr + 1, s2 = vr == 0 ? 2 : vr - 1;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        if constexpr (VRING) {
          f[e + 1] -= (double)visc_face(vring[s0][e][ty][tl], vring[s1][e][ty][tl],
                                        vring[s2][e][ty][tl], vn[e]);
          vring[s0][e][ty][tl] = vn[e];  // node m+2 takes node m-1's slot
        } else {
          f[e + 1] -= (double)visc_face(vw[0][e], vw[1][e], vw[2][e], vn[e]);
          vw[0][e] = vw[1][e];
          vw[1][e] = vw[2][e];
          vw[2][e] = vn[e];
        }
      }
      vr = s1;
    }
    double dfx[5];  // xi-face difference (XI)
    if (XI) {
#pragma unroll
      for (int n = 0; n < 5; n++) dfx[n] = fxr[n] - __shfl_up(fxr[n], 1);
    }
    if (AX == 1 && XI) {
#pragma unroll
      for (int n = 0; n < 5; n++)
        if (own) sa.rhs[n * S + o] = dfx[n] + (f[n] - fprev[n]);
    } else if (AX == 1 && FUSE_XI) {  // xi-faces in the zeta march
#pragma unroll
      for (int n = 0; n < 5; n++) sa.rhs[n * S + o] = f[n] - fprev[n];
    } else if (AX == 1) {
#pragma unroll
      for (int n = 0; n < 5; n++) {
        const double* fxn = sa.fx + n * S;
        sa.rhs[n * S + o] = (fxn[o] - fxn[o - 1]) + (f[n] - fprev[n]);
      }
    } else {
      if (!HOIST) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
          rin[n] = sa.rhs[n * S + o];
          qnv[n] = sa.stage > 1 ? sa.Qn[n * S + o] : 0.0;
        }
      }
      const double inv_j = fast_rcp(jo);
      double qn_out[5];
#pragma unroll
      for (int n = 0; n < 5; n++) {
        double df = rin[n] + (f[n] - fprev[n]);
        if (XI) df += dfx[n];
        df *= inv_j;
        const double qin = q[2][n];
        double out;
        if (sa.stage == 1) out = qin - dt * df;
        else if (sa.stage == 2) out = 0.75 * qnv[n] + 0.25 * (qin - dt * df);
        else out = (qnv[n] + 2.0 * (qin - dt * df)) * (1.0 / 3.0);
        if (own) sa.Qout[n * S + o] = out;
        qn_out[n] = out;
      }
      // primitives of the new state (src/primitive_bc_kernel.cu:63-76)
      if (own) {
        double rr, uu, vv, ww, pp;
        q_to_prim(g, qn_out, rr, uu, vv, ww, pp);
        if (sa.pmask & 1) sa.prim[o] = rr;
        if (sa.pmask & 2) sa.prim[S + o] = uu;
        if (sa.pmask & 4) sa.prim[2 * S + o] = vv;
        if (sa.pmask & 8) sa.prim[3 * S + o] = ww;
        if (sa.pmask & 16) sa.prim[4 * S + o] = pp;
        if (sa.pmask & 32) sa.prim[5 * S + o] = c.gM2 * pp * fast_rcp(rr);
      }
    }
#pragma unroll
    for (int n = 0; n < 5; n++) fprev[n] = f[n];
  }
}

// The troubled faces of axis AX in the flux kernels' face range fb: face m
// is troubled when its flag or a neighbour's along AX is set (the same
// one-cell dilation the marches apply).  Each is appended to list (node
// offset) and its slot (slot_base + position) recorded in slot[face].  A
// thread takes LIST_PLANES faces along k and a workgroup (64 x 4 threads)
// reserves the positions of all its faces with one atomic, so the counter
// sees one request per 4096 faces (one per wavefront and plane measured
// 70 ms per step at 512^3 with 60 % of the faces troubled: same-address
// atomics serialise).  Positions past cap are not written (the host then
// runs full WENO); once the count has passed cap, workgroups stop adding.
constexpr int LIST_PLANES = 16;

template <int AX>
__global__ void __launch_bounds__(256) k_face_list(Consts c, Box fb,
                                                   const uint8_t* __restrict__ fl,
                                                   int* __restrict__ slot, int* __restrict__ list,
                                                   unsigned* __restrict__ count, unsigned cap,
                                                   int slot_base) {
  const int lane = threadIdx.x, w = threadIdx.y;  // 64 x 4: one wavefront per row
  const int i = fb.lo[0] + (int)blockIdx.x * 64 + lane;
  const int j = fb.lo[1] + (int)blockIdx.y * 4 + w;
  const int k0 = fb.lo[2] + (int)blockIdx.z * LIST_PLANES;
  const int np = min(LIST_PLANES, fb.lo[2] + fb.n[2] - k0);
  const bool line = i < fb.lo[0] + fb.n[0] && j < fb.lo[1] + fb.n[1];
  const int st = stride(c, AX), sz = c.PXY;
  const int a0 = idx(c, i, j, k0);
  unsigned mine = 0;       // bit p: face (i, j, k0 + p) is troubled
  unsigned long long bal[LIST_PLANES];
  unsigned nwave = 0;
#pragma unroll
  for (int p = 0; p < LIST_PLANES; p++) {
    bool tr = false;
    if (line && p < np) {
      const int a = a0 + p * sz;
      tr = (fl[a - st] | fl[a] | fl[a + st]) != 0;
    }
    bal[p] = __ballot(tr);
    mine |= (unsigned)tr << p;
    nwave += (unsigned)__popcll(bal[p]);
  }
  __shared__ unsigned wn[4], wbase;
  if (lane == 0) wn[w] = nwave;
  __syncthreads();
  if (lane == 0 && w == 0) {
    const unsigned tot = wn[0] + wn[1] + wn[2] + wn[3];
    // past cap already (the count only grows): list nothing, positions from cap
    unsigned b = cap;
    if (tot && __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= cap)
      b = atomicAdd(count, tot);
    wbase = b;
  }
  __syncthreads();
  unsigned pos = wbase;
  for (int q = 0; q < w; q++) pos += wn[q];
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int p = 0; p < LIST_PLANES; p++) {
    if ((mine >> p) & 1u) {
      const unsigned at = pos + (unsigned)__popcll(bal[p] & below);
      if (at < cap) {
        const int a = a0 + p * sz;
        list[at] = a;
        slot[a] = slot_base + (int)at;
      }
    }
    pos += (unsigned)__popcll(bal[p]);
  }
}

// Characteristic-WENO flux (the stage's arithmetic form) of each listed face
// of axis AX, from the stage input Q, into its slot.
template <int AX, class F>
__global__ void __launch_bounds__(256) k_weno_list(Consts c, const int* __restrict__ list,
                                                   unsigned n, const double* __restrict__ Q,
                                                   const double* __restrict__ met,
                                                   double* __restrict__ flux, long fstride,
                                                   int slot_base) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long S = c.S;
  const int st = stride(c, AX);
  const int a = list[t], bn = a + st;
  double q[6][5], f[5];
#pragma unroll
  for (int s = 0; s < 6; s++) load_q(Q, S, a + (s - 2) * st, q[s]);
  const double Ja = met[JAC * S + a], Jb = met[JAC * S + bn];
  const double* nX = met + (3 * AX + 0) * S;
  const double* nY = met + (3 * AX + 1) * S;
  const double* nZ = met + (3 * AX + 2) * S;
  const double ax = nX[a] * Ja, ay = nY[a] * Ja, az = nZ[a] * Ja;
  const double bx = nX[bn] * Jb, by = nY[bn] * Jb, bz = nZ[bn] * Jb;
  const GasT<double> g{c.gamma, c.gam1, c.gam1_inv};
  face_flux_form<F>(g, q, 0.5 * (ax + bx), 0.5 * (ay + by), 0.5 * (az + bz), f);
#pragma unroll
  for (int m = 0; m < 5; m++) flux[m * fstride + slot_base + t] = f[m];
}

template <int AX>
void launch_face_list_axis(cbns_block* b, const Box& fb) {
  const dim3 grid((fb.n[0] + 63) / 64, (fb.n[1] + 3) / 4, (fb.n[2] + LIST_PLANES - 1) / LIST_PLANES);
  k_face_list<AX><<<grid, dim3(64, 4), 0, b->stream>>>(
      b->c, fb, b->flags + (long)AX * b->c.S, b->tslot + (long)AX * b->c.S, b->tlist + b->toff[AX],
      b->tcount + AX, (unsigned)b->tcap[AX], (int)b->toff[AX]);
}

template <class F>
int launch_weno_lists(cbns_block* b, const double* qin) {
  ProfScope ps(b, CBNS_K_WENO_LIST);
  for (int a = 0; a < 3; a++) {
    const unsigned n = b->h_tcount[a];
    if (n == 0) continue;
    const unsigned nb = (n + 255) / 256;
    const int sb = (int)b->toff[a];
    const int* lst = b->tlist + b->toff[a];
    if (a == 0) k_weno_list<0, F><<<nb, 256, 0, b->stream>>>(b->c, lst, n, qin, b->met, b->tflux, b->ttotal, sb);
    else if (a == 1) k_weno_list<1, F><<<nb, 256, 0, b->stream>>>(b->c, lst, n, qin, b->met, b->tflux, b->ttotal, sb);
    else k_weno_list<2, F><<<nb, 256, 0, b->stream>>>(b->c, lst, n, qin, b->met, b->tflux, b->ttotal, sb);
    CBNS_HIP(b, hipGetLastError());
  }
  return CBNS_OK;
}

}  // namespace

int launch_face_lists(cbns_block* b) {
  CBNS_HIP(b, hipMemsetAsync(b->tcount, 0, 3 * sizeof(unsigned), b->stream));
  for (int a = 0; a < 3; a++) {
    const Box fb = face_box(b->c, a, 0);
    if (fb.count() == 0) continue;
    if (a == 0) launch_face_list_axis<0>(b, fb);
    else if (a == 1) launch_face_list_axis<1>(b, fb);
    else launch_face_list_axis<2>(b, fb);
    CBNS_HIP(b, hipGetLastError());
  }
  CBNS_HIP(b, hipMemcpyAsync(b->h_tcount, b->tcount, 3 * sizeof(unsigned), hipMemcpyDeviceToHost,
                             b->stream));
  CBNS_HIP(b, hipStreamSynchronize(b->stream));
  b->tfull = false;
  for (int a = 0; a < 3; a++)
    if ((long)b->h_tcount[a] > b->tcap[a]) b->tfull = true;
  return CBNS_OK;
}

template <class F, bool HYB>
int launch_marches_eta(cbns_block* b, StageArgs& sa) {
  const Consts& c = b->c;
  const int ni = upd_hi(c, 0) - upd_lo(c, 0) + 1;
  const int nt = upd_hi(c, 2) - upd_lo(c, 2) + 1;
  const int nm = upd_hi(c, 1) - upd_lo(c, 1) + 1;
  const int per_wave = march_width<1, F>();
  const int nwx = (ni + per_wave - 1) / per_wave;
  dim3 grid((nwx + MARCH_Y - 1) / MARCH_Y, nt, (nm + CHUNK - 1) / CHUNK);
  ProfScope ps(b, CBNS_K_MARCH_ETA);
  k_march<1, F, HYB><<<grid, dim3(MARCH_X, MARCH_Y), 0, b->stream>>>(c, sa, b->met);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

template <class F, bool HYB>
int launch_marches_zeta(cbns_block* b, StageArgs& sa) {
  const Consts& c = b->c;
  const int ni = upd_hi(c, 0) - upd_lo(c, 0) + 1;
  const int nt = upd_hi(c, 1) - upd_lo(c, 1) + 1;
  const int nm = upd_hi(c, 2) - upd_lo(c, 2) + 1;
  const int per_wave = march_width<2, F>();
  constexpr int R = march_rows<2>();
  dim3 grid((ni + per_wave - 1) / per_wave, (nt + R - 1) / R, (nm + CHUNK - 1) / CHUNK);
  ProfScope ps(b, CBNS_K_MARCH_ZETA);
  k_march<2, F, HYB><<<grid, dim3(MARCH_X, R), 0, b->stream>>>(c, sa, b->met);
  CBNS_HIP(b, hipGetLastError());
  return CBNS_OK;
}

template <class F>
int viscous_part(cbns_block* b, double* lam, int parts) {
  return sizeof(F) == sizeof(double) ? launch_viscous(b, b->vtmp, b->rhs, 1, lam, parts)
                                     : launch_viscous_f32(b, b->vtmp, b->rhs, lam, parts);
}

// The marches take the viscous divergence in face form (visc_face) when the
// zeta march holds the xi-faces; otherwise k_visc_div subtracts it from the
// right-hand side between the marches.
constexpr int TFULL_EVERY = 8;  // steps between indicator passes while a list overflows

template <class F>
constexpr bool visc_in_marches() { return FUSE_XI && xi_axis<F>() == 2; }

// One fused stage: [indicator and face lists (hybrid, when due)], viscous
// flux pass, eta march, zeta march.  The flux pass reads only the stage
// input's primitives, so it may run before the eta march; the marches
// difference its node fluxes along with the inviscid face fluxes.
template <class F>
int fused_stage(cbns_block* b, double dt, int stage, const double* qin, double* qout,
                bool hyb_form) {
  const Consts& c = b->c;
  double* fx = b->flux;  // first 5 face arrays
  int rc = CBNS_OK;
  if (hyb_form) {
    // once per step on the stage-1 input (or every ind_every steps, or at
    // every stage).  While a list has overflowed, every face takes WENO and
    // the flags go unused: the pass then runs every TFULL_EVERY steps, to
    // see whether the lists fit again.
    const int every = b->tfull ? std::max(b->ind_every, TFULL_EVERY) : b->ind_every;
    const bool due = !b->flags_valid || (b->ind_per_stage && !b->tfull) ||
                     (stage == 1 && b->ind_age >= every);
    if (due) {
      if ((rc = launch_indicator(b, qin)) || (rc = launch_face_lists(b))) return rc;
      b->ind_age = 0;
    }
    if (stage == 1) b->ind_age++;
  }
  // first stage of a folded step: the flux pass also forms dt
  double* lam = c.viscous && b->dt_fold && stage == 1 ? b->lam_part : nullptr;
  if (c.viscous && (rc = viscous_part<F>(b, lam, 1))) return rc;
  // more troubled faces than the lists hold: WENO at every face
  const bool hyb = hyb_form && !b->tfull;
  if (hyb && (rc = launch_weno_lists<F>(b, qin))) return rc;
  const long S = c.S;
  const uint8_t* fl = hyb ? b->flags : nullptr;
  if (!FUSE_XI) {
    Box box;
    for (int a = 0; a < 3; a++) {
      box.lo[a] = upd_lo(c, a);
      box.n[a] = upd_hi(c, a) - upd_lo(c, a) + 1;
    }
    box.lo[0] = periodic(c, 0) ? -1 : 0;
    box.n[0] = c.IM - 1 + (periodic(c, 0) ? 2 : 0);  // faces -1..N-1 or 0..N-2
    ProfScope ps(b, CBNS_K_XFACES);
    const TroubledFaces tx{b->tslot, b->tflux, b->ttotal};
    if (hyb)
      k_xfaces<F, true><<<box_grid(box), box_block(), 0, b->stream>>>(c, box, qin, b->met, fl, tx, fx);
    else
      k_xfaces<F, false><<<box_grid(box), box_block(), 0, b->stream>>>(c, box, qin, b->met, fl, tx, fx);
    CBNS_HIP(b, hipGetLastError());
  }
  const int pmask = stage_prim_mask(b, stage);
  StageArgs sa{dt, b->dt_fold ? b->red + 2 : nullptr, stage, b->prim, b->Q, qin, qout, fx,
               b->rhs, c.viscous && visc_in_marches<F>() ? b->vtmp : nullptr,
               hyb ? fl + S : nullptr, hyb ? fl : nullptr, pmask,
               TroubledFaces{hyb ? b->tslot + S : nullptr, b->tflux, b->ttotal},
               TroubledFaces{hyb ? b->tslot : nullptr, b->tflux, b->ttotal}};
  if ((rc = hyb ? launch_marches_eta<F, true>(b, sa) : launch_marches_eta<F, false>(b, sa)))
    return rc;
  if (c.viscous) {
    if (!visc_in_marches<F>() && (rc = viscous_part<F>(b, nullptr, 2))) return rc;
    if (lam) {
      ProfScope ps(b, CBNS_K_DT);
      if ((rc = launch_dt_from_partials(b, lam, (int)visc_flux_blocks(c), b->dt_cap, b->red + 2)))
        return rc;
    }
  }
  if (hyb) {
    sa.flags = fl + 2 * S;
    sa.tf.slot = b->tslot + 2 * S;
  }
  return hyb ? launch_marches_zeta<F, true>(b, sa) : launch_marches_zeta<F, false>(b, sa);
}

bool fused_stage_uses_face_flux() { return !FUSE_XI; }

int launch_fused_stage(cbns_block* b, double dt, int stage, const double* qin, double* qout) {
  const bool hyb = (b->form & CBNS_FORM_HYBRID) != 0;
  return (b->form & CBNS_FORM_FP32) ? fused_stage<float>(b, dt, stage, qin, qout, hyb)
                                    : fused_stage<double>(b, dt, stage, qin, qout, hyb);
}

}  // namespace cbns
