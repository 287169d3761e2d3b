"""ctypes wrapper of the CPU oracle (oracle/libcbns_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, as the checker.  The product never imports it.

Arrays follow the oracle's ghosted SoA layout: a field is a C-contiguous
float64 array of shape (KM+2G, JM+2G, IM+2G) (i fastest), node (i,j,k) at
[k+G, j+G, i+G]; multi-component arrays add a leading component axis.
"""
import ctypes as C
import os
import subprocess

import numpy as np

G = 4
BC_REF, BC_PERIODIC = 0, 1
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcbns_oracle.so")


class Params(C.Structure):
    """Field-for-field copy of orc_params (oracle/cbns_oracle.h)."""

    _fields_ = [
        ("IM", C.c_int), ("JM", C.c_int), ("KM", C.c_int),
        ("i_begin", C.c_int), ("i_end", C.c_int),
        ("viscous", C.c_int), ("visc_centered", C.c_int),
        ("bc", C.c_int * 3),
        ("CFL", C.c_double), ("gamma", C.c_double), ("Mach", C.c_double),
        ("Re", C.c_double), ("Pr", C.c_double), ("Pr_t", C.c_double),
        ("T_inf", C.c_double), ("angle_attack", C.c_double),
    ]


class Grid(C.Structure):
    _fields_ = [
        ("IM", C.c_int), ("JM", C.c_int), ("KM", C.c_int),
        ("PX", C.c_long), ("PY", C.c_long), ("PZ", C.c_long), ("S", C.c_long),
        ("i_begin", C.c_int), ("i_end", C.c_int),
        ("viscous", C.c_int), ("visc_centered", C.c_int), ("bc", C.c_int * 3),
        ("CFL", C.c_double), ("gamma", C.c_double), ("Mach", C.c_double),
        ("Re", C.c_double), ("Pr", C.c_double), ("Pr_t", C.c_double),
        ("T_inf", C.c_double), ("angle_attack", C.c_double),
        ("Mach2", C.c_double), ("gM2", C.c_double), ("Re_inv", C.c_double),
        ("Pr_inv", C.c_double), ("Pr_t_inv", C.c_double), ("gam1", C.c_double),
        ("gam1_inv", C.c_double), ("gPr", C.c_double), ("p_inf", C.c_double),
        ("C_T_inf", C.c_double), ("C_dt_v", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        D = C.POINTER(C.c_double)
        GP = C.POINTER(Grid)
        L.orc_grid_init.argtypes = [GP, C.POINTER(Params)]
        L.orc_weno5.argtypes = [C.c_double] * 5
        L.orc_weno5.restype = C.c_double
        L.orc_eigenvectors_roe.argtypes = [GP, D, D, C.c_double, C.c_double, C.c_double, D, D]
        L.orc_rec_weno5.argtypes = [GP, D, D, D, D]
        L.orc_lf_flux.argtypes = [GP, D, D, C.c_double, C.c_double, C.c_double, D]
        L.orc_face_flux.argtypes = [GP, D, D, D, C.c_double, C.c_double, C.c_double, D]
        L.orc_gradient_nodes.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D, D]
        L.orc_calc_metrics.argtypes = [GP, D, D, D, D, D]
        L.orc_bc_primitive.argtypes = [GP, D, D, D]
        L.orc_numerical_flux.argtypes = [GP, D, D, D, D]
        L.orc_viscous.argtypes = [GP, D, D, D, D]
        L.orc_update_rk3.argtypes = [GP, C.c_double, C.c_int, D, D, D, D, D]
        L.orc_calc_dt.argtypes = [GP, D, D, D]
        L.orc_calc_dt.restype = C.c_double
        L.orc_rk3_stage.argtypes = [GP, C.c_double, C.c_int, D, D, D, D, D]
        L.orc_rk3_step.argtypes = [GP, C.c_double, D, D, D, D, D, D, D]
        L.orc_conservative.argtypes = [GP, D, D]
        L.orc_threads.restype = C.c_int
        F = C.POINTER(C.c_float)
        L.orc_bf16_round.argtypes = [C.c_float]
        L.orc_bf16_round.restype = C.c_float
        L.orc_face_features.argtypes = [GP, D, D, D, C.c_double, C.c_double, C.c_double, F]
        L.orc_weno5_weights.argtypes = [C.c_double] * 5 + [D]
        L.orc_face_label.argtypes = [GP, D, D, D, C.c_double, C.c_double, C.c_double]
        L.orc_face_label.restype = C.c_double
        L.orc_mlp_forward.argtypes = [F, F, C.c_long, C.c_int, F]
        L.orc_mlp_loss_grad.argtypes = [F, F, F, C.c_long, C.c_double, C.c_int, D]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_face_flux_linear.argtypes = [GP, D, C.c_double, C.c_double, C.c_double, D]
        U8 = C.c_void_p
        L.orc_numerical_flux_hybrid.argtypes = [GP, D, D, D, U8, D]
        L.orc_rk3_stage_hybrid.argtypes = [GP, C.c_double, C.c_int, D, D, D, D, D, U8]
        L.orc_adam.argtypes = [F, D, D, D, C.c_long] + [C.c_double] * 4 + [C.c_int]
        _lib = L
    return _lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous array required"
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(IM, JM, KM, *, viscous=0, visc_centered=0, bc=(BC_PERIODIC,) * 3,
                CFL=0.5, gamma=1.4, Mach=0.5, Re=1600.0, Pr=0.72, Pr_t=0.9,
                T_inf=273.15, angle_attack=0.0, i_begin=0, i_end=0):
    p = Params()
    p.IM, p.JM, p.KM = IM, JM, KM
    p.i_begin, p.i_end = i_begin, i_end
    p.viscous, p.visc_centered = int(viscous), int(visc_centered)
    p.bc[0], p.bc[1], p.bc[2] = bc
    p.CFL, p.gamma, p.Mach, p.Re = CFL, gamma, Mach, Re
    p.Pr, p.Pr_t, p.T_inf, p.angle_attack = Pr, Pr_t, T_inf, angle_attack
    return p


class Oracle:
    """One block on the CPU oracle: owns the ghosted arrays."""

    def __init__(self, params):
        self.params = params
        self.g = Grid()
        rc = lib().orc_grid_init(C.byref(self.g), C.byref(params))
        if rc != 0:
            raise ValueError("orc_grid_init failed (grid too small?)")
        g = self.g
        self.shape = (g.PZ, g.PY, g.PX)
        self.S = g.S
        self.met = np.zeros((10,) + self.shape)
        self.Q = np.zeros((5,) + self.shape)
        self.Qp = np.zeros((5,) + self.shape)
        self.prim = np.zeros((6,) + self.shape)
        self.mu = np.zeros(self.shape)

    # views of the node range (no ghosts)
    def inner(self, a):
        return a[..., G:G + self.g.KM, G:G + self.g.JM, G:G + self.g.IM]

    def set_mesh(self, x, y, z, period=None):
        """x, y, z: node arrays of shape (KM, JM, IM)."""
        per = np.zeros(9) if period is None else np.ascontiguousarray(period, dtype=np.float64).reshape(9)
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        return lib().orc_calc_metrics(C.byref(self.g), ptr(x), ptr(y), ptr(z), ptr(per), ptr(self.met))

    def bc_primitive(self, Q=None):
        Q = self.Q if Q is None else Q
        lib().orc_bc_primitive(C.byref(self.g), ptr(Q), ptr(self.prim), ptr(self.met))

    def conservative(self):
        lib().orc_conservative(C.byref(self.g), ptr(self.prim), ptr(self.Q))

    def numerical_flux(self, Q=None):
        Q = self.Q if Q is None else Q
        flux = np.zeros((15,) + self.shape)
        lib().orc_numerical_flux(C.byref(self.g), ptr(Q), ptr(self.prim), ptr(self.met), ptr(flux))
        return flux.reshape((3, 5) + self.shape)

    def viscous(self):
        dfv = np.zeros((4,) + self.shape)
        lib().orc_viscous(C.byref(self.g), ptr(self.prim), ptr(self.met), ptr(self.mu), ptr(dfv))
        return dfv

    def calc_dt(self):
        return lib().orc_calc_dt(C.byref(self.g), ptr(self.prim), ptr(self.mu), ptr(self.met))

    def rk3_stage(self, dt, stage):
        rc = lib().orc_rk3_stage(C.byref(self.g), dt, stage, ptr(self.Q), ptr(self.Qp),
                                 ptr(self.prim), ptr(self.met), ptr(self.mu))
        if rc:
            raise RuntimeError("orc_rk3_stage failed")

    def rk3_stage_hybrid(self, dt, stage, flags):
        """stage with the learned hybrid's troubled-face bytes (3, *ghosted shape)"""
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        assert flags.size == 3 * self.g.S
        rc = lib().orc_rk3_stage_hybrid(C.byref(self.g), dt, stage, ptr(self.Q), ptr(self.Qp),
                                        ptr(self.prim), ptr(self.met), ptr(self.mu),
                                        flags.ctypes.data)
        if rc:
            raise RuntimeError("orc_rk3_stage_hybrid failed")

    def face_features_all(self, axis, Q=None):
        """features (n, 32) and labels (n,) of every face the flux kernels
        compute on `axis`, in the GPU library's sample order (i fastest)."""
        Q = self.Q if Q is None else Q
        g = self.g
        lo, hi = [], []
        for a, n in enumerate((g.IM, g.JM, g.KM)):
            per = g.bc[a] == BC_PERIODIC
            if a == axis:
                lo.append(-1 if per else 0)
                hi.append(n - 1 if per else n - 2)
            else:
                lo.append(0 if per else 1)
                hi.append(n - 1 if per else n - 2)
        feats, labels = [], []
        L = lib()
        f = np.zeros(32, dtype=np.float32)
        q = np.zeros((6, 5))
        pl = np.zeros(5)
        pr = np.zeros(5)
        for k in range(lo[2], hi[2] + 1):
            for j in range(lo[1], hi[1] + 1):
                for i in range(lo[0], hi[0] + 1):
                    node = [i, j, k]
                    for s in range(6):
                        nn = list(node)
                        nn[axis] += s - 2
                        q[s] = Q[:, nn[2] + G, nn[1] + G, nn[0] + G]
                    nb = list(node)
                    nb[axis] += 1
                    a_ = (slice(None), k + G, j + G, i + G)
                    b_ = (slice(None), nb[2] + G, nb[1] + G, nb[0] + G)
                    for side, (qq, out) in enumerate(((q[2], pl), (q[3], pr))):
                        r = qq[0]
                        u, v, w = qq[1] / r, qq[2] / r, qq[3] / r
                        out[:] = (r, u, v, w, g.gam1 * (qq[4] - 0.5 * r * (u * u + v * v + w * w)))
                    m = self.met
                    n = [0.5 * (m[3 * axis + d][a_[1:]] * m[9][a_[1:]] +
                                m[3 * axis + d][b_[1:]] * m[9][b_[1:]]) for d in range(3)]
                    qf = np.ascontiguousarray(q.reshape(30))
                    L.orc_face_features(C.byref(g), ptr(qf), ptr(pl), ptr(pr), n[0], n[1], n[2],
                                        fptr(f))
                    feats.append(f.copy())
                    labels.append(L.orc_face_label(C.byref(g), ptr(qf), ptr(pl), ptr(pr),
                                                   n[0], n[1], n[2]))
        return np.array(feats), np.array(labels)

    def rk3_step(self, t_end, t_cur):
        tc = C.c_double(t_cur)
        dt = C.c_double(0.0)
        rc = lib().orc_rk3_step(C.byref(self.g), t_end, C.byref(tc), C.byref(dt), ptr(self.Q),
                                ptr(self.Qp), ptr(self.prim), ptr(self.met), ptr(self.mu))
        if rc:
            raise RuntimeError("orc_rk3_step failed (rc=%d)" % rc)
        return tc.value, dt.value


def weno5(a, b, c, d, e):
    return lib().orc_weno5(a, b, c, d, e)


def eigenvectors_roe(grid, pl, pr, n):
    R = np.zeros(25)
    L = np.zeros(25)
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    lib().orc_eigenvectors_roe(C.byref(grid), ptr(pl), ptr(pr), n[0], n[1], n[2], ptr(R), ptr(L))
    return R.reshape(5, 5), L.reshape(5, 5)


def rec_weno5(grid, f, R, L):
    out = np.zeros(5)
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(25)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(25)
    L = np.ascontiguousarray(L, dtype=np.float64).reshape(25)
    lib().orc_rec_weno5(C.byref(grid), ptr(f), ptr(R), ptr(L), ptr(out))
    return out


def lf_flux(grid, pr, pl, n):
    out = np.zeros(5)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    lib().orc_lf_flux(C.byref(grid), ptr(pr), ptr(pl), n[0], n[1], n[2], ptr(out))
    return out


def face_flux(grid, q, pl, pr, n):
    out = np.zeros(5)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(30)
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    lib().orc_face_flux(C.byref(grid), ptr(q), ptr(pl), ptr(pr), n[0], n[1], n[2], ptr(out))
    return out


def gradient_nodes(f):
    """f: (KM, JM, IM) node array -> (f_xi, f_eta, f_zeta)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    KM, JM, IM = f.shape
    outs = [np.zeros_like(f) for _ in range(3)]
    lib().orc_gradient_nodes(IM, JM, KM, ptr(f), *(ptr(o) for o in outs))
    return outs


def grid_for(IM=8, JM=8, KM=8, **kw):
    g = Grid()
    lib().orc_grid_init(C.byref(g), C.byref(make_params(IM, JM, KM, **kw)))
    return g


# ---- learned troubled-face indicator (oracle/orc_learned.c) -------------
MLP_NFEAT, MLP_IN, MLP_HID = 30, 32, 32
MLP_W1, MLP_B1, MLP_W2, MLP_B2, MLP_W3, MLP_B3, MLP_NPARAM = 0, 1024, 1056, 2080, 2112, 2144, 2145


def fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "float32 C-contiguous array required"
    return a.ctypes.data_as(C.POINTER(C.c_float))


def bf16_round(x):
    return np.array([lib().orc_bf16_round(float(v)) for v in np.ravel(x)], dtype=np.float32)


def _face_args(q, pl, pr):
    return (np.ascontiguousarray(q, dtype=np.float64).reshape(30),
            np.ascontiguousarray(pl, dtype=np.float64), np.ascontiguousarray(pr, dtype=np.float64))


def face_features(grid, q, pl, pr, n):
    q, pl, pr = _face_args(q, pl, pr)
    out = np.zeros(32, dtype=np.float32)
    lib().orc_face_features(C.byref(grid), ptr(q), ptr(pl), ptr(pr), n[0], n[1], n[2], fptr(out))
    return out


def weno5_weights(a, b, c, d, e):
    om = np.zeros(3)
    lib().orc_weno5_weights(a, b, c, d, e, ptr(om))
    return om


def face_label(grid, q, pl, pr, n):
    q, pl, pr = _face_args(q, pl, pr)
    return lib().orc_face_label(C.byref(grid), ptr(q), ptr(pl), ptr(pr), n[0], n[1], n[2])


def mlp_forward(params, X, bf16=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
    out = np.zeros(len(X), dtype=np.float32)
    lib().orc_mlp_forward(fptr(params), fptr(X), len(X), int(bf16), fptr(out))
    return out


def mlp_loss_grad(params, X, y, w_pos=1.0, bf16=True, grad=True):
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.zeros(MLP_NPARAM) if grad else None
    loss = lib().orc_mlp_loss_grad(fptr(params), fptr(X), fptr(y), len(X), w_pos, int(bf16),
                                   ptr(g))
    return (loss, g) if grad else loss


def adam(params, grad, m, v, lr, b1=0.9, b2=0.999, eps=1e-8, t=1):
    lib().orc_adam(fptr(params), ptr(grad), ptr(m), ptr(v), len(params), lr, b1, b2, eps, t)


def face_flux_linear(grid, q, n):
    out = np.zeros(5)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(30)
    lib().orc_face_flux_linear(C.byref(grid), ptr(q), n[0], n[1], n[2], ptr(out))
    return out
