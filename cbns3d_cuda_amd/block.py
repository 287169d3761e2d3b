"""Python mirror of the block3d_cuda launcher API over the C ABI (libcbns.so).

Names and argument meanings follow the reference's include/*.h launchers
(calc_metrics, set_bc_calc_primitive, calc_dt, calc_numerical_flux,
calc_viscous_flux_contribution, update_rk3, ...); a non-zero status from the
library raises CbnsError carrying cbns_last_error().  Host arrays at the edge
use the reference's layouts (node arrays IM*JM*KM; 2-ghost-layer arrays;
AoS Q).  The library is the HIP build in this package directory: there is no
CPU fallback, and importing without it raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcbns.so")

G = 4
NG_REF = 2
BC_REF, BC_PERIODIC = 0, 1
BUF_Q, BUF_QP = 0, 1
FORM_FP64, FORM_FP32, FORM_HYBRID = 0, 1, 2
MLP_NFEAT, MLP_NPARAM = 32, 2145
F_Q, F_QP, F_PRIM, F_MET, F_MU, F_FLUX, F_DFV = 0, 5, 10, 16, 26, 27, 42
ERR_NAMES = {1: "ARG", 2: "HIP", 3: "NAN", 4: "MESH", 5: "STATE", 6: "NO_DEVICE", 7: "SIZE"}
ERR_SIZE = 7
NOALLOC_MAX_NODES = 1 << 20


class CbnsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cbns error %s (%d): %s" % (ERR_NAMES.get(code, "?"), code, msg))
        self.code = code


class Params(C.Structure):
    """cbns_params of include/cbns.h."""

    _fields_ = [
        ("IM", C.c_int32), ("JM", C.c_int32), ("KM", C.c_int32),
        ("i_begin", C.c_int32), ("i_end", C.c_int32),
        ("viscous", C.c_int32), ("visc_centered", C.c_int32),
        ("bc", C.c_int32 * 3),
        ("CFL", C.c_double), ("gamma", C.c_double), ("Mach", C.c_double),
        ("Re", C.c_double), ("Pr", C.c_double), ("Pr_t", C.c_double),
        ("T_inf", C.c_double), ("angle_attack", C.c_double),
    ]


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


_lib = None

# every extern "C" symbol include/cbns.h declares
EXPORTS = [
    "cbns_create", "cbns_destroy", "cbns_last_error", "cbns_set_mesh", "cbns_get_metrics",
    "cbns_calc_gradient", "cbns_calc_gradient_noalloc", "cbns_initial_condition", "cbns_calc_conservative",
    "cbns_set_conservative", "cbns_get_conservative", "cbns_calc_primitive",
    "cbns_get_primitive", "cbns_set_bc_calc_primitive", "cbns_calc_dt",
    "cbns_calc_numerical_flux", "cbns_calc_viscous_terms",
    "cbns_calc_viscous_flux_contribution", "cbns_update_rk3", "cbns_rk3_stage",
    "cbns_rk3_step", "cbns_set_form", "cbns_get_field", "cbns_set_field",
    "cbns_synchronize", "cbns_field_size", "cbns_set_profiling", "cbns_kernel_time",
    "cbns_face_count", "cbns_face_samples", "cbns_set_indicator", "cbns_flag_count",
    "cbns_get_flags", "cbns_dev_alloc", "cbns_dev_free", "cbns_dev_upload", "cbns_dev_download",
    "cbns_gather_rows",
    "cbns_mlp_create", "cbns_mlp_destroy", "cbns_mlp_last_error", "cbns_mlp_set_params",
    "cbns_mlp_get_params", "cbns_mlp_forward", "cbns_mlp_loss_grad", "cbns_mlp_adam_step",
]
KERNELS = ["xfaces", "march_eta", "march_zeta", "visc_flux", "visc_div", "bc", "dt", "indicator",
           "weno_list"]


def lib():
    """Load libcbns.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libcbns.so not built: run __graft_entry__.build() or make -C cbns3d_cuda_amd")
    L = C.CDLL(LIB_PATH)
    D = C.POINTER(C.c_double)
    B = C.c_void_p
    I = C.c_int32
    FP = C.POINTER(C.c_float)
    sig = {
        "cbns_create": ([C.POINTER(Params), C.POINTER(C.c_void_p)], C.c_int),
        "cbns_destroy": ([B], None),
        "cbns_last_error": ([B], C.c_char_p),
        "cbns_set_mesh": ([B, D, D, D, D], C.c_int),
        "cbns_get_metrics": ([B] + [D] * 10, C.c_int),
        "cbns_calc_gradient": ([I, I, I, D, D, D, D], C.c_int),
        "cbns_calc_gradient_noalloc": ([I, I, I, D, D, D, D], C.c_int),
        "cbns_initial_condition": ([B], C.c_int),
        "cbns_calc_conservative": ([B], C.c_int),
        "cbns_set_conservative": ([B, D], C.c_int),
        "cbns_get_conservative": ([B, D], C.c_int),
        "cbns_calc_primitive": ([B], C.c_int),
        "cbns_get_primitive": ([B, D, D, D, D, D], C.c_int),
        "cbns_set_bc_calc_primitive": ([B, I], C.c_int),
        "cbns_calc_dt": ([B, D], C.c_int),
        "cbns_calc_numerical_flux": ([B, I], C.c_int),
        "cbns_calc_viscous_terms": ([B], C.c_int),
        "cbns_calc_viscous_flux_contribution": ([B], C.c_int),
        "cbns_update_rk3": ([B, C.c_double, I], C.c_int),
        "cbns_rk3_stage": ([B, C.c_double, I], C.c_int),
        "cbns_rk3_step": ([B, C.c_double, D, D], C.c_int),
        "cbns_set_form": ([B, I], C.c_int),
        "cbns_get_field": ([B, I, D], C.c_int),
        "cbns_set_field": ([B, I, D], C.c_int),
        "cbns_synchronize": ([B], C.c_int),
        "cbns_field_size": ([B], C.c_int64),
        "cbns_set_profiling": ([B, I], C.c_int),
        "cbns_kernel_time": ([B, I, I, D, C.POINTER(C.c_int64)], C.c_int),
        "cbns_face_count": ([B, I, C.POINTER(C.c_int64)], C.c_int),
        "cbns_face_samples": ([B, I, I, C.c_void_p, C.c_void_p], C.c_int),
        "cbns_set_indicator": ([B, FP, C.c_float, I], C.c_int),
        "cbns_flag_count": ([B, C.POINTER(C.c_int64)], C.c_int),
        "cbns_get_flags": ([B, C.c_void_p], C.c_int),
        "cbns_dev_alloc": ([C.c_int64, C.POINTER(C.c_void_p)], C.c_int),
        "cbns_dev_free": ([C.c_void_p], C.c_int),
        "cbns_dev_upload": ([C.c_void_p, C.c_void_p, C.c_int64], C.c_int),
        "cbns_dev_download": ([C.c_void_p, C.c_void_p, C.c_int64], C.c_int),
        "cbns_gather_rows": ([C.c_void_p, C.c_void_p, C.c_int64, I, C.c_void_p], C.c_int),
        "cbns_mlp_create": ([C.POINTER(C.c_void_p)], C.c_int),
        "cbns_mlp_destroy": ([B], None),
        "cbns_mlp_last_error": ([B], C.c_char_p),
        "cbns_mlp_set_params": ([B, FP], C.c_int),
        "cbns_mlp_get_params": ([B, FP], C.c_int),
        "cbns_mlp_forward": ([B, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
        "cbns_mlp_loss_grad": ([B, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, D, FP], C.c_int),
        "cbns_mlp_adam_step": ([B, C.c_float, C.c_float, C.c_float, C.c_float], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Block:
    """One curvilinear block on the GPU (one HIP stream, library-owned HBM)."""

    def __init__(self, params):
        self.params = params
        self.IM, self.JM, self.KM = params.IM, params.JM, params.KM
        h = C.c_void_p()
        rc = lib().cbns_create(C.byref(params), C.byref(h))
        if rc:
            raise CbnsError(rc, "cbns_create failed")
        self._h = h
        self.shape = (self.KM + 2 * G, self.JM + 2 * G, self.IM + 2 * G)
        self.ref_shape = (self.KM + 2 * NG_REF, self.JM + 2 * NG_REF, self.IM + 2 * NG_REF)

    def _chk(self, rc):
        if rc:
            msg = lib().cbns_last_error(self._h)
            raise CbnsError(rc, msg.decode() if msg else "")

    def free(self):
        if self._h:
            lib().cbns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- setup (src/calc_metrics.cu, src/initial_condition.cu, ...) ---------
    def calc_metrics(self, x, y, z, period=None):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        per = None if period is None else np.ascontiguousarray(period, dtype=np.float64).reshape(9)
        self._chk(lib().cbns_set_mesh(self._h, _ptr(x), _ptr(y), _ptr(z),
                                      None if per is None else _ptr(per)))

    def get_metrics(self):
        outs = [np.zeros((self.KM, self.JM, self.IM)) for _ in range(10)]
        self._chk(lib().cbns_get_metrics(self._h, *(_ptr(o) for o in outs)))
        return outs

    def initial_condition(self):
        self._chk(lib().cbns_initial_condition(self._h))

    def calc_conservative(self):
        self._chk(lib().cbns_calc_conservative(self._h))

    def set_conservative(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        assert Q.size == 5 * np.prod(self.ref_shape)
        self._chk(lib().cbns_set_conservative(self._h, _ptr(Q)))

    def get_conservative(self):
        Q = np.zeros(self.ref_shape + (5,))
        self._chk(lib().cbns_get_conservative(self._h, _ptr(Q)))
        return Q

    def calc_primitive(self):
        self._chk(lib().cbns_calc_primitive(self._h))

    def get_primitive(self):
        outs = [np.zeros(self.ref_shape) for _ in range(5)]
        self._chk(lib().cbns_get_primitive(self._h, *(_ptr(o) for o in outs)))
        return outs

    # -- the step (src/Block3d.cpp:237-293) -------------------------------
    def set_bc_calc_primitive(self, buf=BUF_Q):
        self._chk(lib().cbns_set_bc_calc_primitive(self._h, buf))

    def calc_dt(self):
        dt = C.c_double(0.0)
        self._chk(lib().cbns_calc_dt(self._h, C.byref(dt)))
        return dt.value

    def calc_numerical_flux(self, buf=BUF_Q):
        self._chk(lib().cbns_calc_numerical_flux(self._h, buf))

    def calc_viscous_terms(self):
        self._chk(lib().cbns_calc_viscous_terms(self._h))

    def calc_viscous_flux_contribution(self):
        self._chk(lib().cbns_calc_viscous_flux_contribution(self._h))

    def update_rk3(self, dt, stage):
        self._chk(lib().cbns_update_rk3(self._h, dt, stage))

    def rk3_stage(self, dt, stage):
        self._chk(lib().cbns_rk3_stage(self._h, dt, stage))

    def rk3_step(self, t_end, t_cur):
        tc = C.c_double(t_cur)
        dt = C.c_double(0.0)
        self._chk(lib().cbns_rk3_step(self._h, t_end, C.byref(tc), C.byref(dt)))
        return tc.value, dt.value

    def set_form(self, form):
        self._chk(lib().cbns_set_form(self._h, form))

    def synchronize(self):
        self._chk(lib().cbns_synchronize(self._h))

    # -- device timing ------------------------------------------------------
    def set_profiling(self, on):
        self._chk(lib().cbns_set_profiling(self._h, int(on)))

    def kernel_times(self, reset=True):
        """{kernel name: (total ms, launches)} since the last reset."""
        out = {}
        for k, name in enumerate(KERNELS):
            ms = C.c_double(0.0)
            n = C.c_int64(0)
            self._chk(lib().cbns_kernel_time(self._h, k, int(reset), C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    # -- raw ghosted SoA fields ---------------------------------------------
    def get_field(self, f, n=1):
        out = np.zeros((n,) + self.shape)
        for q in range(n):
            self._chk(lib().cbns_get_field(self._h, f + q, _ptr(out[q])))
        return out if n > 1 else out[0]

    def set_field(self, f, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        n = arr.size // int(np.prod(self.shape))
        arr = arr.reshape((n,) + self.shape)
        for q in range(n):
            self._chk(lib().cbns_set_field(self._h, f + q, _ptr(arr[q])))

    # -- learned troubled-face indicator -----------------------------------
    # d_* arguments are device addresses (ints) of this library's HIP
    # runtime: DeviceArray.ptr / .at(offset).  Pointers from another HIP
    # runtime in the same process (a framework's bundled one) must not be
    # passed: the two runtimes do not share allocations.
    def face_count(self, axis):
        n = C.c_int64(0)
        self._chk(lib().cbns_face_count(self._h, axis, C.byref(n)))
        return n.value

    def face_samples(self, axis, d_feat, d_label, buf=BUF_Q):
        """Features [n][32] and WENO-weight labels [n] of axis `axis`."""
        self._chk(lib().cbns_face_samples(self._h, axis, buf, d_feat, d_label))

    def set_indicator(self, params, threshold=0.0, per_stage=False, count=False, reuse=False,
                      full_lists=False):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == MLP_NPARAM
        opts = ((1 if per_stage else 0) | (2 if count else 0) | (4 if reuse else 0) |
                (8 if full_lists else 0))
        self._chk(lib().cbns_set_indicator(self._h, _fptr(params), float(threshold), opts))

    def get_flags(self):
        """troubled-face bytes, shape (3,) + ghosted shape"""
        out = np.zeros((3,) + self.shape, dtype=np.uint8)
        self._chk(lib().cbns_get_flags(self._h, out.ctypes.data))
        return out

    def flag_count(self):
        out = (C.c_int64 * 3)()
        self._chk(lib().cbns_flag_count(self._h, out))
        return list(out)


class DeviceArray:
    """A device buffer of libcbns's own HIP runtime (shape + numpy dtype).
    Its address (`ptr`) is what the d_* arguments of the C ABI take."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        rc = lib().cbns_dev_alloc(max(self.nbytes, 1), C.byref(p))
        if rc:
            raise CbnsError(rc, "cbns_dev_alloc(%d) failed" % self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        rc = lib().cbns_dev_upload(self.ptr, a.ctypes.data, a.nbytes)
        if rc:
            raise CbnsError(rc, "cbns_dev_upload failed")

    def download(self, count=None):
        n = int(np.prod(self.shape)) if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        rc = lib().cbns_dev_download(out.ctypes.data, self.ptr, out.nbytes)
        if rc:
            raise CbnsError(rc, "cbns_dev_download failed")
        return out.reshape(self.shape) if count is None else out

    def at(self, offset):
        """address of element `offset` (flat)"""
        return self.ptr + int(offset) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            lib().cbns_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def calc_gradient(f):
    """calc_gradient(info, f, fx, fy, fz) of include/calc_gradient.h:9-10: the
    4th-order index derivatives (one-sided at the two outermost layers) of a
    node array f of shape (KM, JM, IM), i fastest; returns (fx, fy, fz)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    KM, JM, IM = f.shape
    out = [np.zeros_like(f) for _ in range(3)]
    rc = lib().cbns_calc_gradient(IM, JM, KM, _ptr(f), *(_ptr(o) for o in out))
    if rc:
        raise CbnsError(rc, "cbns_calc_gradient failed")
    return tuple(out)


def calc_gradient_noalloc(f):
    """calc_gradient without device allocation: the library's fixed staging
    array (at most NOALLOC_MAX_NODES nodes; CbnsError SIZE beyond)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    KM, JM, IM = f.shape
    out = [np.zeros_like(f) for _ in range(3)]
    rc = lib().cbns_calc_gradient_noalloc(IM, JM, KM, _ptr(f), *(_ptr(o) for o in out))
    if rc:
        raise CbnsError(rc, "cbns_calc_gradient_noalloc failed")
    return tuple(out)


def gather_rows(d_src, d_idx, n, row, d_dst):
    rc = lib().cbns_gather_rows(d_src, d_idx, int(n), int(row), d_dst)
    if rc:
        raise CbnsError(rc, "cbns_gather_rows failed")


class Mlp:
    """The indicator network alone, trained on the GPU (cbns_mlp_*)."""

    def __init__(self, params=None):
        h = C.c_void_p()
        rc = lib().cbns_mlp_create(C.byref(h))
        if rc:
            raise CbnsError(rc, "cbns_mlp_create failed")
        self._h = h
        if params is not None:
            self.set_params(params)

    def _chk(self, rc):
        if rc:
            msg = lib().cbns_mlp_last_error(self._h)
            raise CbnsError(rc, msg.decode() if msg else "")

    def free(self):
        if self._h:
            lib().cbns_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == MLP_NPARAM
        self._chk(lib().cbns_mlp_set_params(self._h, _fptr(params)))

    def get_params(self):
        out = np.zeros(MLP_NPARAM, dtype=np.float32)
        self._chk(lib().cbns_mlp_get_params(self._h, _fptr(out)))
        return out

    def forward(self, d_feat, n, d_logit):
        self._chk(lib().cbns_mlp_forward(self._h, d_feat, n, d_logit))

    def loss_grad(self, d_feat, d_target, n, w_pos=1.0, want_grad=False):
        loss = C.c_double(0.0)
        g = np.zeros(MLP_NPARAM, dtype=np.float32) if want_grad else None
        self._chk(lib().cbns_mlp_loss_grad(self._h, d_feat, d_target, n, float(w_pos),
                                           C.byref(loss), None if g is None else _fptr(g)))
        return (loss.value, g) if want_grad else loss.value

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self._chk(lib().cbns_mlp_adam_step(self._h, lr, beta1, beta2, eps))
