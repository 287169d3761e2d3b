"""The learned form on the GPU against the CPU oracle (oracle/orc_learned.c,
orc_rk3_stage_hybrid):

  network forward      bf16 matrix-core logits vs the oracle's bf16-mode forward
  loss and gradient    vs the oracle's bf16-mode loss/gradient (same rounding
                       points, fp32 vs double accumulation): relative L2 <= 1e-3
  Adam                 vs the oracle's Adam on the GPU's own gradient
  training samples     features (float) and labels (FP64 weights, stored float)
                       of every face vs the oracle's, per axis
  hybrid stage         with the GPU's own troubled-face bytes, the FP64 hybrid
                       stage (linear marches + listed WENO faces) vs the
                       oracle's hybrid stage (1e-12 relative, the plain-form
                       bar); every face troubled -> the plain stage (through
                       the lists, and through the full-WENO fallback);
                       no face troubled -> the oracle's all-linear stage; the
                       FP32 hybrid within the FP32 form's 1e-7
  flags                the GPU's flags are logit > threshold of the oracle's
                       network on the GPU's features, except within 1e-3 of the
                       threshold
"""
import numpy as np
import pytest

from cbns3d_cuda_amd import cases
from cbns3d_cuda_amd.block import (BUF_Q, F_PRIM, F_Q, F_QP, FORM_FP32, FORM_HYBRID, MLP_NPARAM,
                                   Block, DeviceArray, Mlp)
from helpers import oracle_block, rel_err, rel_err_state

pytestmark = pytest.mark.gpu


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    P = (rng.standard_normal(MLP_NPARAM) * scale).astype(np.float32)
    for r in range(32):
        P[r * 32 + 30:r * 32 + 32] = 0
    return P


def random_features(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 32)).astype(np.float32)
    X[:, 30:] = 0
    y = (rng.random(n) < 0.3).astype(np.float32)
    return X, y


@pytest.mark.parametrize("n", [64, 1000, 20000])
def test_mlp_forward_matches_oracle(oracle, n):
    print("mlp forward n", n, flush=True)
    P = random_params(1)
    X, _ = random_features(2, n)
    m = Mlp(P)
    dX = DeviceArray.from_host(X)
    dS = DeviceArray((n,), np.float32)
    m.forward(dX.ptr, n, dS.ptr)
    s = dS.download()
    ref = oracle.mlp_forward(P, X, bf16=True)
    err = np.abs(s - ref)
    print("max |logit diff|", err.max(), "p99", np.quantile(err, 0.99), flush=True)
    assert err.max() < 1e-3 * (1 + np.abs(ref).max())
    assert np.quantile(err, 0.99) < 1e-5 * (1 + np.abs(ref).max())


@pytest.mark.parametrize("n", [100, 70000])
def test_mlp_gradient_and_adam_match_oracle(oracle, n):
    print("mlp grad n", n, flush=True)
    P = random_params(3)
    X, y = random_features(4, n)
    m = Mlp(P)
    dX, dY = DeviceArray.from_host(X), DeviceArray.from_host(y)
    loss, g = m.loss_grad(dX.ptr, dY.ptr, n, w_pos=4.0, want_grad=True)
    loss_o, g_o = oracle.mlp_loss_grad(P, X, y, w_pos=4.0, bf16=True)
    rel = np.linalg.norm(g - g_o) / np.linalg.norm(g_o)
    print("loss", loss, loss_o, "grad rel L2", rel, flush=True)
    assert loss == pytest.approx(loss_o, rel=1e-5)
    assert rel < 1e-3
    for blk in ((0, 1024), (1024, 1056), (1056, 2080), (2080, 2112), (2112, 2145)):
        a, b = g[blk[0]:blk[1]], g_o[blk[0]:blk[1]]
        assert np.linalg.norm(a - b) <= 2e-3 * np.linalg.norm(b) + 1e-7, blk
    # Adam from zero moments on the GPU's gradient
    m.adam_step(1e-2)
    Pn = m.get_params()
    Pe = P.copy()
    mo, vo = np.zeros(MLP_NPARAM), np.zeros(MLP_NPARAM)
    oracle.adam(Pe, g.astype(np.float64), mo, vo, lr=1e-2, t=1)
    pad = [r * 32 + c for r in range(32) for c in (30, 31)]
    Pe[pad] = 0
    assert np.abs(Pn - Pe).max() < 1e-6


@pytest.mark.parametrize("builder", ["vortex", "channel"])
def test_face_samples_match_oracle(oracle, builder):
    case = cases.isentropic_vortex(12) if builder == "vortex" else cases.c_grid_channel(viscous=0)
    print("case", case.name, flush=True)
    o = oracle_block(oracle, case)
    o.bc_primitive()
    b = Block(case.params)
    try:
        b.calc_metrics(case.x, case.y, case.z, case.period)
        b.set_field(F_Q, o.Q)
        for axis in range(3):
            n = b.face_count(axis)
            dF, dL = DeviceArray((n, 32)), DeviceArray((n,))
            b.face_samples(axis, dF.ptr, dL.ptr)
            feat, lab = dF.download(), dL.download()
            feat_o, lab_o = o.face_features_all(axis)
            assert feat_o.shape == feat.shape
            ferr = np.abs(feat - feat_o).max()
            lerr = np.abs(lab - lab_o).max()
            print("axis", axis, "n", n, "feature err", ferr, "label err", lerr,
                  "label max", lab_o.max(), flush=True)
            assert ferr < 2e-4
            assert lerr < 1e-6
    finally:
        b.free()


def _stage_pair(oracle, case, params, thresh, form, full_lists=True):
    o = oracle_block(oracle, case)
    o.bc_primitive()
    if case.params.viscous:
        o.viscous()
    dt = o.calc_dt()
    b = Block(case.params)
    b.calc_metrics(case.x, case.y, case.z, case.period)
    b.set_indicator(params, thresh, count=True, full_lists=full_lists)
    b.set_form(form)
    b.set_field(F_Q, o.Q)
    b.set_field(F_QP, o.Qp)
    b.set_field(F_PRIM, o.prim)
    b.rk3_stage(dt, 1)
    return o, b, dt


@pytest.mark.parametrize("builder", ["vortex", "tgv", "channel", "box_rpp", "box_odd"])
def test_hybrid_stage_matches_oracle(oracle, builder):
    case = {"vortex": lambda: cases.isentropic_vortex(16), "tgv": lambda: cases.taylor_green(16),
            "channel": lambda: cases.c_grid_channel(viscous=1),
            "box_rpp": lambda: cases.smooth_box(70, 9, 8, viscous=1, bc=(0, 1, 1)),
            "box_odd": lambda: cases.smooth_box(65, 6, 5, viscous=0)}[builder]()
    print("case", case.name, flush=True)
    P = random_params(11)
    # threshold at the median logit of the oracle's network on this state's faces
    o0 = oracle_block(oracle, case)
    o0.bc_primitive()
    feats = np.concatenate([o0.face_features_all(a)[0] for a in range(3)])
    thr = float(np.median(oracle.mlp_forward(P, feats)))
    o, b, dt = _stage_pair(oracle, case, P, thr, FORM_HYBRID)
    try:
        flags = b.get_flags()
        counts = b.flag_count()
        print("troubled faces per axis", counts, flush=True)
        assert 0 < sum(counts) < len(feats) + 6 * 64 ** 2
        o.rk3_stage_hybrid(dt, 1, flags)
        err = rel_err(b.get_field(F_QP, 5), o.Qp)
        print("fp64 hybrid stage relative error", err, flush=True)
        assert err < 1e-12
        # flags agree with the oracle's network on the GPU's own features
        for axis in range(3):
            n = b.face_count(axis)
            dF, dL = DeviceArray((n, 32)), DeviceArray((n,))
            b.face_samples(axis, dF.ptr, dL.ptr)
            s = oracle.mlp_forward(P, dF.download())
            gpu = face_values(b, flags[axis], axis)
            sure = np.abs(s - thr) > 1e-3
            assert np.array_equal(gpu[sure], (s > thr)[sure]), axis
    finally:
        b.free()


def face_values(b, arr, axis):
    """values of a ghosted per-node array at the faces of `axis`, in sample order"""
    from helpers import G
    p = b.params
    lo, hi = [], []
    for a, n in enumerate((p.IM, p.JM, p.KM)):
        per = p.bc[a] == 1
        if a == axis:
            lo.append(-1 if per else 0)
            hi.append(n - 1 if per else n - 2)
        else:
            lo.append(0 if per else 1)
            hi.append(n - 1 if per else n - 2)
    return arr[lo[2] + G:hi[2] + G + 1, lo[1] + G:hi[1] + G + 1, lo[0] + G:hi[0] + G + 1].ravel()


@pytest.mark.parametrize("bias,expect,full_lists", [(100.0, "weno", True), (100.0, "weno", False),
                                                    (-100.0, "linear", True)])
def test_hybrid_extremes(oracle, bias, expect, full_lists):
    """Every face troubled: the listed WENO faces (lists that hold every face)
    or the fallback to WENO at every face (lists of a third of the faces)
    give the plain stage; no face troubled: the all-linear stage."""
    case = cases.isentropic_vortex(16)
    print("case", case.name, "bias", bias, "full lists", full_lists, flush=True)
    P = np.zeros(MLP_NPARAM, dtype=np.float32)
    P[2144] = bias
    o, b, dt = _stage_pair(oracle, case, P, 0.0, FORM_HYBRID, full_lists)
    try:
        flags = b.get_flags()
        if expect == "weno":
            o.rk3_stage(dt, 1)
        else:
            assert not flags.any()
            o.rk3_stage_hybrid(dt, 1, flags)
        err = rel_err(b.get_field(F_QP, 5), o.Qp)
        print(expect, "stage relative error", err, flush=True)
        assert err < 1e-12
    finally:
        b.free()


def test_overflowed_lists_back_off():
    """Every face troubled with lists of a third of an axis' faces: each step
    runs WENO at every face, so the indicator pass runs only every 8 steps
    (steps 1 and 9 of 10), and the state is the plain form's, bit for bit."""
    case = cases.isentropic_vortex(16)
    print("case", case.name, flush=True)
    P = np.zeros(MLP_NPARAM, dtype=np.float32)
    P[2144] = 100.0  # output bias: every logit above the threshold
    states = []
    for hybrid in (True, False):
        b = Block(case.params)
        try:
            b.calc_metrics(case.x, case.y, case.z, case.period)
            b.set_conservative(cases.q_ref_aos(case))
            b.set_bc_calc_primitive(BUF_Q)
            if hybrid:
                b.set_indicator(P, 0.0)
                b.set_form(FORM_HYBRID)
                b.set_profiling(True)
            t = 0.0
            for _ in range(10):
                t, _ = b.rk3_step(1e9, t)
            if hybrid:
                passes = b.kernel_times(reset=False)["indicator"][1]
                print("indicator passes in 10 steps:", passes, flush=True)
                assert passes == 2
            states.append(b.get_field(F_Q, 5))
        finally:
            b.free()
    assert np.array_equal(states[0], states[1])


@pytest.mark.parametrize("builder", ["tgv", "box_rpp"])
def test_fp32_hybrid_stage_close_to_oracle(oracle, builder):
    # box_rpp: lines along i crossing the zeta march's 58-node wavefronts,
    # reference xi axis
    case = {"tgv": lambda: cases.taylor_green(16),
            "box_rpp": lambda: cases.smooth_box(125, 7, 6, viscous=1, bc=(0, 1, 1))}[builder]()
    print("case", case.name, flush=True)
    P = random_params(12)
    o, b, dt = _stage_pair(oracle, case, P, 0.0, FORM_FP32 | FORM_HYBRID)
    try:
        o.rk3_stage_hybrid(dt, 1, b.get_flags())
        err = rel_err_state(b.get_field(F_QP, 5), o.Qp)
        print("fp32 hybrid stage relative error", err, flush=True)
        assert err < 1e-7
    finally:
        b.free()


def test_flag_reuse_over_steps(oracle):
    """options bit 2: the flags of a step's indicator pass serve floor(1/CFL)
    = 2 steps; the second step equals the oracle's hybrid step with them."""
    case = cases.isentropic_vortex(16)  # CFL 0.5
    print("case", case.name, flush=True)
    P = random_params(13)
    o = oracle_block(oracle, case)
    o.bc_primitive()
    b = Block(case.params)
    try:
        b.calc_metrics(case.x, case.y, case.z, case.period)
        b.set_indicator(P, 0.0, count=True, reuse=True, full_lists=True)
        b.set_form(FORM_HYBRID)
        b.set_profiling(True)
        b.set_field(F_Q, o.Q)
        b.set_field(F_QP, o.Qp)
        b.set_field(F_PRIM, o.prim)
        t = 0.0
        t, dt1 = b.rk3_step(1e9, t)
        assert b.kernel_times(reset=False)["indicator"][1] == 1
        f1 = b.get_flags()
        assert 0 < f1.sum()
        # the oracle follows the GPU's first step with the same flags
        dt = o.calc_dt()
        assert dt == pytest.approx(dt1, rel=1e-12)
        for s in (1, 2, 3):
            o.rk3_stage_hybrid(dt, s, f1)
        t, dt2 = b.rk3_step(1e9, t)
        assert np.array_equal(b.get_flags(), f1)  # not re-evaluated
        assert b.kernel_times(reset=False)["indicator"][1] == 1
        dt = o.calc_dt()
        for s in (1, 2, 3):
            o.rk3_stage_hybrid(dt, s, f1)
        err = rel_err(b.get_field(F_Q, 5), o.Q)
        print("two steps with reused flags, relative error", err, flush=True)
        assert err < 1e-11
        t, _ = b.rk3_step(1e9, t)  # third step: a new pass
        assert b.kernel_times(reset=False)["indicator"][1] == 2
    finally:
        b.free()
