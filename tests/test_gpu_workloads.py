"""The learned- and combined-form workloads on the GPU, and what the trained
troubled-face indicator gains over the method without it.

  shock-vortex (Mach-2 shock, vortex, tanh-clustered curvilinear mesh)
    64x32x32       FP64 fused stage vs the oracle (1e-12 relative, the plain
                   form's bar)
    128x64x64      held-out seed (the network was trained on seeds 0-19):
                   FP32 + learned hybrid with the committed network, the
                   all-linear scheme (the method without the network: no
                   face flagged) and full FP32 WENO, each against the FP64
                   full-WENO run after 200 steps.  The network must flag a
                   minority of faces and cut the all-linear error by at
                   least 10x (the all-linear scheme breaks down at the
                   shock: its state turns non-physical within the run).
    256x128x128    the config's own size: hybrid vs FP64 after 40 steps
  isotropic turbulence (solenoidal k^4 exp(-2 (k/4)^2) field, Mt = 0.6,
  Re_lambda = 100, periodic, viscous)
    64^3           FP64 fused stage vs the oracle (1e-12 relative)
    128^3          hybrid vs FP64 after 100 steps: L2 rho <= 1e-6
Every error is L2 / Linf over the updated nodes (learned.compare).
"""
import os

import numpy as np
import pytest

from cbns3d_cuda_amd import cases, learned
from cbns3d_cuda_amd.block import F_PRIM, F_Q, F_QP, FORM_FP32, FORM_HYBRID, Block
from helpers import oracle_block, rel_err

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "cbns3d_cuda_amd", "indicator_weights.npz")
HYB = FORM_FP32 | FORM_HYBRID


def _fused_stages_match(oracle, case):
    """One step's three fused stages on the GPU, each from the oracle's own
    stage input, against the oracle's stage output."""
    o = oracle_block(oracle, case)
    o.bc_primitive()
    if case.params.viscous:
        o.viscous()
    dt = o.calc_dt()
    b = Block(case.params)
    try:
        b.calc_metrics(case.x, case.y, case.z, case.period)
        for s in (1, 2, 3):
            b.set_field(F_Q, o.Q)
            b.set_field(F_QP, o.Qp)
            b.set_field(F_PRIM, o.prim)
            o.rk3_stage(dt, s)
            b.rk3_stage(dt, s)
            ref = o.Q if s == 3 else o.Qp
            err = rel_err(b.get_field(F_Q if s == 3 else F_QP, 5), ref)
            print("%s stage %d: rel err %.3e" % (case.name, s, err), flush=True)
            assert err < 1e-12, s
    finally:
        b.free()


def test_shock_vortex_stage_matches_oracle(oracle):
    case = cases.shock_vortex(64, 32, 32, seed=3)
    print("case", case.name, flush=True)
    _fused_stages_match(oracle, case)


def test_turbulence_stage_matches_oracle(oracle):
    case = cases.isotropic_turbulence(64, seed=2)
    print("case", case.name, flush=True)
    _fused_stages_match(oracle, case)


def test_trained_indicator_gain_shock_vortex():
    """The committed network against the method without it, on a seed it was
    not trained on."""
    P, thr, meta = learned.load(WEIGHTS)
    case = cases.shock_vortex(128, 64, 64, seed=1000)
    print("case", case.name, "threshold", thr, flush=True)
    steps = 200
    hyb = learned.compare(case, steps, HYB, P, thr)
    lin = learned.compare(case, steps, HYB, P, 1e30)  # no face flagged: all linear
    weno = learned.compare(case, steps, FORM_FP32)
    print("trained hybrid: L2 rho %.3e, flagged %.4f" % (hyb["L2_rho"], hyb["troubled_fraction_mean"]))
    print("all-linear:     L2 rho %.3e%s" % (lin["L2_rho"], ", diverged at step %d"
                                            % lin["diverged_step"] if "diverged_step" in lin else ""))
    print("full WENO FP32: L2 rho %.3e" % weno["L2_rho"], flush=True)
    for r in (hyb, weno):
        assert np.isfinite(r["L2_rho"]) and np.isfinite(r["L2_E"])
    # without the network the linear scheme either breaks down at the shock
    # (non-physical state, reported as a NaN time step) or errs far more
    assert "diverged_step" in lin or lin["troubled_fraction_mean"] == 0.0
    assert hyb["troubled_fraction_mean"] < 0.5
    assert hyb["L2_rho"] <= 0.1 * lin["L2_rho"]
    assert weno["L2_rho"] <= hyb["L2_rho"]


def test_shock_vortex_config_size():
    P, thr, _ = learned.load(WEIGHTS)
    case = cases.shock_vortex(256, 128, 128, seed=0)
    print("case", case.name, flush=True)
    r = learned.compare(case, 40, HYB, P, thr)
    print("hybrid vs FP64 after 40 steps: %s" % r, flush=True)
    assert np.isfinite(r["L2_rho"]) and np.isfinite(r["Linf_E"])
    assert 0.0 < r["troubled_fraction_mean"] < 0.5
    assert r["L2_rho"] < 1e-5


def test_turbulence_hybrid_error():
    P, thr, _ = learned.load(WEIGHTS)
    case = cases.isotropic_turbulence(128, seed=0)
    print("case", case.name, flush=True)
    r = learned.compare(case, 100, HYB, P, thr)
    print("hybrid vs FP64 after 100 steps: %s" % r, flush=True)
    assert np.isfinite(r["L2_rho"]) and np.isfinite(r["Linf_E"])
    assert r["L2_rho"] <= 1e-6
