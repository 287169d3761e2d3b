"""Precision form (CBNS_FORM_FP32): face and viscous arithmetic in float on
local deviations, state, metrics, Jacobian, RK combination and dt in double.

  one stage, same input as the FP64 oracle       <= 1e-7 relative to each quantity's
                                                  scale (helpers.rel_err_state)
                                                  (FP32 rounding, 6e-8, of the local
                                                  stencil variation, scaled by the
                                                  CFL number; measured 1.4e-8)
  200 steps of the viscous Taylor-Green vortex   L2(rho) <= 1e-6 against the
  (64^3) against the FP64 GPU run                 plain FP64 form; L2/Linf of rho
                                                  and E reported
"""
import numpy as np
import pytest

from cbns3d_cuda_amd import cases
from cbns3d_cuda_amd.block import BUF_Q, F_PRIM, F_Q, F_QP, FORM_FP32, Block
from helpers import G, oracle_block, rel_err_state

pytestmark = pytest.mark.gpu


FP32_CASES = {
    "vortex": lambda: cases.isentropic_vortex(20),
    "tgv": lambda: cases.taylor_green(16),
    # lines along i crossing the zeta march's 58-node wavefronts (3 of them),
    # periodic and reference xi axis; all-reference C-grid
    "box_130x6x5_ppp_visc": lambda: cases.smooth_box(130, 6, 5, viscous=1),
    "box_125x7x6_rpp_visc": lambda: cases.smooth_box(
        125, 7, 6, viscous=1, bc=(cases.BC_REF, cases.BC_PERIODIC, cases.BC_PERIODIC)),
    "cgrid_viscous": lambda: cases.c_grid_channel(viscous=1),
}


@pytest.mark.parametrize("builder", list(FP32_CASES))
def test_fp32_stage_close_to_oracle(oracle, builder):
    case = FP32_CASES[builder]()
    o = oracle_block(oracle, case)
    o.bc_primitive()
    if case.params.viscous:
        o.viscous()
    dt = o.calc_dt()
    print("case", case.name, flush=True)
    b = Block(case.params)
    try:
        b.calc_metrics(case.x, case.y, case.z, case.period)
        b.set_form(FORM_FP32)
        b.set_field(F_Q, o.Q)
        b.set_field(F_QP, o.Qp)
        b.set_field(F_PRIM, o.prim)
        o.rk3_stage(dt, 1)
        b.rk3_stage(dt, 1)
        err = rel_err_state(b.get_field(F_QP, 5), o.Qp)
        print("fp32 stage-1 relative error", err, flush=True)
        assert err < 1e-7
    finally:
        b.free()


def _run(case, form, steps):
    b = Block(case.params)
    try:
        b.calc_metrics(case.x, case.y, case.z, case.period)
        b.set_form(form)
        b.set_conservative(cases.q_ref_aos(case))
        b.set_bc_calc_primitive(BUF_Q)
        t = 0.0
        for _ in range(steps):
            t, dt = b.rk3_step(1e9, t)
        return t, b.get_field(F_Q, 5)
    finally:
        b.free()


def test_fp32_tgv_200_steps_error():
    N = 64
    case = cases.taylor_green(N)
    print("case", case.name, flush=True)
    t64, q64 = _run(case, 0, 200)
    t32, q32 = _run(case, FORM_FP32, 200)
    inner = (slice(None),) + (slice(G, G + N),) * 3
    a, b = q32[inner], q64[inner]
    l2 = lambda c: float(np.sqrt(np.mean((a[c] - b[c]) ** 2)))
    linf = lambda c: float(np.abs(a[c] - b[c]).max())
    print("t %.6g vs %.6g; L2 rho %.3e E %.3e; Linf rho %.3e E %.3e"
          % (t32, t64, l2(0), l2(4), linf(0), linf(4)), flush=True)
    assert t32 == pytest.approx(t64, rel=1e-6)
    assert l2(0) <= 1e-6
