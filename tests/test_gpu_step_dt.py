"""The time step of cbns_rk3_step, and the primitives a step leaves behind.

The step forms its dt on the device, where the stages read it: on viscous
blocks the first stage's viscous flux pass (the per-node spectral-radius sum
of src/time_step_kernel.cu:43-94 and its maximum), on inviscid blocks the
time-step pass of cbns_calc_dt, with no host round trip before the stages.
It must equal cbns_calc_dt bit for bit, including the limit to the rest of
the run (src/Block3d.cpp:241-246), and a NaN time step must be reported with
the state left as it was (the reference throws before the update,
src/calc_dt.cu:56-58).
"""
import numpy as np
import pytest

from cbns3d_cuda_amd import cases
from cbns3d_cuda_amd.block import BUF_Q, F_PRIM, F_Q, FORM_FP32, Block, CbnsError

pytestmark = pytest.mark.gpu

CASES = {
    "tgv": lambda: cases.taylor_green(16),
    "cgrid_viscous": lambda: cases.c_grid_channel(viscous=1),
    "box_70x9x8_rpp_visc": lambda: cases.smooth_box(
        70, 9, 8, viscous=1, bc=(cases.BC_REF, cases.BC_PERIODIC, cases.BC_PERIODIC)),
    "vortex": lambda: cases.isentropic_vortex(16),
    "cgrid_inviscid": lambda: cases.c_grid_channel(viscous=0),
    "box_70x9x8_rpp": lambda: cases.smooth_box(
        70, 9, 8, viscous=0, bc=(cases.BC_REF, cases.BC_PERIODIC, cases.BC_PERIODIC)),
}


def _block(case, form=0):
    b = Block(case.params)
    b.calc_metrics(case.x, case.y, case.z, case.period)
    b.set_form(form)
    b.set_conservative(cases.q_ref_aos(case))
    b.set_bc_calc_primitive(BUF_Q)
    return b


@pytest.mark.parametrize("form", [0, FORM_FP32])
@pytest.mark.parametrize("name", list(CASES))
def test_step_dt_equals_calc_dt(name, form):
    case = CASES[name]()
    print("case", case.name, "form", form, flush=True)
    b = _block(case, form)
    try:
        t = 0.0
        for _ in range(3):
            dt_ref = b.calc_dt()
            t_new, dt = b.rk3_step(1e9, t)
            assert dt == dt_ref  # bit for bit
            assert t_new == t + dt_ref
            t = t_new
        # limited to the rest of the run
        dt_ref = b.calc_dt()
        t_end = t + 0.25 * dt_ref
        t_new, dt = b.rk3_step(t_end, t)
        assert dt == t_end - t
        assert t_new == t + (t_end - t)
    finally:
        b.free()


@pytest.mark.parametrize("name", ["tgv", "vortex", "box_70x9x8_rpp"])
def test_nan_step_reported_and_state_kept(name):
    case = CASES[name]()
    print("case", case.name, flush=True)
    b = _block(case)
    try:
        t, _ = b.rk3_step(1e9, 0.0)
        q0 = b.get_field(F_Q, 5)
        prim = b.get_field(F_PRIM, 6)
        prim[4][tuple(s // 2 for s in prim[4].shape)] = np.nan  # one node's pressure
        b.set_field(F_PRIM, prim)
        with pytest.raises(CbnsError) as e:
            b.rk3_step(1e9, t)
        assert e.value.code == 3  # CBNS_ERR_NAN
        assert np.array_equal(b.get_field(F_Q, 5), q0)
    finally:
        b.free()


@pytest.mark.parametrize("name", ["tgv", "cgrid_viscous", "box_70x9x8_rpp_visc", "vortex",
                                  "cgrid_inviscid"])
def test_step_leaves_primitives_of_its_state(name):
    """cbns_rk3_step's first two stages write only the primitives the next
    stage reads; after the step every primitive is that of the new state
    (as set_bc_calc_primitive derives them, to rounding: the stage forms
    1/rho by Newton steps, the boundary pass by division)."""
    case = {**CASES, "vortex": lambda: cases.isentropic_vortex(20),
            "cgrid_inviscid": lambda: cases.c_grid_channel(viscous=0)}[name]()
    print("case", case.name, flush=True)
    b = _block(case)
    try:
        t = 0.0
        for _ in range(2):
            t, _ = b.rk3_step(1e9, t)
        n = 6 if case.params.viscous else 5
        p_step = b.get_field(F_PRIM, n)
        b.set_bc_calc_primitive(BUF_Q)
        p_ref = b.get_field(F_PRIM, n)
        for f in range(n):
            scale = np.abs(p_ref[f]).max()
            err = np.abs(p_step[f] - p_ref[f]).max() / scale
            print("field", f, "max relative difference", err, flush=True)
            assert err < 1e-14
    finally:
        b.free()
