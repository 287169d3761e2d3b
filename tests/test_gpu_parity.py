"""GPU parity: the HIP library (through the C ABI) against the CPU oracle on
identical inputs.

Bars:
  metrics, boundary/ghost fill, primitives      bit-exact
  time step                                     bit-exact
  face fluxes, viscous divergence, RK stage     1e-12 relative (per component,
                                                normalised by its max), same
                                                input Q on both sides
  three full steps                              1e-10 relative
"""
import numpy as np
import pytest

from cbns3d_cuda_amd import cases
from cbns3d_cuda_amd.block import (BC_PERIODIC, BC_REF, BUF_Q, BUF_QP, F_DFV, F_FLUX, F_MET, F_MU,
                                   F_PRIM, F_Q, F_QP, Block)
from helpers import G, oracle_block, rel_err, update_mask

pytestmark = pytest.mark.gpu

STAGE_TOL = 1e-12

CASES = {
    "vortex": lambda: cases.isentropic_vortex(20),
    "tgv_ref": lambda: cases.taylor_green(16),
    "tgv_centered": lambda: cases.taylor_green(16, visc_centered=1),
    "cgrid_inviscid": lambda: cases.c_grid_channel(viscous=0),
    "cgrid_viscous": lambda: cases.c_grid_channel(viscous=1),
    # edge sizes: the minimum, lines crossing the 63-node wavefront tiling,
    # mixed boundary kinds
    "box_5x6x7": lambda: cases.smooth_box(5, 6, 7, viscous=0),
    "box_65x6x5_visc": lambda: cases.smooth_box(65, 6, 5, viscous=1),
    "box_70x9x8_rpp_visc": lambda: cases.smooth_box(70, 9, 8, viscous=1,
                                                    bc=(BC_REF, BC_PERIODIC, BC_PERIODIC)),
    "box_7x9x6_rrr_visc": lambda: cases.smooth_box(7, 9, 6, viscous=1, bc=(BC_REF,) * 3),
    # lines crossing the viscous sweep's tiling: 60-column tiles, 4-row
    # tiles, 32-plane chunks
    "box_62x9x34_visc": lambda: cases.smooth_box(62, 9, 34, viscous=1),
    "box_63x10x35_rpr_visc": lambda: cases.smooth_box(63, 10, 35, viscous=1,
                                                      bc=(BC_REF, BC_PERIODIC, BC_REF)),
    # centred divergence on node pairs: lines of 66 nodes (one paired tile)
    "box_66x9x8_centred_visc": lambda: cases.smooth_box(66, 9, 8, viscous=1, visc_centered=1),
}


def gpu_block(case, o):
    """GPU block with the oracle's exact state (all ghosted storage)."""
    print("case", case.name, flush=True)
    b = Block(case.params)
    b.calc_metrics(case.x, case.y, case.z, case.period)
    put_state(b, o)
    return b


def put_state(b, o):
    b.set_field(F_Q, o.Q)
    b.set_field(F_QP, o.Qp)
    b.set_field(F_PRIM, o.prim)


@pytest.fixture(params=list(CASES))
def pair(request, oracle):
    case = CASES[request.param]()
    o = oracle_block(oracle, case)
    b = gpu_block(case, o)
    yield case, o, b
    b.free()


def test_metrics_bit_exact(pair):
    case, o, b = pair
    met = b.get_field(F_MET, 10)
    assert np.array_equal(met, o.met)


def test_bc_primitive_bit_exact(pair):
    case, o, b = pair
    for buf in (BUF_Q, BUF_QP):
        o.bc_primitive(o.Q if buf == BUF_Q else o.Qp)
        b.set_bc_calc_primitive(buf)
        got_q = b.get_field(F_Q if buf == BUF_Q else F_QP, 5)
        assert np.array_equal(got_q, o.Q if buf == BUF_Q else o.Qp)
        nprim = 6 if case.params.viscous else 5
        assert np.array_equal(b.get_field(F_PRIM, 6)[:nprim], o.prim[:nprim])


def test_calc_dt_bit_exact(pair):
    case, o, b = pair
    o.bc_primitive()
    put_state(b, o)
    if case.params.viscous:
        o.viscous()  # fills the oracle's mu array
    assert b.calc_dt() == o.calc_dt()


def test_viscous_terms_mu_bit_exact(pair):
    """cbns_calc_viscous_terms leaves Sutherland's mu over the viscous node
    range [-2, N+1]^3, as calc_viscous_terms does (src/viscous_terms_kernel.cu:61-66)."""
    case, o, b = pair
    if not case.params.viscous:
        pytest.skip("inviscid block: calc_viscous_terms does nothing")
    o.bc_primitive()
    put_state(b, o)
    o.viscous()
    b.calc_viscous_terms()
    p = case.params
    ext = tuple(slice(G - 2, G + n + 2) for n in (p.KM, p.JM, p.IM))
    assert np.array_equal(b.get_field(F_MU)[ext], o.mu[ext])


def _stage_inputs(o, dt):
    """Advance the oracle to each stage's input; yield (stage, snapshot)."""
    snaps = []
    for s in (1, 2, 3):
        snaps.append((s, o.Q.copy(), o.Qp.copy(), o.prim.copy()))
        o.rk3_stage(dt, s)
    return snaps


def test_unfused_stage_matches_oracle(pair, oracle):
    case, o, b = pair
    o.bc_primitive()
    dt = 0.8 * o.calc_dt() if not case.params.viscous else (o.viscous(), o.calc_dt())[1]
    for s, Q, Qp, prim in _stage_inputs(o, dt):
        o2 = oracle_block(oracle, case)
        o2.Q[...], o2.Qp[...], o2.prim[...] = Q, Qp, prim
        put_state(b, o2)
        buf = BUF_Q if s == 1 else BUF_QP
        # face fluxes
        flux_ref = o2.numerical_flux(Q if s == 1 else Qp)
        b.calc_numerical_flux(buf)
        flux = b.get_field(F_FLUX, 15).reshape(flux_ref.shape)
        for d in range(3):
            m = face_mask(case, d)
            assert rel_err(flux[d][:, m], flux_ref[d][:, m]) < STAGE_TOL, (s, d)
        if case.params.viscous:
            dfv_ref = o2.viscous()
            b.calc_viscous_flux_contribution()
            m = update_mask(case)
            assert rel_err(b.get_field(F_DFV, 4)[:, m], dfv_ref[:, m]) < STAGE_TOL
        o2.rk3_stage(dt, s)
        b.update_rk3(dt, s)
        b.set_bc_calc_primitive(BUF_Q if s == 3 else BUF_QP)
        out_ref = o2.Q if s == 3 else o2.Qp
        out = b.get_field(F_Q if s == 3 else F_QP, 5)
        assert rel_err(out, out_ref) < STAGE_TOL, s


def face_mask(case, d):
    p = case.params
    shape = (p.KM + 2 * G, p.JM + 2 * G, p.IM + 2 * G)
    m = np.zeros(shape, dtype=bool)
    sl = []
    for a, n in ((2, p.KM), (1, p.JM), (0, p.IM)):
        per = p.bc[a] == 1
        if a == d:
            lo, hi = (-1, n - 1) if per else (0, n - 2)
        else:
            lo, hi = (0, n - 1) if per else (1, n - 2)
        sl.append(slice(G + lo, G + hi + 1))
    m[tuple(sl)] = True
    return m


def test_fused_stage_matches_oracle(pair, oracle):
    case, o, b = pair
    o.bc_primitive()
    dt = 0.8 * o.calc_dt() if not case.params.viscous else (o.viscous(), o.calc_dt())[1]
    for s, Q, Qp, prim in _stage_inputs(o, dt):
        o2 = oracle_block(oracle, case)
        o2.Q[...], o2.Qp[...], o2.prim[...] = Q, Qp, prim
        put_state(b, o2)
        o2.rk3_stage(dt, s)
        b.rk3_stage(dt, s)
        out_ref = o2.Q if s == 3 else o2.Qp
        out = b.get_field(F_Q if s == 3 else F_QP, 5)
        assert rel_err(out, out_ref) < STAGE_TOL, s
        nprim = 6 if case.params.viscous else 5
        assert rel_err(b.get_field(F_PRIM, 6)[:nprim], o2.prim[:nprim]) < STAGE_TOL, s


def test_three_steps_match_oracle(pair):
    case, o, b = pair
    o.bc_primitive()
    b.set_bc_calc_primitive(BUF_Q)
    t_o = t_b = 0.0
    for n in range(3):
        t_o, dt_o = o.rk3_step(1e9, t_o)
        t_b, dt_b = b.rk3_step(1e9, t_b)
        assert dt_b == pytest.approx(dt_o, rel=1e-12)
    assert rel_err(b.get_field(F_Q, 5), o.Q) < 1e-10
