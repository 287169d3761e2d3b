"""Pins of the oracle's grid drivers (CPU): the reference's own gradient and
metrics gtests, closed forms of the periodic stencils, brute-force periodic
ghost fill, invariants of the reference's boundary rules, closed-form time
step and viscous divergence, the SSP-RK3 stability polynomial, and discrete
conservation / axis-permutation symmetry of a full stage."""
import numpy as np
import pytest

import ref_cases
from helpers import G, oracle_block, oracle_params
from cbns3d_cuda_amd import cases


# -------------------------------------------------- reference gtests
@pytest.mark.parametrize("case", [ref_cases.gradient_fun01, ref_cases.gradient_fun02])
def test_gradient_reference_gtests(oracle, case):
    f, exp, tol = case()
    got = oracle.gradient_nodes(f)
    for g, e in zip(got, exp):
        assert np.abs(g - e).max() < tol
        assert np.abs(g - e).max() < 1e-9  # exact for degree <= 4


@pytest.mark.parametrize("case", [ref_cases.metrics_rectilinear, ref_cases.metrics_rotation,
                                  ref_cases.metrics_conical_shell])
def test_metrics_reference_gtests(oracle, case):
    (x, y, z), exp, tol = case()
    KM, JM, IM = x.shape
    o = oracle.Oracle(oracle.make_params(IM, JM, KM, bc=(0, 0, 0)))
    o.set_mesh(x, y, z)
    got = [o.inner(o.met[c]) for c in range(10)]
    mse = ref_cases.metric_mse(got, exp)
    assert max(mse) <= tol, mse


def test_metrics_ghost_extension_copies_nearest_node(oracle):
    (x, y, z), _, _ = ref_cases.metrics_conical_shell()
    KM, JM, IM = x.shape
    o = oracle.Oracle(oracle.make_params(IM, JM, KM, bc=(0, 0, 0)))
    o.set_mesh(x, y, z)
    met = o.met
    # faces, an edge and a corner: the nearest boundary node, bit for bit
    assert np.array_equal(met[:, G + 3, G + 5, 0:G], np.repeat(met[:, G + 3, G + 5, G:G + 1], G, -1))
    assert np.array_equal(met[:, 0, 1, G + 2], met[:, G, G, G + 2])
    assert np.array_equal(met[:, -1, -1, -1], met[:, G + KM - 1, G + JM - 1, G + IM - 1])


def test_metrics_periodic_closed_form(oracle):
    """x = i h + A sin(w j) sin(w k) (cyclic): the 4th-order central stencil of
    sin(w s) is exactly cos(w s) (4/3 sin w - 1/6 sin 2w)."""
    N, A, h = 16, 0.3, 0.5
    w = 2 * np.pi / N
    k, j, i = np.meshgrid(np.arange(N, dtype=float), np.arange(N, dtype=float),
                          np.arange(N, dtype=float), indexing="ij")
    x = i * h + A * np.sin(w * j) * np.sin(w * k)
    y = j * h + A * np.sin(w * k) * np.sin(w * i)
    z = k * h + A * np.sin(w * i) * np.sin(w * j)
    o = oracle.Oracle(oracle.make_params(N, N, N, bc=(1, 1, 1)))
    o.set_mesh(x, y, z, np.eye(3) * N * h)
    D = 4.0 / 3.0 * np.sin(w) - 1.0 / 6.0 * np.sin(2 * w)
    Jm = np.zeros((N, N, N, 3, 3))  # d(x,y,z)/d(xi,eta,zeta)
    Jm[..., 0, 0] = h
    Jm[..., 0, 1] = A * D * np.cos(w * j) * np.sin(w * k)
    Jm[..., 0, 2] = A * D * np.sin(w * j) * np.cos(w * k)
    Jm[..., 1, 1] = h
    Jm[..., 1, 2] = A * D * np.cos(w * k) * np.sin(w * i)
    Jm[..., 1, 0] = A * D * np.sin(w * k) * np.cos(w * i)
    Jm[..., 2, 2] = h
    Jm[..., 2, 0] = A * D * np.cos(w * i) * np.sin(w * j)
    Jm[..., 2, 1] = A * D * np.sin(w * i) * np.cos(w * j)
    inv = np.linalg.inv(Jm)
    exp = [inv[..., r, c] for r in range(3) for c in range(3)] + [np.linalg.det(Jm)]
    for c in range(10):
        assert np.abs(o.inner(o.met[c]) - exp[c]).max() < 1e-12 * max(1.0, np.abs(exp[c]).max())
    # periodic ghosts are the wrapped nodes
    assert np.array_equal(o.met[:, 0, G + 2, G + 3], o.met[:, G + N - G, G + 2, G + 3])


# ------------------------------------------------------- boundary rules
def _prim_of(q, g):
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    p = g.gam1 * (q[4] - 0.5 * rho * (u * u + v * v + w * w))
    return np.stack([rho, u, v, w, p, g.gM2 * p / rho])


def test_bc_periodic_bruteforce(oracle):
    case = cases.taylor_green(12)
    o = oracle_block(oracle, case)
    rng = np.random.default_rng(0)
    o.Q[...] *= 1 + 0.01 * rng.standard_normal(o.Q.shape)
    q_nodes = o.inner(o.Q).copy()
    o.bc_primitive()
    N = 12
    idx = (np.arange(N + 2 * G) - G) % N
    expect_q = q_nodes[:, idx][:, :, idx][:, :, :, idx]
    assert np.array_equal(o.Q, expect_q)
    assert np.array_equal(o.prim, _prim_of(expect_q, o.g))


@pytest.mark.parametrize("viscous", [0, 1])
def test_bc_reference_rules(oracle, viscous):
    case = cases.c_grid_channel(viscous=viscous)
    o = oracle_block(oracle, case)
    p = case.params
    IM, JM, KM = p.IM, p.JM, p.KM
    g = o.g
    before = o.Q.copy()
    o.bc_primitive()
    Q, P = o.Q, o.prim

    def at(a, i, j, k):
        return a[..., G + k, G + j, G + i]

    for k in range(1, KM - 1):
        for j in range(1, JM - 1):
            # xi far field: rho, momentum from the neighbour, p = p_inf
            for i0, src, sgn in ((0, 1, -1), (IM - 1, IM - 2, 1)):
                for li in range(3):
                    q = at(Q, i0 + sgn * li, j, k)
                    assert np.array_equal(q[:4], at(Q, src, j, k)[:4])
                assert at(P, i0, j, k)[4] == pytest.approx(g.p_inf, rel=1e-14)
        for i in range(IM):
            if i < p.i_begin - 1:  # wake cut: mirror-averaged and cross-copied
                assert np.array_equal(at(Q, i, 0, k), at(Q, IM - 1 - i, 0, k))
                pa, pb = _prim_of(at(Q, i, 1, k)[:, None], g)[:5, 0], _prim_of(at(Q, IM - 1 - i, 1, k)[:, None], g)[:5, 0]
                assert np.allclose(at(P, i, 0, k)[:5], 0.5 * (pa + pb), rtol=1e-14, atol=1e-14)
                for lj in (1, 2):
                    assert np.array_equal(at(Q, i, -lj, k), at(Q, IM - 1 - i, lj, k))
            elif i < p.i_end:
                if viscous:  # no-slip: zero momentum, odd-reflected ghosts
                    assert np.all(at(Q, i, 0, k)[1:4] == 0.0)
                    assert at(P, i, 0, k)[4] == pytest.approx(at(P, i, 1, k)[4], rel=1e-14)
                    for lj in (1, 2):
                        assert np.array_equal(at(Q, i, -lj, k)[1:4], -at(Q, i, lj, k)[1:4])
                else:  # slip: wall velocity along the tangent (eta_y, -eta_x)
                    ex, ey = at(o.met, i, 0, k)[3], at(o.met, i, 0, k)[4]
                    u, v = at(P, i, 0, k)[1], at(P, i, 0, k)[2]
                    assert abs(u * ex + v * ey) < 1e-12 * np.hypot(ex, ey) * max(1, np.hypot(u, v))
                    for lj in (1, 2):
                        mi, mg = at(Q, i, lj, k)[1:3], at(Q, i, -lj, k)[1:3]
                        t = np.array([ey, -ex]) / np.hypot(ex, ey)
                        nn = np.array([ex, ey]) / np.hypot(ex, ey)
                        assert np.dot(mg, t) == pytest.approx(np.dot(mi, t), abs=1e-12)
                        assert np.dot(mg, nn) == pytest.approx(-np.dot(mi, nn), abs=1e-12)
            # eta = JM-1: primitives at JM-1 and JM from Q(JM-1), Q frozen
            ptop = _prim_of(at(Q, i, JM - 1, k)[:, None], g)[:, 0]
            for lj in (0, 1):
                n = 6 if viscous else 5
                assert np.array_equal(at(P, i, JM - 1 + lj, k)[:n], ptop[:n])
            assert np.array_equal(at(Q, i, JM - 1, k), at(before, i, JM - 1, k))
    for j in range(JM):
        for i in range(IM):  # zeta copies
            for lk in range(3):
                assert np.array_equal(at(Q, i, j, -lk), at(Q, i, j, 1))
                assert np.array_equal(at(Q, i, j, KM - 1 + lk), at(Q, i, j, KM - 2))
    # nodes no rule touches keep their values (here: an edge ghost)
    assert np.array_equal(at(Q, -1, -1, 3), at(before, -1, -1, 3))


def flat_wall_case(IM=20, JM=10, KM=7, seed=5):
    """Cartesian mesh (eta along +y at the wall) with the reference's boundary
    set, inviscid, random interior states."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k, j, i = np.meshgrid(np.arange(KM, dtype=float), np.arange(JM, dtype=float),
                          np.arange(IM, dtype=float), indexing="ij")
    shape = i.shape
    prim = np.stack([rng.uniform(0.6, 1.6, shape), rng.uniform(-1, 1, shape),
                     rng.uniform(-1, 1, shape), rng.uniform(-1, 1, shape),
                     rng.uniform(1.0, 4.0, shape)])
    params = cases.make_params(IM, JM, KM, viscous=0, bc=(0, 0, 0), CFL=0.5, gamma=1.4,
                               Mach=0.5, i_begin=5, i_end=IM - 4)
    return cases.Case("flat_wall", params, i * 0.1, j * 0.08, k * 0.12, None, prim)


def test_bc_slip_wall_ghost_energy_flat_wall(oracle):
    """The slip wall's ghost energy (src/primitive_bc_kernel.cu:331-333) is
    E_g = E_1 - (m_x^2 + m_y^2 - m_gx^2 + m_gy^2) / (2 rho) with the ghost
    momentum m_g mirrored across the wall tangent t.  On a flat wall with
    eta along +y, t = (1, 0), so m_g = (m_x, -m_y) and the expression reduces
    to E_g = E_1 - m_y^2 / rho exactly; a sign-corrected kinetic-energy
    balance (-m_gx^2 - m_gy^2) would leave E_g = E_1.  The mirrored state and
    the slip node are pinned beside it."""
    case = flat_wall_case()
    o = oracle_block(oracle, case)
    p = case.params
    before = o.Q.copy()
    o.bc_primitive()
    Q = o.Q

    def at(a, i, j, k):
        return a[..., G + k, G + j, G + i]

    checked = 0
    for k in range(1, p.KM - 1):
        for i in range(p.i_begin - 1, p.i_end):
            for lj in (1, 2):
                src = at(before, i, lj, k)
                rho, mx, my, mz, E = src
                ghost = at(Q, i, -lj, k)
                assert np.array_equal(ghost[:4], np.array([rho, mx, -my, mz]))
                assert ghost[4] == pytest.approx(E - my * my / rho, rel=1e-14, abs=1e-14)
                assert abs(ghost[4] - E) > 1e-3 * abs(my * my / rho)
                checked += 1
            # the wall node takes the first interior node's state with v = 0
            rho, mx, my, mz, E = at(before, i, 1, k)
            u, v, w = mx / rho, my / rho, mz / rho
            pres = 0.4 * (E - 0.5 * rho * (u * u + v * v + w * w))
            wall = at(Q, i, 0, k)
            assert wall[0] == rho and wall[2] == 0.0
            assert wall[1] == pytest.approx(mx, rel=1e-14)
            assert wall[4] == pytest.approx(pres / 0.4 + 0.5 * rho * (u * u + w * w), rel=1e-13)
    assert checked == 2 * (p.KM - 2) * (p.i_end - p.i_begin + 1)


# ----------------------------------------------------------- time step
@pytest.mark.parametrize("viscous", [0, 1])
def test_calc_dt_closed_form(oracle, viscous):
    N, h = 8, 0.25
    k, j, i = np.meshgrid(*(np.arange(N, dtype=float),) * 3, indexing="ij")
    prm = oracle.make_params(N, N, N, viscous=viscous, bc=(1, 1, 1), CFL=0.7, Mach=0.3,
                             Re=100.0, Pr=0.72, T_inf=300.0)
    o = oracle.Oracle(prm)
    o.set_mesh(i * h, j * h, k * h, np.eye(3) * N * h)
    rho, u, v, w, p = 1.2, 0.3, -0.4, 0.5, 9.0
    o.prim[0], o.prim[1], o.prim[2], o.prim[3], o.prim[4] = rho, u, v, w, p
    T = o.g.gM2 * p / rho
    o.prim[5] = T
    C = 110.5 / 300.0
    mu = (1 + C) / (T + C) * T ** 1.5
    o.mu[...] = mu
    c = np.sqrt(1.4 * p / rho)
    lam = (abs(u) + c + abs(v) + c + abs(w) + c) / h
    if viscous:
        cdtv = max(4.0 / 3.0, 1.4 / 0.72) / 100.0
        lam += 3 * (1 / h ** 2) * 2 * mu * cdtv * np.sqrt(u * u + v * v + w * w) / c / rho
    assert o.calc_dt() == pytest.approx(0.7 / lam, rel=1e-13)


# ----------------------------------------------------------- viscous
def _cartesian_block(oracle, N, h, visc_centered):
    k, j, i = np.meshgrid(*(np.arange(N, dtype=float),) * 3, indexing="ij")
    prm = oracle.make_params(N, N, N, viscous=1, visc_centered=visc_centered, bc=(0, 0, 0),
                             Mach=0.5, Re=50.0, T_inf=300.0)
    o = oracle.Oracle(prm)
    o.set_mesh(i * h, j * h, k * h)
    s = (np.arange(N + 2 * G) - G) * h
    Z, Y, X = np.meshgrid(s, s, s, indexing="ij")
    return o, X, Y, Z


@pytest.mark.parametrize("visc_centered", [0, 1])
def test_viscous_divergence_closed_form(oracle, visc_centered):
    """u = x^2 y, v = w = 0, rho = 1, T = 1 (mu = 1) on a Cartesian mesh:
    D(mom x) = 8/3 y/Re, D(mom y) = 2x/Re - 4/3 x/Re, each derivative taken
    at the reference's transverse offset (-1 node) or centred."""
    N, h = 10, 0.1
    o, X, Y, Z = _cartesian_block(oracle, N, h, visc_centered)
    g = o.g
    o.prim[0] = 1.0
    o.prim[1] = X * X * Y
    o.prim[2] = 0.0
    o.prim[3] = 0.0
    o.prim[4] = g.p_inf
    o.prim[5] = 1.0
    dfv = o.viscous()
    Re = 50.0
    sh = 0 if visc_centered else -h
    J = h ** 3
    sl = (slice(G + 1, G + N - 1),) * 3
    d0 = dfv[0][sl] / J
    d1 = dfv[1][sl] / J
    Xs, Ys = X[sl], Y[sl]
    assert np.allclose(d0, 8.0 / 3.0 * (Ys + sh) / Re, rtol=0, atol=1e-10)
    assert np.allclose(d1, 2 * Xs / Re - 4.0 / 3.0 * (Xs + sh) / Re, rtol=0, atol=1e-10)
    assert np.allclose(dfv[2][sl], 0.0, atol=1e-14)


def test_viscous_energy_closed_form(oracle):
    """u = x^2 (1-D), T = 1: D(energy) = 8 x^2 / Re."""
    N, h = 10, 0.1
    o, X, Y, Z = _cartesian_block(oracle, N, h, 0)
    o.prim[0] = 1.0
    o.prim[1] = X * X
    o.prim[2] = o.prim[3] = 0.0
    o.prim[4] = o.g.p_inf
    o.prim[5] = 1.0
    dfv = o.viscous()
    sl = (slice(G + 1, G + N - 1),) * 3
    assert np.allclose(dfv[0][sl] / h ** 3, 8.0 / 3.0 / 50.0, atol=1e-10)
    assert np.allclose(dfv[3][sl] / h ** 3, 8 * X[sl] ** 2 / 50.0, atol=1e-10)


# ------------------------------------------------------------------ RK3
def test_rk3_stability_polynomial(oracle):
    """dq/dt = -a q through the flux-difference update: one SSP-RK3 step
    multiplies q by 1 - z + z^2/2 - z^3/6 (z = a dt)."""
    N = 6
    o = oracle.Oracle(oracle.make_params(N, N, N, bc=(0, 0, 0)))
    k, j, i = np.meshgrid(*(np.arange(N, dtype=float),) * 3, indexing="ij")
    o.set_mesh(i * 0.5, j * 0.5, k * 0.5)
    rng = np.random.default_rng(1)
    o.Q[...] = rng.uniform(0.5, 1.5, o.Q.shape)
    q0 = o.Q.copy()
    a, dt = 0.7, 0.3
    J = o.met[9]
    for stage in (1, 2, 3):
        src = o.Q if stage == 1 else o.Qp
        flux = np.zeros((15,) + o.shape)
        flux[0:5] = np.cumsum(J * a * src, axis=-1)  # E(i) - E(i-1) = J a q(i)
        import ctypes
        oracle.lib().orc_update_rk3(ctypes.byref(o.g), dt, stage, oracle.ptr(o.met), oracle.ptr(flux),
                                    None, oracle.ptr(o.Q), oracle.ptr(o.Qp))
    z = a * dt
    mask = (slice(None),) + (slice(G + 1, G + N - 1),) * 3
    assert np.allclose(o.Q[mask], q0[mask] * (1 - z + z * z / 2 - z ** 3 / 6), rtol=1e-14)


# ------------------------------------------- properties of a full step
def _total(o, Q):
    J = o.inner(o.met[9])
    return np.array([(J * o.inner(Q[c])).sum() for c in range(5)])


@pytest.mark.parametrize("builder", ["vortex", "tgv_ref", "tgv_centered"])
def test_step_conserves_on_periodic_mesh(oracle, builder):
    if builder == "vortex":
        case = cases.isentropic_vortex(12)
    else:
        case = cases.taylor_green(12, visc_centered=int(builder == "tgv_centered"))
    o = oracle_block(oracle, case)
    o.bc_primitive()
    before = _total(o, o.Q)
    o.rk3_step(1e9, 0.0)
    after = _total(o, o.Q)
    scale = np.abs(before).max()
    assert np.abs(after - before).max() < 1e-13 * scale


def test_freestream_preserved_on_cartesian_mesh(oracle):
    N = 10
    k, j, i = np.meshgrid(*(np.arange(N, dtype=float),) * 3, indexing="ij")
    case = cases.isentropic_vortex(N)
    case.x, case.y, case.z = i * 1.0, j * 1.0, k * 1.0
    case.period = np.eye(3) * N
    case.prim = np.stack([np.full(i.shape, v) for v in (1.3, 0.5, -0.2, 0.1, 2.0)])
    o = oracle_block(oracle, case)
    o.bc_primitive()
    q0 = o.Q.copy()
    o.rk3_step(1e9, 0.0)
    assert np.abs(o.Q - q0).max() < 1e-14


@pytest.mark.parametrize("viscous", [0, 1])
def test_stage_same_on_every_axis(oracle, viscous):
    """A 1-D flow along xi, eta or zeta of a Cartesian periodic mesh gives the
    same profile: every sweep, its metrics and its index ranges agree (the
    shear waves vanish, so the tangent-basis choice does not enter)."""
    N, h = 12, 0.4
    L = N * h
    k, j, i = np.meshgrid(*(np.arange(N, dtype=float),) * 3, indexing="ij")
    coords = (i * h, j * h, k * h)
    res = []
    for a in range(3):
        s = coords[a]
        rho = 1.0 + 0.2 * np.sin(2 * np.pi * s / L)
        un = 0.5 + 0.1 * np.cos(2 * np.pi * s / L)
        p = 1.0 + 0.1 * np.sin(4 * np.pi * s / L + 0.3)
        vel = [np.zeros_like(s)] * 3
        vel[a] = un
        case = cases.isentropic_vortex(N)
        case.params.viscous = viscous
        case.params.Re, case.params.T_inf, case.params.Pr = 80.0, 300.0, 0.72
        case.x, case.y, case.z = coords
        case.period = np.eye(3) * L
        case.prim = np.stack([rho, vel[0], vel[1], vel[2], p])
        o = oracle_block(oracle, case)
        o.bc_primitive()
        o.rk3_stage(1e-2, 1)
        q = o.inner(o.Qp)
        prof = np.stack([q[0], q[1 + a], q[4]])
        # bring axis a to the xi position
        res.append(np.moveaxis(prof, 3 - a, 3))
        assert np.abs(np.delete(q[1:4], a, axis=0)).max() < 1e-15
    assert np.abs(res[1] - res[0]).max() < 1e-14
    assert np.abs(res[2] - res[0]).max() < 1e-14
