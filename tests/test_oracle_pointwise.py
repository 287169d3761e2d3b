"""Pins of the oracle's per-point functions (CPU).

Each check comes from mathematics the reference's method fixes, not from
re-typing the oracle's own formula: known answers of WENO5-JS, the Roe
property of the Roe-averaged eigen-system, textbook Euler fluxes, the
Rusanov flux for axis-aligned faces, and exact reconstruction of linear data.
"""
import numpy as np
import pytest

GAMMA = 1.4


def euler_flux_n(q, n, gamma=GAMMA):
    """Textbook Euler flux F(q).n for conservative q and (any) vector n."""
    rho, mx, my, mz, E = q
    u, v, w = mx / rho, my / rho, mz / rho
    p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v + w * w))
    qn = u * n[0] + v * n[1] + w * n[2]
    return np.array([rho * qn, mx * qn + p * n[0], my * qn + p * n[1], mz * qn + p * n[2],
                     (E + p) * qn])


def cons(prim, gamma=GAMMA):
    rho, u, v, w, p = prim
    return np.array([rho, rho * u, rho * v, rho * w, p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def prim_of(q, gamma=GAMMA):
    rho, mx, my, mz, E = q
    u, v, w = mx / rho, my / rho, mz / rho
    return np.array([rho, u, v, w, (gamma - 1) * (E - 0.5 * rho * (u * u + v * v + w * w))])


def rand_prim(rng):
    return np.array([rng.uniform(0.5, 2.0), rng.uniform(-1, 1), rng.uniform(-1, 1),
                     rng.uniform(-1, 1), rng.uniform(0.5, 3.0)])


# ---------------------------------------------------------------- weno5
def test_weno5_known_answers(oracle):
    # symmetric stencil: all three candidates are 1/6 and IS1 = IS2 = IS3
    assert oracle.weno5(4, 1, 0, 1, 4) == pytest.approx(1.0 / 6.0, rel=4e-16, abs=0)
    # nonlinear weights; value from an mpmath evaluation of WENO5-JS
    assert oracle.weno5(-8, -1, 0, 1, 8) == pytest.approx(0.49954946749288827, rel=5e-16, abs=0)


def test_weno5_constant_and_linear(oracle):
    rng = np.random.default_rng(1)
    for _ in range(50):
        c = rng.uniform(-5, 5)
        assert oracle.weno5(c, c, c, c, c) == pytest.approx(c, rel=1e-15, abs=1e-15)
        a, b = rng.uniform(-3, 3, size=2)
        f = a + b * np.arange(5.0)
        # linear data: equal smoothness indicators -> linear weights -> the
        # 5th-order upwind value (2a-13b+47c+27d-3e)/60, exact at s = 2.5
        lin = (2 * f[0] - 13 * f[1] + 47 * f[2] + 27 * f[3] - 3 * f[4]) / 60.0
        assert lin == pytest.approx(a + 2.5 * b, rel=1e-13, abs=1e-13)
        assert oracle.weno5(*f) == pytest.approx(a + 2.5 * b, rel=1e-13, abs=1e-13)


def test_weno5_picks_smooth_substencil(oracle):
    # jump between c and d: only S1 = (a,b,c) is smooth -> candidate 1
    assert abs(oracle.weno5(0, 0, 0, 1, 1)) < 1e-9
    # jump between b and c: only S3 = (c,d,e) is smooth -> candidate 3 = 1
    assert oracle.weno5(0, 0, 1, 1, 1) == pytest.approx(1.0, abs=1e-9)
    # reversed roles: a jump between a and b leaves S2 and S3 smooth
    assert oracle.weno5(5, 0, 0, 0, 0) == pytest.approx(0.0, abs=1e-9)


def test_weno5_translation_invariant(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        f = rng.uniform(-1, 1, 5)
        s = rng.uniform(-10, 10)
        assert oracle.weno5(*(f + s)) - s == pytest.approx(oracle.weno5(*f), abs=5e-14)


def test_weno5_rounding_vs_mpmath(oracle):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 50
    rng = np.random.default_rng(3)

    def ref(a, b, c, d, e):
        a, b, c, d, e = (mpmath.mpf(float(t)) for t in (a, b, c, d, e))
        eps = mpmath.mpf("1e-6")
        IS1 = mpmath.mpf(13) / 12 * (a - 2 * b + c) ** 2 + (a - 4 * b + 3 * c) ** 2 / 4 + eps
        IS2 = mpmath.mpf(13) / 12 * (b - 2 * c + d) ** 2 + (b - d) ** 2 / 4 + eps
        IS3 = mpmath.mpf(13) / 12 * (c - 2 * d + e) ** 2 + (3 * c - 4 * d + e) ** 2 / 4 + eps
        w = [mpmath.mpf("0.1") / IS1 ** 2, mpmath.mpf("0.6") / IS2 ** 2, mpmath.mpf("0.3") / IS3 ** 2]
        f = [(2 * a - 7 * b + 11 * c) / 6, (-b + 5 * c + 2 * d) / 6, (2 * c + 5 * d - e) / 6]
        return float(sum(wi * fi for wi, fi in zip(w, f)) / sum(w))

    for _ in range(200):
        f = rng.uniform(-1, 1, 5) * 10 ** rng.uniform(-3, 1)
        assert oracle.weno5(*f) == pytest.approx(ref(*f), rel=1e-13, abs=1e-15)


# ---------------------------------------------------- eigenvectors (Roe)
NORMALS = [(0.3, 0.5, -0.2), (1.0, 0.0, 0.0), (-2.0, 0.0, 0.0), (0.0, 1.0, 0.0),
           (0.0, 0.0, 1.0), (0.1, -0.7, 0.7), (1e-3, 2e-3, -5e-4)]


@pytest.mark.parametrize("n", NORMALS)
def test_roe_left_right_inverse(oracle, n):
    g = oracle.grid_for()
    rng = np.random.default_rng(4)
    for _ in range(20):
        R, L = oracle.eigenvectors_roe(g, rand_prim(rng), rand_prim(rng), n)
        assert np.abs(L @ R - np.eye(5)).max() < 1e-13


@pytest.mark.parametrize("n", NORMALS)
def test_roe_property(oracle, n):
    """R diag(lambda) L (qR - qL) = F(qR).n^ - F(qL).n^ (Roe's U-property)."""
    g = oracle.grid_for()
    rng = np.random.default_rng(5)
    nh = np.array(n) / np.linalg.norm(n)
    for _ in range(20):
        pl, pr = rand_prim(rng), rand_prim(rng)
        R, L = oracle.eigenvectors_roe(g, pl, pr, n)
        uvw = R[1:4, 1]
        c = np.dot(R[1:4, 2] - R[1:4, 0], nh) / 2.0
        qn = np.dot(uvw, nh)
        A = R @ np.diag([qn - c, qn, qn + c, qn, qn]) @ L
        dq = cons(pr) - cons(pl)
        dF = euler_flux_n(cons(pr), nh) - euler_flux_n(cons(pl), nh)
        assert np.abs(A @ dq - dF).max() < 1e-12 * max(1.0, np.abs(dF).max())
        # Roe velocity: sqrt(rho)-weighted average
        wl, wr = np.sqrt(pl[0]), np.sqrt(pr[0])
        assert np.allclose(uvw, (wl * pl[1:4] + wr * pr[1:4]) / (wl + wr), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("n", NORMALS[:3])
def test_roe_eigenvectors_of_flux_jacobian(oracle, n):
    """Columns of R are eigenvectors of dF.n^/dq at the Roe state."""
    g = oracle.grid_for()
    p = np.array([1.3, 0.4, -0.3, 0.2, 1.1])
    R, L = oracle.eigenvectors_roe(g, p, p, n)
    nh = np.array(n) / np.linalg.norm(n)
    q0 = cons(p)
    h = 1e-6
    A = np.zeros((5, 5))
    for k in range(5):
        e = np.zeros(5)
        e[k] = h
        A[:, k] = (euler_flux_n(q0 + e, nh) - euler_flux_n(q0 - e, nh)) / (2 * h)
    c = np.sqrt(GAMMA * p[4] / p[0])
    qn = np.dot(p[1:4], nh)
    lam = [qn - c, qn, qn + c, qn, qn]
    for k in range(5):
        assert np.abs(A @ R[:, k] - lam[k] * R[:, k]).max() < 1e-7


# ---------------------------------------------------------- rec_weno5
def test_rec_weno5_constant_and_linear(oracle):
    g = oracle.grid_for()
    rng = np.random.default_rng(6)
    for _ in range(20):
        R, L = oracle.eigenvectors_roe(g, rand_prim(rng), rand_prim(rng), rng.uniform(-1, 1, 3))
        q = cons(rand_prim(rng))
        out = oracle.rec_weno5(g, np.tile(q, 5), R, L)
        assert np.allclose(out, prim_of(q), rtol=1e-13, atol=1e-13)
        dq = 0.01 * rng.uniform(-1, 1, 5)
        sten = np.array([q + s * dq for s in range(5)])
        out = oracle.rec_weno5(g, sten, R, L)
        assert np.allclose(out, prim_of(q + 2.5 * dq), rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------- lf_flux
def test_lf_flux_equal_states_is_euler_flux(oracle):
    g = oracle.grid_for()
    rng = np.random.default_rng(7)
    for _ in range(20):
        p = rand_prim(rng)
        n = rng.uniform(-1, 1, 3)
        f = oracle.lf_flux(g, p, p, n)
        assert np.allclose(f, euler_flux_n(cons(p), n), rtol=1e-13, atol=1e-13)


def test_lf_flux_axis_aligned_is_rusanov(oracle):
    g = oracle.grid_for()
    rng = np.random.default_rng(8)
    for _ in range(20):
        pl, pr = rand_prim(rng), rand_prim(rng)
        nx = rng.uniform(0.1, 2.0)
        n = (nx, 0.0, 0.0)
        cl, cr = np.sqrt(GAMMA * pl[4] / pl[0]), np.sqrt(GAMMA * pr[4] / pr[0])
        a = max(abs(pl[1]) + cl, abs(pr[1]) + cr) * nx
        ref = 0.5 * (euler_flux_n(cons(pl), n) + euler_flux_n(cons(pr), n)) - 0.5 * a * (cons(pr) - cons(pl))
        assert np.allclose(oracle.lf_flux(g, pr, pl, n), ref, rtol=1e-13, atol=1e-13)


def _central(pl, pr, n):
    return 0.5 * (euler_flux_n(cons(pl), n) + euler_flux_n(cons(pr), n))


def test_lf_flux_dissipation_speed_cancels_on_a_diagonal_normal(oracle):
    """src/lf_flux.cu:37-52 takes the dissipation speed as
    |sum_d max_LR(|u_d| + c) n_d|, not the textbook max(|q.n| + c|n|).  With
    |u| = |v| on each side the per-axis maxima are equal, so for n = (a, -a, 0)
    the speed cancels to exactly 0 and the flux is the central one; the
    textbook Rusanov speed would be positive there."""
    g = oracle.grid_for()
    rng = np.random.default_rng(81)
    for _ in range(20):
        pl, pr = rand_prim(rng), rand_prim(rng)
        pl[2] = pl[1] * rng.choice((-1.0, 1.0))
        pr[2] = pr[1] * rng.choice((-1.0, 1.0))
        a = rng.uniform(0.2, 2.0)
        n = (a, -a, 0.0)
        f = oracle.lf_flux(g, pr, pl, n)
        assert np.allclose(f, _central(pl, pr, n), rtol=1e-13, atol=1e-13)
        # the textbook speed differs from zero by O(1): the check can tell them apart
        cl, cr = np.sqrt(GAMMA * pl[4] / pl[0]), np.sqrt(GAMMA * pr[4] / pr[0])
        nn = np.hypot(a, a)
        rus = max(abs(pl[1] * a - pl[2] * a) + cl * nn, abs(pr[1] * a - pr[2] * a) + cr * nn)
        jump = np.abs(cons(pr) - cons(pl)).max()
        assert 0.5 * rus * jump > 1e-3


def test_lf_flux_dissipation_speed_at_rest_sums_the_normal(oracle):
    """Both states at rest with the same sound speed c: the per-axis maxima
    are all c, so the speed is c |n_x + n_y + n_z| (the textbook speed would
    be c |n|).  Normals with mixed signs pin the sum and its outer |.|."""
    g = oracle.grid_for()
    rng = np.random.default_rng(82)
    for _ in range(20):
        rl = rng.uniform(0.5, 2.0)
        rr = rng.uniform(0.5, 2.0)
        pres_l = rng.uniform(0.5, 3.0)
        pl = np.array([rl, 0.0, 0.0, 0.0, pres_l])
        pr = np.array([rr, 0.0, 0.0, 0.0, pres_l * rr / rl])  # same c^2 = gamma p / rho
        c = np.sqrt(GAMMA * pres_l / rl)
        n = rng.uniform(-1, 1, 3)
        speed = c * abs(n.sum())
        ref = _central(pl, pr, n) - 0.5 * speed * (cons(pr) - cons(pl))
        assert np.allclose(oracle.lf_flux(g, pr, pl, n), ref, rtol=1e-12, atol=1e-12)


def test_lf_flux_negative_axis_normal_dissipates(oracle):
    """n = (-a, 0, 0): the speed is +a max(|u| + c) (the outer |.|); without it
    the dissipation would change sign and the flux would anti-diffuse."""
    g = oracle.grid_for()
    rng = np.random.default_rng(83)
    for _ in range(20):
        pl, pr = rand_prim(rng), rand_prim(rng)
        a = rng.uniform(0.1, 2.0)
        n = (-a, 0.0, 0.0)
        cl, cr = np.sqrt(GAMMA * pl[4] / pl[0]), np.sqrt(GAMMA * pr[4] / pr[0])
        speed = a * max(abs(pl[1]) + cl, abs(pr[1]) + cr)
        ref = _central(pl, pr, n) - 0.5 * speed * (cons(pr) - cons(pl))
        assert np.allclose(oracle.lf_flux(g, pr, pl, n), ref, rtol=1e-13, atol=1e-13)


# ----------------------------------------------------------- face flux
def test_face_flux_uniform_and_linear(oracle):
    g = oracle.grid_for()
    rng = np.random.default_rng(9)
    for _ in range(20):
        p = rand_prim(rng)
        n = rng.uniform(-1, 1, 3)
        q = cons(p)
        f = oracle.face_flux(g, np.tile(q, 6), p, p, n)
        assert np.allclose(f, euler_flux_n(q, n), rtol=1e-13, atol=1e-13)
        # linear data: both sides reconstruct q at the face exactly, no dissipation
        dq = 0.02 * rng.uniform(-1, 1, 5)
        sten = np.array([q + s * dq for s in range(6)])  # nodes m-2 .. m+3
        f = oracle.face_flux(g, sten, prim_of(sten[2]), prim_of(sten[3]), n)
        assert np.allclose(f, euler_flux_n(q + 2.5 * dq, n), rtol=1e-12, atol=1e-12)


def test_face_flux_upwinds_a_step(oracle):
    """A contact at rest between m and m+1: L state from the left, R from the right."""
    g = oracle.grid_for()
    ql = cons(np.array([1.0, 0.0, 0.0, 0.0, 1.0]))
    qr = cons(np.array([0.5, 0.0, 0.0, 0.0, 1.0]))
    sten = np.array([ql, ql, ql, qr, qr, qr])
    f = oracle.face_flux(g, sten, prim_of(ql), prim_of(qr), (1.0, 0.0, 0.0))
    cl, cr = np.sqrt(GAMMA), np.sqrt(GAMMA / 0.5)
    a = max(cl, cr)
    ref = 0.5 * (euler_flux_n(ql, (1, 0, 0)) + euler_flux_n(qr, (1, 0, 0))) - 0.5 * a * (qr - ql)
    assert np.allclose(f, ref, rtol=1e-8, atol=1e-8)


def test_face_operation_count():
    """The reference algorithm's floating-point work per face, counted on the
    oracle with an operation-counting scalar (oracle/opcount.cpp); bench.py
    divides this count by the measured face-kernel time."""
    import json
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    subprocess.run(["make", "-s", "-C", here, "opcount"], check=True)
    out = subprocess.run([os.path.join(here, "opcount")], capture_output=True, text=True, check=True)
    rows = [json.loads(l) for l in out.stdout.splitlines()]
    for r in rows:
        assert (r["add_sub"], r["mul"], r["div"], r["sqrt"]) == (673, 795, 102, 10)
        assert r["flops"] == 1580
