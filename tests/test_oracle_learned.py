"""Pins of the learned-indicator oracle (oracle/orc_learned.c) and of the
linear-weight face flux (oracle/cbns_oracle.c orc_face_flux_linear):

  bf16_round        torch's float32 -> bfloat16 conversion (round to nearest even)
  weno5_weights     sum_k omega_k f_k reproduces orc_weno5 (pinned on its own);
                    constant and linear stencils give the linear weights exactly;
                    a step leaves all weight on the smooth sub-stencil
  face_label        0 on uniform and on linearly varying states; O(1) at a jump
  face_features     zero on uniform states; closed form on linear states
                    (s - 5/2) delta / sqrt(5 delta^2 + 1e-6) with delta = L dq from
                    orc_eigenvectors_roe; invariant under a constant shift of q
  mlp_loss_grad     central finite differences of the loss (double mode)
  adam              first step from zero moments moves by lr * g / (|g| + eps)
  face_flux_linear  equals orc_face_flux where the WENO weights are linear
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def state(rho, u, v, w, p, gamma=1.4):
    return np.array([rho, rho * u, rho * v, rho * w, p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def prim(q, gamma=1.4):
    r = q[0]
    u, v, w = q[1] / r, q[2] / r, q[3] / r
    return np.array([r, u, v, w, (gamma - 1) * (q[4] - 0.5 * r * (u * u + v * v + w * w))])


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        # exact ties: low 16 bits 0x8000 with even and odd upper parts
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 3.0, 0.0])])
    x = x.astype(np.float32)
    import torch  # CPU conversion only
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(po.bf16_round(x), ref)


def test_weno5_weights_reproduce_weno5_and_limits():
    rng = np.random.default_rng(1)
    for _ in range(200):
        s = rng.standard_normal(5) * 10.0 ** rng.integers(-4, 2)
        a, b, c, d, e = s
        om = po.weno5_weights(*s)
        f = np.array([(2 * a - 7 * b + 11 * c) / 6, (-b + 5 * c + 2 * d) / 6, (2 * c + 5 * d - e) / 6])
        assert om.sum() == pytest.approx(1.0, abs=1e-15)
        assert om @ f == pytest.approx(po.weno5(*s), rel=1e-12, abs=1e-14)
    for sten in ([3.0] * 5, [0.0, 1.0, 2.0, 3.0, 4.0], [4.0, 1.0, 0.0, 1.0, 4.0]):
        assert np.allclose(po.weno5_weights(*sten), [0.1, 0.6, 0.3], rtol=0, atol=1e-15)
    om = po.weno5_weights(0.0, 0.0, 0.0, 1.0, 1.0)  # jump between c and d
    assert om[0] > 0.99


def _face(grid, qs, n=(1.0, 0.0, 0.0)):
    q = np.array(qs)
    return q, prim(q[2]), prim(q[3]), n


@pytest.mark.parametrize("n", [(1.0, 0.0, 0.0), (0.3, -0.8, 0.5), (0.0, 0.0, 2.0)])
def test_label_zero_on_uniform_and_linear_states(n):
    g = po.grid_for(Mach=0.5)
    q0 = state(1.1, 0.3, -0.2, 0.1, 2.5)
    dq = state(1.0, 0.2, 0.1, -0.1, 0.05) * 1e-2
    for qs in ([q0] * 6, [q0 + s * dq for s in range(6)]):
        q, pl, pr, _ = _face(g, qs)
        assert po.face_label(g, q, pl, pr, n) < 1e-12


def test_label_large_at_contact_jump():
    g = po.grid_for(Mach=0.5)
    lo, hi = state(1.0, 0.5, 0.0, 0.0, 1.0), state(4.0, 0.5, 0.0, 0.0, 1.0)
    q, pl, pr, n = _face(g, [lo, lo, lo, hi, hi, hi])
    assert po.face_label(g, q, pl, pr, n) > 0.5


@pytest.mark.parametrize("n", [(1.0, 0.0, 0.0), (0.3, -0.8, 0.5)])
def test_features_closed_form_on_linear_states(n):
    g = po.grid_for(Mach=0.5)
    q0 = state(1.1, 0.3, -0.2, 0.1, 2.5)
    z = po.face_features(g, *_face(g, [q0] * 6)[:3], n)
    assert np.array_equal(z, np.zeros(32, dtype=np.float32))
    for scale in (1e-4, 1e-2):
        dq = state(1.0, 0.2, 0.1, -0.1, 0.05) * scale
        q, pl, pr, _ = _face(g, [q0 + s * dq for s in range(6)])
        f = po.face_features(g, q, pl, pr, n)
        _, L = po.eigenvectors_roe(g, pl, pr, n)
        delta = L @ dq
        expect = np.array([(s - 2.5) * dc / np.sqrt(5 * dc * dc + 1e-6) for dc in delta for s in range(6)])
        assert np.abs(f[:30] - expect).max() < 1e-6
        assert f[30] == 0 and f[31] == 0
        # a constant shift of the stencil (with the same node primitives) changes nothing
        shifted = po.face_features(g, q + 0.25 * q0, pl, pr, n)
        assert np.abs(shifted - f).max() < 1e-6


def _mlp_case(seed, n=300):
    rng = np.random.default_rng(seed)
    P = (rng.standard_normal(po.MLP_NPARAM) * 0.3).astype(np.float32)
    for r in range(32):
        P[r * 32 + 30:r * 32 + 32] = 0
    X = rng.standard_normal((n, 32)).astype(np.float32)
    X[:, 30:] = 0
    y = (rng.random(n) < 0.3).astype(np.float32)
    return P, X, y


def test_mlp_gradient_matches_finite_differences():
    P, X, y = _mlp_case(3)
    loss, grad = po.mlp_loss_grad(P, X, y, w_pos=4.0, bf16=False)
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.integers(0, 30, 4) + 32 * rng.integers(0, 32, 4),  # W1
                          po.MLP_B1 + rng.integers(0, 32, 3), po.MLP_W2 + rng.integers(0, 1024, 4),
                          po.MLP_B2 + rng.integers(0, 32, 3), po.MLP_W3 + rng.integers(0, 32, 3),
                          [po.MLP_B3]])
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += np.float32(1e-5)
        Pm[i] -= np.float32(1e-5)
        fd = (po.mlp_loss_grad(Pp, X, y, 4.0, False, False) -
              po.mlp_loss_grad(Pm, X, y, 4.0, False, False)) / (float(Pp[i]) - float(Pm[i]))
        assert grad[i] == pytest.approx(fd, rel=2e-5, abs=1e-8), i
    # padding input columns of W1 get no gradient
    assert not grad[[r * 32 + c for r in range(32) for c in (30, 31)]].any()


def test_mlp_bf16_mode_close_to_double_mode():
    P, X, y = _mlp_case(5)
    a = po.mlp_forward(P, X, bf16=False)
    b = po.mlp_forward(P, X, bf16=True)
    assert np.abs(a - b).max() < 0.05 * (1 + np.abs(a).max())
    assert not np.array_equal(a, b)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(6)
    P = rng.standard_normal(50).astype(np.float32)
    g = rng.standard_normal(50)
    m, v = np.zeros(50), np.zeros(50)
    P0 = P.copy()
    po.adam(P, g, m, v, lr=1e-2, eps=1e-8, t=1)
    assert np.allclose(P, P0 - 1e-2 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-6)


@pytest.mark.parametrize("n", [(1.0, 0.0, 0.0), (0.3, -0.8, 0.5)])
def test_linear_flux_equals_weno_where_weights_are_linear(n):
    g = po.grid_for(Mach=0.5)
    q0 = state(1.1, 0.3, -0.2, 0.1, 2.5)
    dq = state(1.0, 0.2, 0.1, -0.1, 0.05) * 1e-2
    q = np.array([q0 + s * dq for s in range(6)])
    a = po.face_flux_linear(g, q, n)
    b = po.face_flux(g, q, prim(q[2]), prim(q[3]), n)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-14)
    # the coefficients are sum_k d_k (candidate stencils of src/weno.cu:39-41)
    rng = np.random.default_rng(7)
    q = np.array([q0 + rng.standard_normal(5) * 1e-2 for _ in range(6)])
    qL = (2 * q[0] - 13 * q[1] + 47 * q[2] + 27 * q[3] - 3 * q[4]) / 60
    cand = (0.1 * (2 * q[0] - 7 * q[1] + 11 * q[2]) + 0.6 * (-q[1] + 5 * q[2] + 2 * q[3]) +
            0.3 * (2 * q[2] + 5 * q[3] - q[4])) / 6
    assert np.allclose(qL, cand, rtol=1e-14)
