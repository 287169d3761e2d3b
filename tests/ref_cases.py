"""The reference's own known-answer cases (test/gradient_test.cpp,
test/metrics_test.cpp), restated as data generators with their tolerances.

Each returns node arrays in (KM, JM, IM) order (i fastest, as the reference
stores them) plus the analytic answer.
"""
import numpy as np


def gradient_fun01():
    """test/gradient_test.cpp:20-84: f = i^2 + j^2 + k^2 on 50 x 100 x 150."""
    IM, JM, KM = 50, 100, 150
    k, j, i = np.meshgrid(np.arange(KM, dtype=float), np.arange(JM, dtype=float),
                          np.arange(IM, dtype=float), indexing="ij")
    return i * i + j * j + k * k, (2 * i, 2 * j, 2 * k), 1e-5


def gradient_fun02():
    """test/gradient_test.cpp:86-150: f = i j k."""
    IM, JM, KM = 50, 100, 150
    k, j, i = np.meshgrid(np.arange(KM, dtype=float), np.arange(JM, dtype=float),
                          np.arange(IM, dtype=float), indexing="ij")
    return i * j * k, (j * k, i * k, i * j), 1e-5


def metrics_rectilinear():
    """test/metrics_test.cpp:77-225: tanh-stretched 10^3 box, MSE <= 2e-2."""
    n, L, s = 10, 5.0, 1.2
    k, j, i = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float),
                          np.arange(n, dtype=float), indexing="ij")
    x = 0.5 * L * (np.tanh(s * (-1 + 2.0 * i / (n - 1))) + 1)
    y = 0.5 * L * (np.tanh(s * (-1 + 2.0 * j / (n - 1))) + 1)
    z = 0.5 * L * (np.tanh(s * (-1.0 + 2.0 * k / (n - 1))) + 1)
    ux = np.tanh(s * (-1 + 2.0 * i / (n - 1)))
    uy = np.tanh(s * (-1 + 2.0 * j / (n - 1)))
    uz = np.tanh(s * (-1.0 + 2.0 * k / (n - 1)))
    x_xi = L * s * (1 - ux * ux) / (n - 1)
    y_eta = L * s * (1 - uy * uy) / (n - 1)
    z_zeta = L * s * (1 - uz * uz) / (n - 1)
    zero = np.zeros_like(x)
    exp = [1.0 / x_xi, zero, zero, zero, 1.0 / y_eta, zero, zero, zero, 1.0 / z_zeta,
           x_xi * y_eta * z_zeta]
    return (x, y, z), exp, 2.0e-2


def _rotation_matrix():
    a = b = g = np.pi / 4.0
    ca, sa, cb, sb, cg, sg = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(g), np.sin(g)
    return np.array([[cb * cg, sa * sb * cg - sg * ca, sa * sg + sb * ca * cg],
                     [sg * cb, sa * sb * sg + ca * cg, -sa * cg + sb * sg * ca],
                     [-sb, sa * cb, ca * cb]])


def metrics_rotation():
    """test/metrics_test.cpp:227-409: rotated uniform 10^3 box, MSE <= 1e-8."""
    n = 10
    R = _rotation_matrix()
    k, j, i = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float),
                          np.arange(n, dtype=float), indexing="ij")
    v = np.stack([i / (n - 1.0), j / (n - 1.0), k / (n - 1.0)])
    X = np.einsum("ab,b...->a...", R, v)
    JA = R @ np.diag([1.0 / (n - 1)] * 3)
    JB = np.linalg.inv(JA)
    ones = np.ones((n, n, n))
    exp = [JB[r, c] * ones for r in range(3) for c in range(3)]
    exp.append(np.linalg.det(R) / (n - 1) ** 3 * ones)
    return (X[0], X[1], X[2]), exp, 1.0e-8


def metrics_conical_shell():
    """test/metrics_test.cpp:411-601: 10 x 20 x 5 conical shell, MSE <= 1e-3."""
    IM, JM, KM = 10, 20, 5
    Rx, Ry, Rz = 3.0, 6.0, 5.0
    th, ph = 5.0 * np.pi / 12.0, 1.0 * np.pi / 12.0
    k, j, i = np.meshgrid(np.arange(KM, dtype=float), np.arange(JM, dtype=float),
                          np.arange(IM, dtype=float), indexing="ij")
    xi, eta, zeta = i / (IM - 1), j / (JM - 1), k / (KM - 1)
    x = (Rx - (Rx - 1.0) * xi) * np.cos(th * (2 * eta - 1)) * np.sin(ph * (2 * zeta + 1))
    y = (Ry - (Ry - 1.0) * xi) * np.sin(th * (2 * eta - 1)) * np.sin(ph * (2 * zeta + 1))
    z = (Rz - (Rz - 1.0) * xi) * np.cos(ph * (2 * zeta + 1))
    Jm = np.zeros((3, 3) + x.shape)
    Jm[0, 0] = (1 - Rx) * np.sin(ph * (2 * zeta + 1)) * np.cos(th * (2 * eta - 1))
    Jm[0, 1] = -2 * th * (Rx - xi * (Rx - 1)) * np.sin(ph * (2 * zeta + 1)) * np.sin(th * (2 * eta - 1))
    Jm[0, 2] = 2 * ph * (Rx - xi * (Rx - 1)) * np.cos(ph * (2 * zeta + 1)) * np.cos(th * (2 * eta - 1))
    Jm[1, 0] = (1 - Ry) * np.sin(ph * (2 * zeta + 1)) * np.sin(th * (2 * eta - 1))
    Jm[1, 1] = 2 * th * (Ry - xi * (Ry - 1)) * np.sin(ph * (2 * zeta + 1)) * np.cos(th * (2 * eta - 1))
    Jm[1, 2] = 2 * ph * (Ry - xi * (Ry - 1)) * np.sin(th * (2 * eta - 1)) * np.cos(ph * (2 * zeta + 1))
    Jm[2, 0] = (1 - Rz) * np.cos(ph * (2 * zeta + 1))
    Jm[2, 1] = 0.0
    Jm[2, 2] = -2 * ph * (Rz - xi * (Rz - 1)) * np.sin(ph * (2 * zeta + 1))
    scale = np.array([1.0 / (IM - 1), 1.0 / (JM - 1), 1.0 / (KM - 1)])
    JA = np.moveaxis(Jm, (0, 1), (-2, -1)) * scale[None, :]
    JB = np.linalg.inv(JA)
    det = np.linalg.det(np.moveaxis(Jm, (0, 1), (-2, -1))) * scale.prod()
    exp = [JB[..., r, c] for r in range(3) for c in range(3)] + [det]
    return (x, y, z), exp, 1.0e-3


def metric_mse(got, exp):
    """Per-component mean squared error, as the reference's gtests compute it."""
    return [float(np.mean((g - e) ** 2)) for g, e in zip(got, exp)]
