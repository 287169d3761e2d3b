"""Volume-averaged diagnostics of a periodic block (host side, numpy).

kinetic_energy  (1/V) sum J rho |u|^2 / 2
enstrophy       (1/V) sum J rho |omega|^2 / 2, omega = curl u with the
                chain rule d/dx_c = sum_d (d xi_d / d x_c) d/d xi_d and 4th-order
                central index derivatives (periodic axes)
Both are computed from the block's node fields (primitives and metrics).
"""
import numpy as np

from .block import F_MET, F_PRIM, G


def _nodes(a, IM, JM, KM):
    return a[..., G:G + KM, G:G + JM, G:G + IM]


def _d4(f, axis):
    # periodic 4th-order central derivative in index space
    return (np.roll(f, 2, axis) - np.roll(f, -2, axis)) / 12.0 \
        - 2.0 / 3.0 * (np.roll(f, 1, axis) - np.roll(f, -1, axis))


def fields(block):
    IM, JM, KM = block.IM, block.JM, block.KM
    prim = _nodes(block.get_field(F_PRIM, 6), IM, JM, KM)
    met = _nodes(block.get_field(F_MET, 10), IM, JM, KM)
    return prim, met


def kinetic_energy(prim, met):
    J = met[9]
    rho, u, v, w = prim[0], prim[1], prim[2], prim[3]
    return float((J * 0.5 * rho * (u * u + v * v + w * w)).sum() / J.sum())


def enstrophy(prim, met):
    J = met[9]
    rho = prim[0]
    grad = []  # grad[q][c] = d vel_q / d x_c
    for q in (1, 2, 3):
        dxi = [_d4(prim[q], ax) for ax in (2, 1, 0)]  # d/dxi, d/deta, d/dzeta
        grad.append([sum(met[3 * d + c] * dxi[d] for d in range(3)) for c in range(3)])
    wx = grad[2][1] - grad[1][2]
    wy = grad[0][2] - grad[2][0]
    wz = grad[1][0] - grad[0][1]
    return float((J * 0.5 * rho * (wx * wx + wy * wy + wz * wz)).sum() / J.sum())
