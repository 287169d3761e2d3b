"""Shared setup for tests: build the same block on the oracle and on the GPU."""
import numpy as np

from cbns3d_cuda_amd import cases

G = 4


def oracle_params(po, case):
    p = case.params
    return po.make_params(p.IM, p.JM, p.KM, viscous=p.viscous, visc_centered=p.visc_centered,
                          bc=tuple(p.bc), CFL=p.CFL, gamma=p.gamma, Mach=p.Mach, Re=p.Re, Pr=p.Pr,
                          Pr_t=p.Pr_t, T_inf=p.T_inf, angle_attack=p.angle_attack,
                          i_begin=p.i_begin, i_end=p.i_end)


def ghosted_q(case):
    """Full ghosted SoA conservative state: nodes from the case, ghosts by
    wrap (periodic) or edge replication (so every ghost holds a valid state)."""
    q = cases.conservative(case.prim, case.params.gamma)
    mode = cases.pad_modes(case.params) if case.period is not None else "edge"
    return np.ascontiguousarray(cases.ghosted(q, G, mode=mode))


def oracle_block(po, case):
    o = po.Oracle(oracle_params(po, case))
    o.set_mesh(case.x, case.y, case.z, case.period)
    o.Q[...] = ghosted_q(case)
    o.Qp[...] = o.Q
    # primitives start defined everywhere (nodes no boundary rule writes keep them)
    q = o.Q
    rho = q[0]
    u, v, w = q[1] / rho, q[2] / rho, q[3] / rho
    p = (o.g.gamma - 1) * (q[4] - 0.5 * rho * (u * u + v * v + w * w))
    o.prim[...] = np.stack([rho, u, v, w, p, o.g.gM2 * p / rho])
    return o


def rel_err(a, b):
    """Per-component max |a-b| / max |b| over the leading component axis."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for c in range(a.shape[0]):
        den = max(np.abs(b[c]).max(), 1e-300)
        out.append(float(np.abs(a[c] - b[c]).max() / den))
    return max(out)


def rel_err_state(a, b):
    """max |a-b| per conservative component over the scale of its quantity:
    rho and E by their own max, the three momentum components by the largest
    momentum magnitude (a component that is ~0 everywhere, like w of the
    Taylor-Green vortex, is measured against the momentum scale)."""
    a = np.asarray(a)
    b = np.asarray(b)
    mom = max(np.abs(b[1:4]).max(), 1e-300)
    scale = [max(np.abs(b[0]).max(), 1e-300), mom, mom, mom, max(np.abs(b[4]).max(), 1e-300)]
    return max(float(np.abs(a[c] - b[c]).max() / scale[c]) for c in range(5))


def update_mask(case):
    """Boolean mask (ghosted shape) of the RK-updated nodes."""
    p = case.params
    shape = (p.KM + 2 * G, p.JM + 2 * G, p.IM + 2 * G)
    m = np.zeros(shape, dtype=bool)
    sl = []
    for a, n in ((2, p.KM), (1, p.JM), (0, p.IM)):
        per = p.bc[a] == 1
        lo, hi = (0, n - 1) if per else (1, n - 2)
        sl.append(slice(G + lo, G + hi + 1))
    m[tuple(sl)] = True
    return m
