"""Meshes and initial states of the workloads the solver is measured on.

Each builder returns a Case: the cbns_params, node coordinate arrays
(KM, JM, IM), the period vectors of the periodic axes and the primitive
initial state at the nodes.  Nothing here runs per step.

  isentropic_vortex  inviscid vortex advection in [0,10]^3, periodic,
                     curvilinear mesh x = xi + 0.3 sin(2 pi eta/10) sin(2 pi zeta/10)
                     and cyclic permutations; vortex strength 5 along z
                     centred at (5,5), freestream u=v=w=1, rho=p=1
  taylor_green       viscous Taylor-Green vortex on [0,2 pi]^3, periodic,
                     Re=1600, M=0.1, Pr=0.72, Sutherland viscosity; mesh
                     x = xi + 0.05 L sin(eta + phi1) sin(zeta + phi2) and
                     cyclic permutations, phases drawn from a seeded PCG64
  c_grid_channel     a small mesh for the reference's own boundary set
                     (far field on xi, wake cut / wall on eta = 0, frozen far
                     field on eta = JM-1, copies on zeta) with a perturbed
                     freestream
  shock_vortex       Mach-2 normal shock hit by a compressible columnar vortex
                     (strength 0.3, core radius 0.05), inviscid, on a
                     [0,2]x[0,1]x[0,1] mesh clustered (spacing ratio 1.5) at
                     the shock plane and gently curved; xi takes the
                     reference's far-field rule, eta and zeta are periodic;
                     vortex centre and sense of rotation from a seeded PCG64
  isotropic_turbulence  decaying compressible isotropic turbulence on
                     [0,2 pi]^3, periodic, uniform mesh: a random solenoidal
                     Fourier velocity field with E(k) ~ k^4 exp(-2 (k/k0)^2),
                     k0 = 4, seeded phases, scaled to turbulent Mach 0.6 and
                     Re_lambda = 100; uniform rho and p; viscous
"""
from dataclasses import dataclass

import numpy as np

from .block import BC_PERIODIC, BC_REF, make_params


@dataclass
class Case:
    name: str
    params: object
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    period: np.ndarray  # (3, 3) shift of coordinate c across axis a, or None
    prim: np.ndarray    # (5, KM, JM, IM): rho, u, v, w, p at the nodes


def _node_coords(n, length):
    return np.arange(n, dtype=np.float64) * (length / n)


def isentropic_vortex(N=64, *, CFL=0.5, strength=5.0, gamma=1.4):
    L = 10.0
    s = _node_coords(N, L)
    zeta, eta, xi = np.meshgrid(s, s, s, indexing="ij")  # (K, J, I)
    A, w = 0.3, 2.0 * np.pi / L
    x = xi + A * np.sin(w * eta) * np.sin(w * zeta)
    y = eta + A * np.sin(w * zeta) * np.sin(w * xi)
    z = zeta + A * np.sin(w * xi) * np.sin(w * eta)
    dx, dy = x - 5.0, y - 5.0
    r2 = dx * dx + dy * dy
    e = strength / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    u = 1.0 - e * dy
    v = 1.0 + e * dx
    wv = np.ones_like(u)
    T = 1.0 - (gamma - 1.0) * strength ** 2 / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho * T
    # p_inf = 1/(gamma M^2) = 1 for the freestream rho = p = 1
    params = make_params(N, N, N, viscous=0, bc=(BC_PERIODIC,) * 3, CFL=CFL, gamma=gamma,
                         Mach=1.0 / np.sqrt(gamma))
    period = np.eye(3) * L
    return Case("isentropic_vortex_%d" % N, params, x, y, z, period, np.stack([rho, u, v, wv, p]))


def taylor_green(N=128, *, CFL=0.5, Re=1600.0, Mach=0.1, Pr=0.72, gamma=1.4, T_inf=273.15,
                 seed=2024, perturb=0.05, visc_centered=0):
    L = 2.0 * np.pi
    s = _node_coords(N, L)
    zeta, eta, xi = np.meshgrid(s, s, s, indexing="ij")
    rng = np.random.Generator(np.random.PCG64(seed))
    ph = rng.uniform(0.0, 2.0 * np.pi, size=6)
    A = perturb * L
    x = xi + A * np.sin(eta + ph[0]) * np.sin(zeta + ph[1])
    y = eta + A * np.sin(zeta + ph[2]) * np.sin(xi + ph[3])
    z = zeta + A * np.sin(xi + ph[4]) * np.sin(eta + ph[5])
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    p0 = 1.0 / (gamma * Mach ** 2)
    p = p0 + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    rho = gamma * Mach ** 2 * p  # T = gamma M^2 p / rho = 1
    params = make_params(N, N, N, viscous=1, visc_centered=visc_centered,
                         bc=(BC_PERIODIC,) * 3, CFL=CFL, gamma=gamma, Mach=Mach, Re=Re, Pr=Pr,
                         T_inf=T_inf)
    period = np.eye(3) * L
    return Case("taylor_green_%d" % N, params, x, y, z, period, np.stack([rho, u, v, w, p]))


def c_grid_channel(IM=24, JM=16, KM=8, *, viscous=0, seed=7, Mach=0.5, CFL=0.5,
                   angle_attack=2.31):
    """Curved sheared mesh with the reference's boundary set on all axes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k, j, i = np.meshgrid(np.arange(KM, dtype=float), np.arange(JM, dtype=float),
                          np.arange(IM, dtype=float), indexing="ij")
    x = i * 0.1 + 0.02 * np.sin(0.3 * j) * np.cos(0.2 * k)
    y = j * 0.08 * (1.0 + 0.05 * np.cos(0.25 * i)) + 0.01 * i
    z = k * 0.12 + 0.01 * np.sin(0.2 * i + 0.1 * j)
    gamma = 1.4
    aoa = angle_attack * np.pi / 180.0
    shape = x.shape
    rho = 1.0 + 0.05 * rng.standard_normal(shape)
    u = np.cos(aoa) + 0.05 * rng.standard_normal(shape)
    v = np.sin(aoa) + 0.05 * rng.standard_normal(shape)
    w = 0.05 * rng.standard_normal(shape)
    p = 1.0 / (gamma * Mach ** 2) * (1.0 + 0.05 * rng.standard_normal(shape))
    params = make_params(IM, JM, KM, viscous=viscous, bc=(BC_REF,) * 3, CFL=CFL, gamma=gamma,
                         Mach=Mach, Re=5.0e3, Pr=0.71, T_inf=255.556, angle_attack=angle_attack,
                         i_begin=max(2, IM // 4), i_end=IM - max(2, IM // 4) + 1)
    return Case("c_grid_channel", params, x, y, z, None, np.stack([rho, u, v, w, p]))


def _clustered(n, length, x_s, ratio=1.5, width=0.05):
    """Node positions of n points on [0, length) whose spacing is `ratio`
    times finer at x_s than far from it: dx/ds ~ 1 - k sech^2((s - s_s)/w)
    with k = 1 - 1/ratio, integrated in closed form; s_s solves x(s_s) = x_s."""
    k = 1.0 - 1.0 / ratio
    s = np.arange(n, dtype=np.float64) / n

    def raw(t, ss):
        return t - k * width * (np.tanh((t - ss) / width) - np.tanh(-ss / width))

    def mapped(t, ss):
        return length * raw(t, ss) / raw(1.0, ss)

    lo, hi = 0.0, 1.0
    for _ in range(200):  # bisection on the monotone map s_s -> x(s_s)
        mid = 0.5 * (lo + hi)
        if mapped(mid, mid) < x_s:
            lo = mid
        else:
            hi = mid
    return mapped(s, 0.5 * (lo + hi))


def shock_vortex(IM=256, JM=128, KM=128, *, seed=0, CFL=0.5, Mach=2.0, gamma=1.4,
                 strength=0.3, core=0.05, x_shock=0.6):
    """Mach-2 stationary normal shock at x_shock with a columnar (z-aligned)
    vortex upstream, moving with the freestream into it.  Vortex:
    u_theta = s eps (r/rc) exp((1 - (r/rc)^2)/2), eps = strength (freestream
    speed 1), isentropic core T = 1 - (gamma-1) M^2 eps^2 exp(1 - (r/rc)^2)/2.
    Downstream: the Rankine-Hugoniot state of the freestream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xv = rng.uniform(0.25, 0.35)
    yv = rng.uniform(0.35, 0.65)
    sense = 1.0 if rng.uniform() < 0.5 else -1.0
    Lx, Ly, Lz = 2.0, 1.0, 1.0
    xs = _clustered(IM, Lx, x_shock)
    ys = _node_coords(JM, Ly)
    zs = _node_coords(KM, Lz)
    zeta, eta, xi = np.meshgrid(zs, ys, xs, indexing="ij")
    bump = np.sin(np.pi * xi / Lx)
    a = 0.01
    x = xi + a * np.sin(2 * np.pi * eta / Ly) * np.sin(2 * np.pi * zeta / Lz)
    y = eta + a * np.sin(2 * np.pi * zeta / Lz) * bump
    z = zeta + a * np.sin(2 * np.pi * eta / Ly) * bump
    # freestream (rho = 1, |u| = 1, p = 1/(gamma M^2)) plus the vortex
    dx, dy = x - xv, y - yv
    r2 = (dx * dx + dy * dy) / (core * core)
    g = np.exp(0.5 * (1.0 - r2))
    u = 1.0 - sense * strength * (dy / core) * g
    v = sense * strength * (dx / core) * g
    w = np.zeros_like(u)
    T = 1.0 - 0.5 * (gamma - 1.0) * Mach ** 2 * strength ** 2 * g * g
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho * T / (gamma * Mach ** 2)
    # Rankine-Hugoniot jump for the freestream Mach number
    M2 = Mach * Mach
    r_ratio = (gamma + 1.0) * M2 / ((gamma - 1.0) * M2 + 2.0)
    p_ratio = 1.0 + 2.0 * gamma / (gamma + 1.0) * (M2 - 1.0)
    down = x > x_shock
    rho = np.where(down, r_ratio, rho)
    u = np.where(down, 1.0 / r_ratio, u)
    v = np.where(down, 0.0, v)
    p = np.where(down, p_ratio / (gamma * M2), p)
    params = make_params(IM, JM, KM, viscous=0, bc=(BC_REF, BC_PERIODIC, BC_PERIODIC), CFL=CFL,
                         gamma=gamma, Mach=Mach)
    period = np.diag([0.0, Ly, Lz])
    return Case("shock_vortex_%dx%dx%d_s%d" % (IM, JM, KM, seed), params, x, y, z, period,
                np.stack([rho, u, v, w, p]))


def solenoidal_field(N, seed, k0=4.0):
    """Real velocity field (3, N, N, N) on the uniform grid x = 2 pi i / N
    (axis order k, j, i), divergence-free mode by mode, with shell energy
    E(k) ~ k^4 exp(-2 (k/k0)^2) and complex-Gaussian mode amplitudes from a
    seeded PCG64; normalised to <|u|^2> = 1."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kf = np.fft.fftfreq(N, 1.0 / N)
    kr = np.fft.rfftfreq(N, 1.0 / N)
    kz = kf[:, None, None]
    ky = kf[None, :, None]
    kx = kr[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    kk = np.sqrt(k2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(kk > 0, np.sqrt(kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2) / (4.0 * np.pi * k2)),
                       0.0)
    shape = amp.shape
    uh = [amp * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)]
    # solenoidal projection: u_hat -= k (k . u_hat) / |k|^2  (components z, y, x -> w, v, u)
    kvec = (kx, ky, kz)
    with np.errstate(divide="ignore", invalid="ignore"):
        dot = (kx * uh[0] + ky * uh[1] + kz * uh[2]) / np.where(k2 > 0, k2, 1.0)
    for c in range(3):
        uh[c] = uh[c] - kvec[c] * dot
    u = np.stack([np.fft.irfftn(uh[c], s=(N, N, N), axes=(0, 1, 2)) for c in range(3)])  # u, v, w
    del uh
    u /= np.sqrt(np.mean(np.sum(u * u, axis=0)))
    return u


def taylor_microscale(u):
    """Longitudinal Taylor microscale of a periodic field on [0, 2 pi]^3:
    lambda = u' / sqrt(<(du_a/dx_a)^2>), u' = sqrt(<|u|^2>/3)."""
    N = u.shape[-1]
    kf = np.fft.fftfreq(N, 1.0 / N)
    kr = np.fft.rfftfreq(N, 1.0 / N)
    ks = (kr[None, None, :], kf[None, :, None], kf[:, None, None])  # x, y, z
    g2 = []
    for a in range(3):
        d = np.fft.irfftn(1j * ks[a] * np.fft.rfftn(u[a]), s=u.shape[1:], axes=(0, 1, 2))
        g2.append(np.mean(d * d))
    uprime = np.sqrt(np.mean(np.sum(u * u, axis=0)) / 3.0)
    return uprime, uprime / np.sqrt(np.mean(g2))


def isotropic_turbulence(N=512, *, seed=0, Mt=0.6, Re_lambda=100.0, k0=4.0, CFL=0.5,
                         gamma=1.4, Pr=0.72, T_inf=273.15):
    """Velocities scaled to rms magnitude 1 and Mach = Mt, so the turbulent
    Mach number sqrt(<|u|^2>)/c is Mt; rho = 1, T = 1 (mu = 1); Re from
    Re_lambda = u' lambda / nu with nu = 1/Re."""
    L = 2.0 * np.pi
    u = solenoidal_field(N, seed, k0)
    uprime, lam = taylor_microscale(u)
    Re = Re_lambda / (uprime * lam)
    s = _node_coords(N, L)
    zeta, eta, xi = np.meshgrid(s, s, s, indexing="ij")
    rho = np.ones_like(xi)
    p = np.full_like(xi, 1.0 / (gamma * Mt * Mt))
    params = make_params(N, N, N, viscous=1, bc=(BC_PERIODIC,) * 3, CFL=CFL, gamma=gamma,
                         Mach=Mt, Re=Re, Pr=Pr, T_inf=T_inf)
    period = np.eye(3) * L
    return Case("isotropic_turbulence_%d_s%d" % (N, seed), params, xi, eta, zeta, period,
                np.concatenate([rho[None], u, p[None]]))


def smooth_box(IM, JM, KM, *, viscous=1, bc=(BC_PERIODIC,) * 3, seed=11, Mach=0.3, CFL=0.5,
               visc_centered=0):
    """Any size and boundary mix: [0, 2 pi)^3 in computational coordinates,
    curved by x = xi + 0.2 sin(eta + a) sin(zeta + b) and its cyclic
    permutations (periodic in every index), with a smooth seeded state of a
    few Fourier modes around a freestream (rho, p > 0 everywhere)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = 2.0 * np.pi
    zeta, eta, xi = np.meshgrid(_node_coords(KM, L), _node_coords(JM, L), _node_coords(IM, L),
                                indexing="ij")
    ph = rng.uniform(0.0, 2.0 * np.pi, size=6)
    x = xi + 0.2 * np.sin(eta + ph[0]) * np.sin(zeta + ph[1])
    y = eta + 0.2 * np.sin(zeta + ph[2]) * np.sin(xi + ph[3])
    z = zeta + 0.2 * np.sin(xi + ph[4]) * np.sin(eta + ph[5])
    gamma = 1.4

    def modes(amp):
        a = rng.uniform(-1.0, 1.0, size=(3, 3))
        return amp * (np.sin(xi + a[0, 0]) * np.cos(eta + a[0, 1]) +
                      np.cos(eta + a[1, 0]) * np.sin(zeta + a[1, 1]) +
                      np.sin(zeta + a[2, 0] + xi) * a[2, 2])

    rho = 1.0 + modes(0.1)
    u = 1.0 + modes(0.2)
    v = 0.2 + modes(0.2)
    w = -0.1 + modes(0.2)
    p = 1.0 / (gamma * Mach ** 2) * (1.0 + modes(0.05))
    n = (IM, JM, KM)
    params = make_params(IM, JM, KM, viscous=viscous, visc_centered=visc_centered, bc=tuple(bc),
                         CFL=CFL, gamma=gamma, Mach=Mach, Re=800.0, Pr=0.72, T_inf=273.15,
                         i_begin=max(1, IM // 4), i_end=IM - max(1, IM // 4))
    period = np.eye(3) * L
    name = "smooth_box_%dx%dx%d_%s%s" % (IM, JM, KM, "".join("p" if b else "r" for b in bc),
                                         "_visc" if viscous else "")
    return Case(name, params, x, y, z, period, np.stack([rho, u, v, w, p]))


def pad_modes(params):
    """np.pad mode per node axis (K, J, I): wrap on periodic axes, edge else."""
    return ["wrap" if params.bc[a] == BC_PERIODIC else "edge" for a in (2, 1, 0)]


def conservative(prim, gamma):
    """(5, ...) primitives -> (5, ...) conservative (src/conservative_kernel.cu:42-46)."""
    rho, u, v, w, p = prim
    return np.stack([rho, rho * u, rho * v, rho * w,
                     p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)])


def ghosted(arr, ng, mode="edge"):
    """Pad node arrays (..., K, J, I) with ng ghost layers on each side; mode
    is one np.pad mode or a list of three (for K, J, I)."""
    modes = [mode] * 3 if isinstance(mode, str) else list(mode)
    out = arr
    for ax, m in zip(range(arr.ndim - 3, arr.ndim), modes):
        pad = [(0, 0)] * arr.ndim
        pad[ax] = (ng, ng)
        out = np.pad(out, pad, mode=m)
    return out


def q_ref_aos(case):
    """Conservative state in the reference's host layout: AoS, 2 ghost layers."""
    q = conservative(case.prim, case.params.gamma)
    qg = ghosted(q, 2, mode=pad_modes(case.params) if case.period is not None else "edge")
    return np.ascontiguousarray(np.moveaxis(qg, 0, -1))
