"""The workload builders (cbns3d_cuda_amd/cases.py): closed-form properties of
the initial states that the GPU runs start from."""
import numpy as np
import pytest

from cbns3d_cuda_amd import cases


def spectral_div(u):
    N = u.shape[-1]
    kf = np.fft.fftfreq(N, 1.0 / N)
    kr = np.fft.rfftfreq(N, 1.0 / N)
    ks = (kr[None, None, :], kf[None, :, None], kf[:, None, None])
    return sum(np.fft.irfftn(1j * ks[a] * np.fft.rfftn(u[a]), s=u.shape[1:], axes=(0, 1, 2))
               for a in range(3))


def test_isotropic_turbulence_initial_state():
    c = cases.isotropic_turbulence(48, seed=3)
    p = c.params
    rho, u, pr = c.prim[0], c.prim[1:4], c.prim[4]
    assert np.abs(spectral_div(u)).max() < 1e-12           # solenoidal
    assert np.mean(np.sum(u * u, axis=0)) == pytest.approx(1.0, rel=1e-12)
    c_sound = np.sqrt(p.gamma * pr / rho)
    assert np.sqrt(np.mean(np.sum(u * u, axis=0))) / c_sound.mean() == pytest.approx(0.6, rel=1e-12)
    up, lam = cases.taylor_microscale(u)
    assert up * lam * p.Re == pytest.approx(100.0, rel=1e-12)  # Re_lambda, nu = 1/Re
    # shell spectrum follows k^4 exp(-2 (k/4)^2) up to the sampling noise of
    # the random mode amplitudes
    N = 48
    uh = [np.fft.fftn(u[a]) / N ** 3 for a in range(3)]
    kf = np.fft.fftfreq(N, 1.0 / N)
    kk = np.sqrt(kf[:, None, None] ** 2 + kf[None, :, None] ** 2 + kf[None, None, :] ** 2)
    e = 0.5 * sum(np.abs(x) ** 2 for x in uh)
    shells = np.bincount(np.rint(kk).astype(int).ravel(), weights=e.ravel())
    # expected shell sums of the mode energies k^4 exp(-2 (k/4)^2) / (4 pi k^2)
    with np.errstate(divide="ignore", invalid="ignore"):
        em = np.where(kk > 0, kk ** 2 * np.exp(-2.0 * (kk / 4.0) ** 2) / (4 * np.pi), 0.0)
    expect = np.bincount(np.rint(kk).astype(int).ravel(), weights=em.ravel())
    k = np.arange(2, 9)
    ratio = shells[k] / expect[k]
    assert ratio.max() / ratio.min() < 1.5, ratio
    # seeded: same seed, same field; another seed, another field
    assert np.array_equal(cases.isotropic_turbulence(16, seed=5).prim,
                          cases.isotropic_turbulence(16, seed=5).prim)
    assert not np.array_equal(cases.isotropic_turbulence(16, seed=5).prim,
                              cases.isotropic_turbulence(16, seed=6).prim)


def test_shock_vortex_initial_state():
    c = cases.shock_vortex(64, 32, 32, seed=2)
    p = c.params
    g, M = p.gamma, p.Mach
    rho, u, v, w, pr = c.prim
    x = c.x
    down = x > 0.6
    # Rankine-Hugoniot for Mach 2: rho 8/3, u 3/8, p 4.5 p_inf
    assert np.allclose(rho[down], 8.0 / 3.0) and np.allclose(u[down], 3.0 / 8.0)
    assert np.allclose(pr[down], 4.5 / (g * M * M))
    # mass, momentum and energy fluxes match across the jump (freestream side)
    def fluxes(r, uu, pp):
        return np.array([r * uu, r * uu * uu + pp, uu * (pp * g / (g - 1) + 0.5 * r * uu * uu)])
    assert np.allclose(fluxes(1.0, 1.0, 1.0 / (g * M * M)),
                       fluxes(8.0 / 3.0, 3.0 / 8.0, 4.5 / (g * M * M)), rtol=1e-12)
    # the vortex is isentropic: p / rho^gamma is the freestream's upstream
    up = ~down
    assert np.allclose(pr[up] / rho[up] ** g, 1.0 / (g * M * M), rtol=1e-12)
    # spacing ratio 1.5 between far field and the shock plane
    dx = np.diff(c.x[0, 0])
    assert dx.max() / dx.min() == pytest.approx(1.5, rel=2e-2)
    assert not np.any(w)
