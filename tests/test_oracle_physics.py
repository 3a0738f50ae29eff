"""Physics pins of the oracle's extensions (no reference code exists for them):
Taylor-Green decay exp(-2 nu k^2 t) on periodic edges, exact linear Couette
flow between a half-way wall and a Ladd moving lid, Zou-He inlet/outlet
values, developed Poiseuille profile and mass-flux conservation, FP64 mass
conservation."""
import numpy as np
import pytest

from chemsim_amd import workloads as W
from oracle import pyoracle as O


def test_taylor_green_decay_rate():
    n, tau, u0, steps = 64, 0.8, 0.01, 400
    rho, ux, uy = W.taylor_green(n, u0)
    p = O.params(n, n, tau=tau, edge=O.EDGE_PERIODIC)
    g = O.run(p, O.equilibrium(p, rho, ux, uy), steps)
    vx, vy = O.velocity(g)
    amp = (vx * ux + vy * uy).sum() / (ux * ux + uy * uy).sum()
    rate = -np.log(amp) / steps
    rate_exact = 2 * W.viscosity(tau) * (2 * np.pi / n) ** 2
    assert abs(rate / rate_exact - 1) < 5e-3
    # the field stays the Taylor-Green mode
    resid = np.sqrt(((vx - amp * ux) ** 2 + (vy - amp * uy) ** 2).sum() / ((amp * ux) ** 2 + (amp * uy) ** 2).sum())
    assert resid < 1e-3


def test_periodic_mass_and_momentum_conservation_f64():
    n = 24
    rho, ux, uy = W.taylor_green(n, 0.05)
    p = O.params(n, n, tau=0.8, edge=O.EDGE_PERIODIC)
    f = O.equilibrium(p, rho, ux, uy)
    s0 = O.sums(f)
    s1 = O.sums(O.run(p, f, 200))
    assert abs(s1[0] - s0[0]) < 1e-11 * s0[0]
    assert abs(s1[1] - s0[1]) < 1e-12 and abs(s1[2] - s0[2]) < 1e-12


@pytest.mark.parametrize("tau", [0.6, 1.0, 1.7])
def test_couette_exact_linear_profile(tau):
    h, w, U = 18, 4, 0.05
    fl = np.zeros((h, w), np.uint8)
    fl[0, :] = O.LID
    fl[-1, :] = O.WALL
    p = O.params(w, h, tau=tau, edge=O.EDGE_PERIODIC, boundary=O.BB_HALFWAY, lid=(0.0, U))
    r0, a, b = (x.astype(np.float64) for x in W.rest_state(w, h))
    g = O.run(p, O.equilibrium(p, r0, a, b), int(60 * (h - 2) ** 2 / W.viscosity(tau) / 10), fl)
    vx, vy = O.velocity(g)
    y = np.arange(1, h - 1) - 0.5
    np.testing.assert_allclose(vy[1:-1, 0], U * (1 - y / (h - 2)), atol=1e-10)
    assert np.abs(vx[1:-1]).max() < 1e-12


def test_zou_he_prescribed_values_after_one_step():
    ch = W.Channel(w=16, h=10, re=10.0, u_in=0.03)
    fl = ch.flags()
    p = O.params(ch.w, ch.h, tau=ch.tau, boundary=O.BB_HALFWAY, inlet=ch.inlet(), outlet_rho=1.02)
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.standard_normal((ch.h, ch.w))
    ux = 0.01 * rng.standard_normal((ch.h, ch.w))
    uy = 0.01 * rng.standard_normal((ch.h, ch.w))
    g = O.step(p, O.equilibrium(p, rho, ux, uy), fl)
    r, vx, vy = O.density(g), *O.velocity(g)
    np.testing.assert_allclose(vx[1:-1, 0], 0.0, atol=1e-15)
    np.testing.assert_allclose(vy[1:-1, 0], 0.03, atol=1e-15)
    np.testing.assert_allclose(r[1:-1, -1], 1.02, atol=1e-14)
    np.testing.assert_allclose(vx[1:-1, -1], 0.0, atol=1e-15)


def test_poiseuille_profile_and_flux():
    ch = W.Channel(w=96, h=18, re=10.0, u_in=0.02)
    fl = ch.flags()
    p = O.params(ch.w, ch.h, tau=ch.tau, boundary=O.BB_HALFWAY, inlet=ch.inlet(), outlet_rho=1.0)
    r0, a, b = (x.astype(np.float64) for x in W.rest_state(ch.w, ch.h))
    g = O.run(p, O.equilibrium(p, r0, a, b), 30000, fl)
    rho = O.density(g)
    vx, vy = O.velocity(g)
    prof = vy[1:-1, ch.w // 2]
    y = np.arange(1, ch.h - 1) - 0.5
    H = ch.h - 2
    ub = prof.mean()
    np.testing.assert_allclose(prof, 6 * ub * (y / H) * (1 - y / H), atol=0.01 * prof.max())
    # constant mass flux away from the outlet's boundary layer (~10 columns)
    flux = (rho * vy)[1:-1, 1:ch.w - 32].sum(0)
    np.testing.assert_allclose(flux, flux[0], rtol=1e-5)
    np.testing.assert_allclose(flux[0], ch.u_in * H * rho[1:-1, 0].mean(), rtol=1e-3)
