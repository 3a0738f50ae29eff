"""Pins of the oracle's Regularized and KBC operators (src/lbm.rs:606-661,
468-572) by what the mathematics fixes, independent of the oracle's code:

Regularized (Latt & Chopard projection): equilibrium is a fixed point; mass and
momentum are conserved (the Hermite tensors Q_i = c_i c_i - cs^2 I carry none);
the operator is a projector (applying it twice equals applying it once), which
fails for a dropped or transposed stress component or a wrong cs^2 term.

KBC: equilibrium is a fixed point (gamma* = 2 whatever the shear part is), and
on inputs whose velocity is exactly zero, purely along x, purely along y, or
exactly diagonal (u == v), the reference's shear-part formulas reduce by hand
to short closed forms (below) that the oracle's output must follow.  The four
cases separate u from v in every delta_s term and exercise the u v and
pi~ = sum f dx^2 - u v cross terms.
"""
import numpy as np
import pytest

from oracle import pyoracle as O

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.float64)
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
W = np.array([16, 4, 4, 4, 4, 1, 1, 1, 1], dtype=np.float64) / 36.0


def textbook_feq(rho, ux, uy, cs2):
    cu = CX[:, None] * ux[None] + CY[:, None] * uy[None]
    u2 = ux * ux + uy * uy
    return W[:, None] * rho[None] * (1 + cu / cs2 + cu * cu / (2 * cs2 * cs2) - u2 / (2 * cs2))


def moments(f):
    rho = f.sum(0)
    return rho, (f * CX[:, None]).sum(0) / rho, (f * CY[:, None]).sum(0) / rho


def near_equilibrium(n, seed, eps=2e-3):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.05 * rng.standard_normal(n)
    ux = 0.05 * rng.standard_normal(n)
    uy = 0.05 * rng.standard_normal(n)
    f = textbook_feq(rho, ux, uy, 1.0 / 3.0) + eps * rng.standard_normal((9, n)) * W[:, None]
    return f


# ---------------------------------------------------------------- Regularized
# The projection properties need cs^2 = sum_i w_i c_ix^2 = 1/3, i.e. dx = dt:
# the reference's lattice vectors are the integer c_i whatever dx and dt are.
@pytest.mark.parametrize("dx,dt", [(1.0, 1.0), (0.5, 0.5)])
def test_regularized_equilibrium_fixed_point(dx, dt):
    p = O.params(64, 1, regularized=True, delta_x=dx, delta_t=dt)
    rng = np.random.default_rng(1)
    rho = 1.0 + 0.05 * rng.standard_normal(64)
    ux, uy = 0.03 * rng.standard_normal(64), 0.03 * rng.standard_normal(64)
    feq = O.equilibrium(p, rho, ux, uy).reshape(9, -1)
    out = O.regularized(p, feq)
    np.testing.assert_allclose(out, feq, rtol=0, atol=1e-15)


@pytest.mark.parametrize("dx,dt", [(1.0, 1.0), (0.5, 0.5)])
def test_regularized_conserves_mass_and_momentum(dx, dt):
    p = O.params(256, 1, regularized=True, delta_x=dx, delta_t=dt)
    f = near_equilibrium(256, 2, eps=1e-2)
    out = O.regularized(p, f)
    np.testing.assert_allclose(out.sum(0), f.sum(0), rtol=0, atol=1e-14)
    np.testing.assert_allclose((out * CX[:, None]).sum(0), (f * CX[:, None]).sum(0), rtol=0, atol=1e-14)
    np.testing.assert_allclose((out * CY[:, None]).sum(0), (f * CY[:, None]).sum(0), rtol=0, atol=1e-14)
    assert np.abs(out - f).max() > 1e-4  # it did change the ghost moments


def test_regularized_is_a_projector():
    p = O.params(256, 1, regularized=True)
    f = near_equilibrium(256, 3, eps=1e-2)
    once = O.regularized(p, f)
    twice = O.regularized(p, once)
    np.testing.assert_allclose(twice, once, rtol=0, atol=2e-15)
    # the second-order stress of the non-equilibrium part is kept
    rho, ux, uy = moments(f)
    feq = textbook_feq(rho, ux, uy, 1.0 / 3.0)
    for a, b in [(CX, CX), (CX, CY), (CY, CY)]:
        sa = ((f - feq) * (a * b)[:, None]).sum(0)
        sb = ((once - feq) * (a * b)[:, None]).sum(0)
        np.testing.assert_allclose(sb, sa, rtol=0, atol=1e-14)


def test_regularized_f32_follows_f64():
    p = O.params(256, 1, regularized=True)
    f = near_equilibrium(256, 4, eps=1e-2)
    o64 = O.regularized(p, f)
    o32 = O.regularized(p, f.astype(np.float32))
    np.testing.assert_allclose(o32, o64, rtol=0, atol=2e-7)


# ------------------------------------------------------------------------ KBC
def dyadic(rng, lo, hi, n, bits=12):
    return rng.integers(int(lo * 2**bits), int(hi * 2**bits), n).astype(np.float64) / 2**bits


def symmetric_populations(kind, n, seed):
    """Populations whose velocity is exactly 0, (u, 0), (0, v) or (u, u):
    dyadic values with mirror-symmetric pairs, so every moment sum is exact."""
    rng = np.random.default_rng(seed)
    f = np.empty((9, n))
    f[0] = dyadic(rng, 0.40, 0.47, n)
    for i in range(1, 5):
        f[i] = dyadic(rng, 0.09, 0.13, n)
    for i in range(5, 9):
        f[i] = dyadic(rng, 0.02, 0.035, n)
    if kind == "rest":     # f1 = f3, f2 = f4, f5 = f6 = f7 = f8
        f[3], f[4] = f[1], f[2]
        f[6], f[7], f[8] = f[5], f[5], f[5]
    elif kind == "x":      # y-mirror: f2 = f4, f5 = f8, f6 = f7
        f[4], f[8], f[7] = f[2], f[5], f[6]
    elif kind == "y":      # x-mirror: f1 = f3, f5 = f6, f7 = f8
        f[3], f[6], f[8] = f[1], f[5], f[7]
    elif kind == "diag":   # x <-> y swap: f1 = f2, f3 = f4, f6 = f8
        f[2], f[4], f[8] = f[1], f[3], f[6]
    return f


def kbc_closed_form(kind, f, dx, dt, nu):
    """The reference's KBC with delta_s reduced by hand for the case."""
    cs2 = (dx / (np.sqrt(3.0) * dt)) ** 2
    rho, u, v = moments(f)
    if kind in ("rest", "x"):
        v = np.zeros_like(v)
    if kind in ("rest", "y"):
        u = np.zeros_like(u)
    feq = textbook_feq(rho, u, v, cs2)
    pit0 = rho * dx * dx                                   # sum_i f_i dx^2
    z = np.zeros_like(rho)
    if kind == "rest":    # u = v = 0: only the diagonal part, rho^2 dx^4 / 4
        d0, d13, d24 = z, z, z
        d58 = rho * rho * dx**4 / 4
    elif kind == "x":     # v = 0: n~ = -u^2, pi~ = rho dx^2
        d0 = u**4 * rho / 2
        d13 = -u * u * (1 + u * dx + u * u) * rho / 4
        d24 = (u * u * (1 - u * u) - 4 * u * dx * pit0) * rho / 4
        d58 = ((4 * u * dx + 2 * dx * dx) * pit0 + u * u * (u * u - u * dx)) * rho / 8
    elif kind == "y":     # u = 0: n~ = v^2, pi~ = rho dx^2
        d0 = v**4 * rho / 2
        d13 = ((1 - v * v) * v * v - 4 * v * dx * pit0) * rho / 4
        d24 = -(v * dx + v * v + 1) * v * v * rho / 4
        d58 = (2 * dx * dx * pit0 + (v * v - v * dx) * v * v) * rho / 8
    else:                 # u = v: n~ = 0, pi~ = rho dx^2 - u^2
        pit = pit0 - u * u
        d0 = 4 * u * u * pit * rho
        d13 = -(4 * u * dx + 8 * u * u) * pit * rho / 4
        d24 = d13
        d58 = (8 * u * u + 4 * u * dx + 2 * dx * dx) * pit * rho / 8
    ds = np.stack([d0, d13, d24, d13, d24, d58, d58, d58, d58])
    dh = f - feq - ds
    beta = 1.0 / (2.0 * nu / cs2 + 1.0)
    num = (ds * dh / feq).sum(0)
    den = (dh * dh / feq).sum(0)
    gamma = -((num / den) * (2.0 - 1.0 / beta) - 1.0 / beta)
    return f - 2.0 * beta * ds - beta * gamma * dh


@pytest.mark.parametrize("kind", ["rest", "x", "y", "diag"])
@pytest.mark.parametrize("dx,dt,nu", [(1.0, 1.0, 0.05), (0.5, 0.25, 0.3)])
def test_kbc_reduced_cases(kind, dx, dt, nu):
    f = symmetric_populations(kind, 512, {"rest": 5, "x": 6, "y": 7, "diag": 8}[kind])
    rho, u, v = moments(f)
    assert np.abs(u).max() > 1e-3 or kind in ("rest", "y")
    p = O.params(512, 1, kbc=nu, delta_x=dx, delta_t=dt)
    got = O.kbc(p, f)
    want = kbc_closed_form(kind, f, dx, dt, nu)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)
    assert np.abs(got - f).max() > 1e-4


@pytest.mark.parametrize("dx,dt", [(1.0, 1.0), (0.5, 0.5)])
def test_kbc_equilibrium_fixed_point(dx, dt):
    """(dx = dt: otherwise the reference's f_eq does not reproduce its own
    velocity, sum_i w_i c_i c_i = I/3 != cs^2 I, and is no fixed point.)"""
    p = O.params(64, 1, kbc=0.1, delta_x=dx, delta_t=dt)
    rng = np.random.default_rng(9)
    rho = 1.0 + 0.05 * rng.standard_normal(64)
    ux, uy = 0.03 * rng.standard_normal(64), 0.03 * rng.standard_normal(64)
    feq = O.equilibrium(p, rho, ux, uy).reshape(9, -1)
    out = O.kbc(p, feq)
    np.testing.assert_allclose(out, feq, rtol=0, atol=1e-13)


def test_kbc_f32_follows_f64():
    p = O.params(256, 1, kbc=0.1)
    f = near_equilibrium(256, 10, eps=1e-2)
    o64 = O.kbc(p, f)
    o32 = O.kbc(p, f.astype(np.float32))
    np.testing.assert_allclose(o32, o64, rtol=0, atol=5e-6)


@pytest.mark.parametrize("coll", ["regularized", "kbc"])
def test_fused_step_matches_unfused(coll):
    """The fused step (stream, bounce-back, collide per cell) and the unfused
    phases agree for the new operators as for BGK."""
    from chemsim_amd import workloads as Wk
    n = 24
    kw = {"regularized": True} if coll == "regularized" else {"kbc": 0.05}
    p = O.params(n, n, tau=0.8, lid=Wk.lid_velocity(0.05), **kw)
    flags = Wk.cavity_flags(n, n)
    rho, ux, uy = Wk.rest_state(n, n)
    f0 = O.equilibrium(p, rho, ux, uy)
    a = O.run(p, f0.copy(), 5, flags)
    b = O.run_unfused(p, f0.copy(), 5, flags)
    np.testing.assert_array_equal(a, b)
