"""Pins of the oracle's TRT operator (lbm.rs:374-451), next to the BGK hot
path.  The reference's TRT::evaluate takes swap_equilibrium, which keeps
f_eq_0 and puts the *populations* of the opposite directions in 1..8
(lbm.rs:311-323); its result therefore reduces to
    o_0 = f_0 - dt wp (f_0 - f_eq_0),
    o_i = f_i - dt (wp + wm) / 2 (f_i - f_eq_i),  i >= 1,
which is what these tests pin (a textbook TRT, swapped rates or the
equilibrium in the swap would all fail them)."""
import numpy as np
import pytest

from oracle import pyoracle as O


def rand_f(rng, h, w, dtype=np.float64):
    rho = 1 + 0.02 * rng.standard_normal((h, w))
    ux, uy = 0.04 * rng.standard_normal((2, h, w))
    p = O.params(w, h)
    return (O.equilibrium(p, rho, ux, uy) + 0.003 * rng.standard_normal((9, h, w))).astype(dtype)


@pytest.mark.parametrize("tp,tm", [(0.8, 0.8), (0.7, 1.3), (1.9, 0.55)])
def test_reduced_form(tp, tm):
    rng = np.random.default_rng(4)
    h, w = 6, 7
    f = rand_f(rng, h, w)
    p = O.params(w, h, trt=(tp, tm))
    o = O.trt(p, f)
    feq = O.equilibrium(p, O.density(f), *O.velocity(f))
    wp, wm = 1.0 / tp, 1.0 / tm
    ref = f - 0.5 * (wp + wm) * (f - feq)
    ref[0] = f[0] - wp * (f[0] - feq[0])
    np.testing.assert_allclose(o, ref, rtol=0, atol=2e-16)


def test_equal_rates_is_bgk():
    rng = np.random.default_rng(5)
    f = rand_f(rng, 5, 9)
    for tau in (0.6, 1.0, 1.7):
        o_trt = O.trt(O.params(9, 5, trt=(tau, tau)), f)
        pb = O.params(9, 5, tau=tau)
        o_bgk = O.bgk(pb, f, O.equilibrium(pb, O.density(f), *O.velocity(f)))
        np.testing.assert_allclose(o_trt, o_bgk, rtol=0, atol=1e-16)


def test_trt_new_inverts_lambda_and_viscosity():
    """TRT::new (lbm.rs:380-390) against its inverse lambda() (lbm.rs:392-399)
    and the viscosity cs^2 (tau+ - 1/2) (lbm.rs:447-450)."""
    for lam, nu in ((0.25, 10.0), (3.0 / 16, 0.1), (0.1, 0.02)):
        tp, tm = O.trt_new(lam, nu)
        assert abs((tp - 0.5) * (tm - 0.5) - lam) < 2e-6 * max(1.0, lam)
        assert abs((tp - 0.5) / 3.0 - nu) < 2e-6 * max(1.0, nu)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
def test_fused_equals_unfused(dtype, edge):
    rng = np.random.default_rng(6)
    h, w = 9, 11
    f = rand_f(rng, h, w, dtype)
    flags = (rng.random((h, w)) < 0.15).astype(np.uint8)
    p = O.params(w, h, edge=edge, trt=(0.7, 1.2))
    g = O.run(p, f, 5, flags)
    if dtype == np.float32:
        np.testing.assert_array_equal(g, O.run_unfused(p, f, 5, flags))
    # one step by hand: stream, bounce back, TRT on every cell
    q = O.bounce_back(O.stream(f, edge), flags)
    np.testing.assert_array_equal(O.step(p, f, flags), O.trt(p, q))
