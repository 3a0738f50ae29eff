"""Pins of two oracle functions against facts from outside the oracle.

1. The deviation-form Zou-He inlet/outlet (oracle/lbm_oracle.c
   zou_he_inlet_dev / zou_he_outlet_dev, used by every f32dev and 16-bit
   channel run).  Zou-He fixes the unknown populations of a boundary cell so
   that its moments are the prescribed ones: at an inlet cell rho u equals
   rho (inlet_ux, inlet_uy), at an outlet cell rho equals outlet_rho with no
   tangential velocity.  The collision that follows conserves mass and
   momentum, so after one step the stored populations' own moments -- here
   computed with numpy in float64 from the deviations, rho = rho0 + sum g_i,
   m = sum g_i c_i -- must carry the prescribed values to f32 rounding.  A
   dropped term, a wrong sign or a 1/6 <-> 2/3 swap in those formulas moves
   them by the size of that term (checked by planting such mistakes in a
   numpy copy of the formulas below).  Second, the deviation-form channel step
   is the full-form update computed more precisely: against the FP64
   full-form step it is at least as accurate as the FP32 full form.
2. The FP64 reductions orc_sums / orc_residual (the conservation sums of
   Lattice::density / momentum_density, lbm.rs:117-131, and the steady-state
   residual ||u - u_prev|| / ||u|| over non-solid cells) against math.fsum of
   numpy float64 values, including u = 0 and all-solid lattices.
"""
import math

import numpy as np
import pytest

from chemsim_amd import workloads as W
from oracle import pyoracle as O

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def channel_case(seed, w=24, h=12, u_in=0.04, outlet_rho=1.015):
    rng = np.random.default_rng(seed)
    ch = W.Channel(w=w, h=h, circles=[(h // 2, w // 2, 2)], re=12.0, u_in=u_in)
    fl = ch.flags()
    p = O.params(w, h, tau=ch.tau, boundary=O.BB_HALFWAY, inlet=ch.inlet(), outlet_rho=outlet_rho)
    rho = 1 + 0.01 * rng.standard_normal((h, w))
    ux = 0.01 * rng.standard_normal((h, w))
    uy = u_in + 0.01 * rng.standard_normal((h, w))
    f64 = O.equilibrium(p, rho, ux, uy) + 2e-4 * rng.standard_normal((9, h, w))
    return p, fl, f64


def moments_of_dev(p, g):
    """(rho, ux, uy) in float64 of f32 deviations g_i = f_i - w_i rho0."""
    g = g.astype(np.float64)
    rho = p.rho0 + g.sum(0)
    mx = np.tensordot(CX, g, axes=1)
    my = np.tensordot(CY, g, axes=1)
    return rho, mx / rho, my / rho


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dev_zou_he_prescribes_inlet_velocity_and_outlet_density(seed):
    p, fl, f64 = channel_case(seed)
    g = O.step_devf32(p, O.to_dev(p, f64.astype(np.float32)), fl)
    rho, ux, uy = moments_of_dev(p, g)
    inlet, outlet = fl == O.INLET, fl == O.OUTLET
    assert inlet.sum() == p.h - 2 and outlet.sum() == p.h - 2
    # f32 rounding of deviations of size ~1e-2: momentum to ~1e-8
    np.testing.assert_allclose(ux[inlet], p.inlet_ux, atol=2e-7)
    np.testing.assert_allclose(uy[inlet], p.inlet_uy, atol=2e-7)
    np.testing.assert_allclose(rho[outlet], p.outlet_rho, atol=2e-7)
    np.testing.assert_allclose(ux[outlet], 0.0, atol=2e-7)
    # the prescribed values are not already there before the step
    r0, a0, b0 = moments_of_dev(p, O.to_dev(p, f64.astype(np.float32)))
    assert np.abs(b0[inlet] - p.inlet_uy).max() > 1e-3 and np.abs(r0[outlet] - p.outlet_rho).max() > 1e-3


def zou_he_inlet_numpy(B, rho0, ux, uy, plant=None):
    """Zou-He at a velocity inlet on post-stream deviations B[9] (moving +y
    is the along-channel direction), float64, written from the moment
    conditions: rho u = given, non-equilibrium bounce-back of the normal
    population.  `plant` introduces one mistake."""
    G = B[0] + B[1] + B[3] + 2 * (B[4] + B[7] + B[8])
    rho = (rho0 + G) / (1 - uy)
    ru_t, ru_a, d13 = rho * ux, rho * uy, B[1] - B[3]
    k_a, k_d = 2.0 / 3.0, 1.0 / 6.0
    if plant == "swap":
        k_a, k_d = 1.0 / 6.0, 2.0 / 3.0
    B = B.copy()
    B[2] = B[4] + k_a * ru_a
    sgn = -1.0 if plant == "sign" else 1.0
    B[5] = B[7] - 0.5 * d13 + sgn * 0.5 * ru_t + k_d * ru_a
    B[6] = B[8] + 0.5 * d13 - 0.5 * ru_t + (0.0 if plant == "drop" else k_d) * ru_a
    return B


def test_zou_he_moment_conditions_reject_planted_mistakes():
    """The moment check of the test above separates the correct formulas from
    plausible mistakes (on a numpy copy, in float64)."""
    rng = np.random.default_rng(5)
    B = 0.01 * rng.standard_normal(9)
    ux, uy = 0.01, 0.04

    def mom(Bz):
        rho = 1.0 + Bz.sum()
        return rho, (CX @ Bz) / rho, (CY @ Bz) / rho

    _, gx, gy = mom(zou_he_inlet_numpy(B, 1.0, ux, uy))
    assert abs(gx - ux) < 1e-15 and abs(gy - uy) < 1e-15
    for plant in ("swap", "sign", "drop"):
        _, bx, by = mom(zou_he_inlet_numpy(B, 1.0, ux, uy, plant))
        assert max(abs(bx - ux), abs(by - uy)) > 1e-4, plant


@pytest.mark.parametrize("steps", [1, 30])
def test_dev_zou_he_channel_matches_fp64_more_closely_than_full_f32(steps):
    """Deviation-form channel steps (Zou-He inlet/outlet, half-way walls, an
    obstacle) against the FP64 full-form steps: at least as accurate as the
    full-form FP32 steps."""
    p, fl, f64 = channel_case(11)
    f32 = f64.astype(np.float32)
    ref = O.run(p, f32.astype(np.float64), steps, fl)
    full = O.run(p, f32, steps, fl).astype(np.float64)
    dev = O.from_dev(p, O.run_devf32(p, O.to_dev(p, f32), steps, fl)).astype(np.float64)
    base = np.array(O.consts_f64(p).base)[:, None, None]
    fluid = (fl != O.WALL)[None]
    scale = np.abs((ref - base) * fluid).max()
    err_full = np.abs((full - ref) * fluid).max() / scale
    err_dev = np.abs((dev - ref) * fluid).max() / scale
    assert err_dev < 2e-5, err_dev
    assert err_dev <= err_full * 1.5, (err_dev, err_full)


# ---------------------------------------------------------- FP64 reductions --

def f32_moments(f):
    """Per-cell density and momentum in f32, summed left to right in direction
    order as Lattice::density / momentum_density do (products by c_i are
    exact), with numpy's IEEE float32 adds."""
    f = np.asarray(f, np.float32)
    rho = np.zeros(f.shape[1:], np.float32)
    mx = np.zeros(f.shape[1:], np.float32)
    my = np.zeros(f.shape[1:], np.float32)
    for i in range(9):
        rho = (rho + f[i]).astype(np.float32)
        mx = (mx + np.float32(CX[i]) * f[i]).astype(np.float32)
        my = (my + np.float32(CY[i]) * f[i]).astype(np.float32)
    return rho, mx, my


def fluid_mask(flags, shape):
    if flags is None:
        return np.ones(shape, bool)
    return (flags != O.WALL) & (flags != O.LID)


@pytest.mark.parametrize("case", ["open", "masked", "all_solid", "at_rest"])
def test_sums_equal_exact_sums_of_cell_moments(case):
    rng = np.random.default_rng({"open": 1, "masked": 2, "all_solid": 3, "at_rest": 4}[case])
    h, w = 37, 53
    p = O.params(w, h)
    rho = 1 + 0.05 * rng.standard_normal((h, w))
    ux = 0.08 * rng.standard_normal((h, w))
    uy = 0.08 * rng.standard_normal((h, w))
    if case == "at_rest":
        ux[:] = 0
        uy[:] = 0
    f = (O.equilibrium(p, rho, ux, uy) + 1e-3 * rng.standard_normal((9, h, w))).astype(np.float32)
    flags = None
    if case == "masked":
        flags = rng.choice(np.array([O.FLUID, O.WALL, O.LID, O.INLET, O.OUTLET], np.uint8), size=(h, w),
                           p=[0.6, 0.2, 0.1, 0.05, 0.05])
    elif case == "all_solid":
        flags = np.full((h, w), O.WALL, np.uint8)
    got = O.sums(f, flags)
    r, mx, my = f32_moments(f)
    keep = fluid_mask(flags, (h, w))
    for k, m in enumerate((r, mx, my)):
        vals = m[keep].astype(np.float64)
        want = math.fsum(vals)
        bound = 1e-13 * math.fsum(np.abs(vals)) + 1e-300
        assert abs(got[k] - want) <= bound, (case, k, got[k], want)
    if case == "all_solid":
        assert (got == 0).all()
    # the FP64 storage sums the same moments in f64 arithmetic
    f64 = f.astype(np.float64)
    got64 = O.sums(f64, flags)
    want_rho = math.fsum(f64.sum(0)[keep]) if keep.any() else 0.0
    assert abs(got64[0] - want_rho) <= 1e-12 * max(1.0, abs(want_rho))


@pytest.mark.parametrize("case", ["moving", "masked", "zero_field", "zero_now", "all_solid"])
def test_residual_equals_exact_norm_ratio(case):
    rng = np.random.default_rng(7)
    h, w = 29, 41
    ux = (0.05 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.05 * rng.standard_normal((h, w))).astype(np.float32)
    px = (ux + 1e-4 * rng.standard_normal((h, w))).astype(np.float32)
    py = (uy + 1e-4 * rng.standard_normal((h, w))).astype(np.float32)
    flags = None
    if case == "masked":
        flags = rng.choice(np.array([O.FLUID, O.WALL, O.LID, O.INLET], np.uint8), size=(h, w),
                           p=[0.7, 0.15, 0.1, 0.05])
    elif case == "all_solid":
        flags = np.full((h, w), O.LID, np.uint8)
    elif case == "zero_field":
        ux[:] = uy[:] = px[:] = py[:] = 0
    elif case == "zero_now":
        ux[:] = uy[:] = 0
    got = O.residual(ux, uy, px, py, flags)
    keep = fluid_mask(flags, (h, w))
    dx = ux[keep].astype(np.float64) - px[keep].astype(np.float64)
    dy = uy[keep].astype(np.float64) - py[keep].astype(np.float64)
    num = math.fsum(np.concatenate([dx * dx, dy * dy]))
    den = math.fsum(np.concatenate([ux[keep].astype(np.float64) ** 2, uy[keep].astype(np.float64) ** 2]))
    if case in ("zero_field", "all_solid"):
        assert got == 0.0
    elif case == "zero_now":
        assert got == math.inf
    else:
        want = math.sqrt(num / den)
        assert abs(got - want) <= 1e-13 * want, (got, want)
        assert 1e-4 < want < 1e-1
