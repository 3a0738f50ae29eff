"""Independent numpy transcriptions (float32, the reference's operation order)
of TRT::evaluate (src/lbm.rs:401-444), Regularized::evaluate (:606-661) and
KBC::evaluate (:468-572), compared byte for byte with the oracle's C
restatement on seeded states: two separate transcriptions of the same
reference code must agree to the last bit."""
import numpy as np
import pytest

from oracle import pyoracle as O

F32 = np.float32
CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
CY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]
WS = [F32(16) / F32(36)] + [F32(4) / F32(36)] * 4 + [F32(1) / F32(36)] * 4


def states(seed, n=4096):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.03 * rng.standard_normal(n)
    ux, uy = 0.05 * rng.standard_normal(n), 0.05 * rng.standard_normal(n)
    p = O.params(n, 1)
    f = O.equilibrium(p, rho, ux, uy) + 0.004 * rng.standard_normal((9, n))
    return f.astype(np.float32)


def lattice_equilibrium(f, dx=F32(1), dt=F32(1)):
    """Lattice::density / velocity + compute_equilibrium (lbm.rs:117-160, 43-71)."""
    rho = np.zeros_like(f[0])
    mx = np.zeros_like(f[0])
    my = np.zeros_like(f[0])
    for i in range(9):
        rho = rho + f[i]
        mx = mx + f[i] * F32(CX[i])
        my = my + f[i] * F32(CY[i])
    inv = F32(1) / rho
    vx, vy = inv * mx, inv * my
    v2 = vx * vx + vy * vy
    cs = dx / (np.sqrt(F32(3)) * dt)
    cs2 = cs * cs
    cs4 = cs2 * cs2
    feq = np.empty_like(f)
    for i in range(9):
        vc = vx * F32(CX[i]) + vy * F32(CY[i])
        s = ((F32(1) + vc * (F32(1) / cs2)) + (vc * vc) * (F32(1) / (F32(2) * cs4))) + v2 * (F32(-1) / (F32(2) * cs2))
        feq[i] = (rho * WS[i]) * s
    return feq, rho, vx, vy, cs


def trt(f, tp, tm, dt=F32(1)):
    feq, *_ = lattice_equilibrium(f)
    f_sw = f[OPP]                                   # swap_populations (lbm.rs:298-309)
    feq_sw = np.concatenate([feq[:1], f[OPP][1:]])  # swap_equilibrium (lbm.rs:311-323)
    fp, fm = f + f_sw, f - f_sw
    feqp, feqm = feq + feq_sw, feq - feq_sw
    wp, wm = F32(1) / F32(tp), F32(1) / F32(tm)
    omega = ((fp - feqp) * wp + (fm - feqm) * wm) * (-dt * F32(0.5))
    return f + omega


def regularized(f):
    feq, _, _, _, cs = lattice_equilibrium(f)
    fneq = f - feq
    cs2 = cs * cs
    cs4 = cs2 * cs2
    z = np.zeros_like(f[0])
    sxx, sxy, syx, syy = z, z, z, z
    for i in range(9):
        c0, c1 = F32(CX[i]), F32(CY[i])
        sxx = sxx + fneq[i] * (c0 * c0)
        sxy = sxy + fneq[i] * (c0 * c1)
        syx = syx + fneq[i] * (c1 * c0)
        syy = syy + fneq[i] * (c1 * c1)
    out = np.empty_like(f)
    for i in range(9):
        c0, c1 = F32(CX[i]), F32(CY[i])
        sf = WS[i] / (F32(2) * cs4)
        reg = feq[i]
        reg = reg + sxx * ((c0 * c0 - cs2) * sf)
        reg = reg + sxy * ((c0 * c1) * sf)
        reg = reg + syx * ((c1 * c0) * sf)
        reg = reg + syy * ((c1 * c1 - cs2) * sf)
        out[i] = reg
    return out


def kbc(f, nu, dx=F32(1)):
    feq, rho, u, v, cs = lattice_equilibrium(f, dx=dx)
    uv, u2, v2 = u * v, u * u, v * v
    temp = np.zeros_like(f[0])
    for i in range(9):
        temp = temp + f[i] * (dx * dx)
    pit = temp - uv
    nt = v2 - u2
    d0 = (((uv * F32(8)) * pit + nt * nt) * rho) * F32(0.5)
    d13 = (((((u * dx) - nt) + F32(1)) * nt - ((v * (dx * F32(4))) + uv * F32(8)) * pit) * rho) * F32(0.25)
    d24 = (((((v * (-dx)) - nt) + F32(-1)) * nt - ((u * (dx * F32(4))) + uv * F32(8)) * pit) * rho) * F32(0.25)
    d58 = (((((uv * F32(8)) + u * (F32(4) * dx)) + (F32(2) * dx) * dx) * pit + (nt - (v - u) * dx) * nt) * rho) * F32(0.125)
    ds = [d0, d13, d24, d13, d24, d58, d58, d58, d58]
    dh = [(f[i] - feq[i]) - ds[i] for i in range(9)]
    beta = F32(1) / ((F32(2) * F32(nu) / (cs * cs)) + F32(1))
    num = np.zeros_like(f[0])
    den = np.zeros_like(f[0])
    for i in range(9):
        num = num + (ds[i] * dh[i]) / feq[i]
        den = den + (dh[i] * dh[i]) / feq[i]
    gamma = (((num / den) * (F32(2) - F32(1) / beta)) + (F32(-1) / beta)) * F32(-1)
    return np.stack([f[i] + (ds[i] * (F32(2) * -beta) + (dh[i] * gamma) * (-beta)) for i in range(9)])


@pytest.mark.parametrize("seed", [1, 2])
def test_trt_transcription(seed):
    f = states(seed)
    tp, tm = O.trt_new(0.25, 0.1)
    got = O.trt(O.params(f.shape[1], 1, trt=(tp, tm)), f)
    np.testing.assert_array_equal(got, trt(f, tp, tm))


@pytest.mark.parametrize("seed", [3, 4])
def test_regularized_transcription(seed):
    f = states(seed)
    np.testing.assert_array_equal(O.regularized(O.params(f.shape[1], 1, regularized=True), f), regularized(f))


@pytest.mark.parametrize("seed,nu", [(5, 0.05), (6, 0.3)])
def test_kbc_transcription(seed, nu):
    f = states(seed)
    with np.errstate(all="ignore"):
        want = kbc(f, nu)
    np.testing.assert_array_equal(O.kbc(O.params(f.shape[1], 1, kbc=nu), f), want)
