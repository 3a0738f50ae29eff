"""The committed step fixtures (tests/golden/step_fixtures.npz, written by
tests/golden/make_step_fixtures.py): 3 reference steps of seeded FP32 states.
Reproduced byte for byte by the oracle and by an independent numpy
transcription of src/lbm.rs (stream, bounce_back, Lattice::density /
momentum_density / velocity, compute_equilibrium, BGK::evaluate, in the
reference's operation order); the HIP kernels are held to the same bytes in
test_gpu_golden_step."""
import os

import numpy as np
import pytest

from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "step_fixtures.npz"))
NAMES = sorted({k.rsplit("_", 1)[0] for k in FIX.files})
F32 = np.float32
CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
CY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]


def shifted(p, dr, dc, periodic):
    """q[r][c] = p[r + dr][c + dc], zero outside (or wrapped)."""
    if periodic:
        return np.roll(p, (-dr, -dc), axis=(0, 1))
    h, w = p.shape
    q = np.zeros_like(p)
    rs, re = max(0, -dr), min(h, h - dr)
    cs, ce = max(0, -dc), min(w, w - dc)
    q[rs:re, cs:ce] = p[rs + dr:re + dr, cs + dc:ce + dc]
    return q


def numpy_reference_step(f, flags, tau, periodic, dx=F32(1), dt=F32(1)):
    # State::stream (lbm.rs:716-729): f_i(r, c) <- f_i(r + cx_i, c - cy_i)
    q = np.stack([shifted(f[i], CX[i], -CY[i], periodic) for i in range(9)])
    # State::bounce_back (lbm.rs:741-751): opposite population at solid cells
    solid = (flags == O.WALL) | (flags == O.LID)
    b = np.stack([np.where(solid, q[OPP[i]], q[i]) for i in range(9)])
    # Lattice::density / momentum_density / velocity (lbm.rs:117-138)
    rho = np.zeros_like(b[0])
    mx = np.zeros_like(b[0])
    my = np.zeros_like(b[0])
    for i in range(9):
        rho = rho + b[i]
        mx = mx + b[i] * F32(CX[i])
        my = my + b[i] * F32(CY[i])
    inv = F32(1) / rho
    vx, vy = inv * mx, inv * my
    # compute_equilibrium (lbm.rs:43-71), weights (lbm.rs:209-219)
    v2 = vx * vx + vy * vy
    cs = dx / (np.sqrt(F32(3)) * dt)
    cs2 = cs * cs
    cs4 = cs2 * cs2
    ws = [F32(16) / F32(36)] + [F32(4) / F32(36)] * 4 + [F32(1) / F32(36)] * 4
    factor = (-dt) / F32(tau)  # BGK::evaluate (lbm.rs:350-364)
    out = np.empty_like(b)
    for i in range(9):
        vc = vx * F32(CX[i]) + vy * F32(CY[i])
        vc2 = vc * vc
        s = ((F32(1) + vc * (F32(1) / cs2)) + vc2 * (F32(1) / (F32(2) * cs4))) + v2 * (F32(-1) / (F32(2) * cs2))
        feq = (rho * ws[i]) * s
        out[i] = b[i] + (b[i] - feq) * factor
    return out


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    tau, edge, steps = FIX[name + "_meta"]
    f0, flags = FIX[name + "_in"], FIX[name + "_flags"]
    h, w = flags.shape
    op = O.params(w, h, tau=float(tau), edge=int(edge))
    got = O.run(op, f0, int(steps), flags if flags.any() else None)
    np.testing.assert_array_equal(got, FIX[name + "_out"])


@pytest.mark.parametrize("name", NAMES)
def test_numpy_transcription_reproduces_fixture(name):
    tau, edge, steps = FIX[name + "_meta"]
    f, flags = FIX[name + "_in"].copy(), FIX[name + "_flags"]
    with np.errstate(all="ignore"):
        for _ in range(int(steps)):
            f = numpy_reference_step(f, flags, F32(tau), int(edge) == O.EDGE_PERIODIC)
    np.testing.assert_array_equal(f, FIX[name + "_out"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_golden_step(ctx, name):
    from chemsim_amd import Lattice, Params, capi
    tau, edge, steps = FIX[name + "_meta"]
    f0, flags = FIX[name + "_in"], FIX[name + "_flags"]
    h, w = flags.shape
    p = Params(tau=float(tau), edge=int(edge))
    with Lattice(ctx, w, h, "f32", p, flags if flags.any() else None) as lat:
        lat.upload(f0)
        lat.step(int(steps))
        np.testing.assert_array_equal(lat.download(), FIX[name + "_out"])
