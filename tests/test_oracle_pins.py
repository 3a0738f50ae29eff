"""Pins of the CPU oracle against facts that do not come from the oracle itself.

- Streaming: scipy.signal.convolve2d (true 2-D convolution, 'same' size, zero
  fill -- the semantics of af::convolve2 with ConvMode::DEFAULT) applied to the
  one-hot stencils as written in src/lbm.rs:233-269 and transposed as at
  src/lbm.rs:718, plus single-pulse displacement checks.
- Constants: the f32 bit patterns of 1/cs^2, 1/(2cs^4), -1/(2cs^2)
  (src/lbm.rs:54-66,84) and the weights (src/lbm.rs:209-219).
- Equilibrium and BGK: closed forms and conservation laws.
- The fused step equals the unfused op-by-op restatement bit for bit.
- 16-bit conversions: numpy float16 and torch bfloat16 rounding.
"""
import numpy as np
import pytest
from scipy.signal import convolve2d

from oracle import pyoracle as O

CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
CY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]
# One-hot positions (row, col) of the 3x3 stencils as written at src/lbm.rs:233-269.
STENCIL_ONE = [(1, 1), (1, 0), (2, 1), (1, 2), (0, 1), (2, 0), (2, 2), (0, 2), (0, 0)]


def f32bits(x):
    return int(np.float32(x).view(np.uint32))


def test_equilibrium_constants_bit_patterns():
    k = O.consts_f32(O.params(4, 4, tau=0.6))
    assert f32bits(k.A) == 0x40400001       # 3.0000002
    assert f32bits(k.B) == 0x40900001       # 4.5000005
    assert f32bits(k.C) == 0xBFC00001       # -1.5000001
    assert f32bits(k.cs2) == 0x3EAAAAAA
    assert k.w[0] == np.float32(16) / np.float32(36)
    assert k.w[1] == np.float32(4) / np.float32(36)
    assert k.w[5] == np.float32(1) / np.float32(36)
    assert np.float32(k.K) == np.float32(-1.0) / np.float32(0.6)
    k8 = O.consts_f32(O.params(4, 4, tau=0.8))
    assert np.float32(k8.K) == np.float32(-1.25)


def test_deviation_constants_are_exact():
    k = O.consts_devf32(O.params(4, 4, tau=0.6))
    assert (k.A, k.B, k.C) == (3.0, 4.5, -1.5)
    k32 = O.consts_f32(O.params(4, 4, tau=0.6))
    assert k.K == k32.K and list(k.w) == list(k32.w)


def test_f64_constants():
    k = O.consts_f64(O.params(4, 4, tau=0.8))
    assert abs(k.A - 3.0) < 1e-15 and abs(k.B - 4.5) < 1e-14 and abs(k.C + 1.5) < 1e-15
    assert k.K == -1.25
    assert abs(sum(k.w) - 1.0) < 1e-15


@pytest.mark.parametrize("shape", [(7, 5), (16, 16), (9, 13)])
def test_stream_matches_true_convolution(shape):
    h, w = shape
    rng = np.random.default_rng(1)
    f = rng.random((9, h, w)).astype(np.float32)
    q = O.stream(f, O.EDGE_ZERO)
    for i in range(9):
        s = np.zeros((3, 3), np.float32)
        s[STENCIL_ONE[i]] = 1.0
        ref = convolve2d(f[i], s.T, mode="same", boundary="fill", fillvalue=0.0)
        np.testing.assert_array_equal(q[i], ref.astype(np.float32))


def test_stream_single_pulse_displacement():
    h, w = 9, 11
    for i in range(9):
        f = np.zeros((9, h, w), np.float64)
        f[i, 4, 5] = 1.0
        q = O.stream(f, O.EDGE_ZERO)
        r, c = np.argwhere(q[i] == 1.0)[0]
        assert (r - 4, c - 5) == (-CX[i], CY[i])
        assert q.sum() == 1.0


def test_stream_periodic_is_roll():
    rng = np.random.default_rng(2)
    f = rng.random((9, 6, 8))
    q = O.stream(f, O.EDGE_PERIODIC)
    for i in range(9):
        np.testing.assert_array_equal(q[i], np.roll(f[i], (-CX[i], CY[i]), axis=(0, 1)))


def test_bounce_back_select_and_involution():
    rng = np.random.default_rng(3)
    f = rng.random((9, 8, 8)).astype(np.float32)
    flags = (rng.random((8, 8)) < 0.3).astype(np.uint8)
    b = O.bounce_back(f, flags)
    for i in range(9):
        np.testing.assert_array_equal(b[i], np.where(flags == 1, f[OPP[i]], f[i]))
    np.testing.assert_array_equal(O.bounce_back(b, flags), f)
    lid = flags * 2  # moving-lid cells are solid in the reference mode too
    np.testing.assert_array_equal(O.bounce_back(f, lid), b)


def test_moments_against_float64_sums():
    rng = np.random.default_rng(4)
    f = (rng.random((9, 10, 12)) * 0.2 + 0.01).astype(np.float32)
    rho = O.density(f)
    mx, my = O.momentum(f)
    f64 = f.astype(np.float64)
    np.testing.assert_allclose(rho, f64.sum(0), rtol=2e-7)
    np.testing.assert_allclose(mx, np.tensordot(CX, f64, 1), atol=2e-7)
    np.testing.assert_allclose(my, np.tensordot(CY, f64, 1), atol=2e-7)
    ux, uy = O.velocity(f)
    inv = np.float32(1) / rho
    np.testing.assert_array_equal(ux, inv * mx)
    np.testing.assert_array_equal(uy, inv * my)
    # exact on dyadic values
    d = np.full((9, 3, 3), 0.125, np.float32)
    assert np.all(O.density(d) == 1.125)


def test_equilibrium_closed_forms():
    p = O.params(5, 4)
    rng = np.random.default_rng(5)
    rho = (1 + 0.1 * rng.random((4, 5))).astype(np.float32)
    zero = np.zeros_like(rho)
    feq = O.equilibrium(p, rho, zero, zero)
    k = O.consts_f32(p)
    for i in range(9):
        np.testing.assert_array_equal(feq[i], rho * np.float32(k.w[i]))
    ux = (0.1 * rng.standard_normal((4, 5))).astype(np.float32)
    uy = (0.1 * rng.standard_normal((4, 5))).astype(np.float32)
    feq = O.equilibrium(p, rho, ux, uy).astype(np.float64)
    v2 = ux.astype(np.float64) ** 2 + uy.astype(np.float64) ** 2
    # sum_i w_i s_i = 1 + v2 (B/3 + C): 1 + ~1e-7 v2 with the f32 constants
    np.testing.assert_allclose(feq.sum(0), rho * (1 + v2 * (k.B / 3 + k.C)), rtol=3e-7)
    np.testing.assert_allclose(np.tensordot(CX, feq, 1), rho * ux * k.A / 3, rtol=2e-6, atol=5e-8)
    np.testing.assert_allclose(np.tensordot(CY, feq, 1), rho * uy * k.A / 3, rtol=2e-6, atol=5e-8)
    # second moment: sum w_i s_i c_x c_y = rho ux uy (B/9 * 2) -> rho ux uy
    pxy = np.tensordot(np.array(CX) * np.array(CY), feq, 1)
    np.testing.assert_allclose(pxy, rho * ux * uy * (2 * k.B / 9), rtol=1e-5, atol=5e-8)


def test_bgk_conserves_and_fixed_point():
    p = O.params(6, 6, tau=0.6)
    rng = np.random.default_rng(6)
    f = (rng.random((9, 6, 6)) * 0.1 + 0.05).astype(np.float64)
    rho = O.density(f)
    ux, uy = O.velocity(f)
    feq = O.equilibrium(p, rho, ux, uy)
    out = O.bgk(p, f, feq)
    np.testing.assert_allclose(out.sum(0), f.sum(0), rtol=1e-13)
    np.testing.assert_allclose(np.tensordot(CX, out, 1), np.tensordot(CX, f, 1), atol=1e-14)
    np.testing.assert_array_equal(O.bgk(p, f, f), f)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
def test_fused_equals_unfused_reference(dtype, edge):
    rng = np.random.default_rng(7)
    h, w = 13, 17
    f = (rng.random((9, h, w)) * 0.1 + 0.05).astype(dtype)
    flags = (rng.random((h, w)) < 0.2).astype(np.uint8)
    p = O.params(w, h, tau=0.7, edge=edge)
    for fl in (None, flags):
        a = O.step(p, f, fl)
        b = O.step_unfused(p, f, fl)
        np.testing.assert_array_equal(a, b)


def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(8)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 1e-2,
        rng.standard_normal(20000).astype(np.float32) * 1e-6,   # half subnormals
        (rng.standard_normal(2000) * 7e4).astype(np.float32),   # overflow
        np.array([0.0, -0.0, 65504.0, 65519.996, 65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                  np.inf, -np.inf], np.float32)])
    with np.errstate(over="ignore"):  # the overflow cases are intended (to +-inf)
        ref = x.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(O.to_f16(x), ref)
    # exact ties at the 13-bit boundary
    t = (np.arange(1, 4000, dtype=np.uint32) << 12 | 0x3C000000).view(np.float32)
    np.testing.assert_array_equal(O.to_f16(t), t.astype(np.float16).view(np.uint16))


def test_bf16_conversion_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.standard_normal(50000).astype(np.float32) * 1e-2,
                        ((np.arange(1, 3000, dtype=np.uint32) << 15) | 0x3F800000).view(np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(O.to_bf16(x), ref)


def widen16(kind, d):
    d = np.asarray(d, np.uint16)
    if kind == O.BF16:
        return (d.astype(np.uint32) << 16).view(np.float32)
    return d.view(np.float16).astype(np.float32)


def test_dev16_roundtrip_and_step():
    p = O.params(8, 6, tau=0.8, edge=O.EDGE_PERIODIC)
    rng = np.random.default_rng(10)
    rho = (1 + 0.01 * rng.standard_normal((6, 8))).astype(np.float32)
    ux = (0.05 * rng.standard_normal((6, 8))).astype(np.float32)
    uy = (0.05 * rng.standard_normal((6, 8))).astype(np.float32)
    f = O.equilibrium(p, rho, ux, uy)
    k = O.consts_f32(p)
    dev = f - np.array(k.base, np.float32)[:, None, None]
    for kind, narrow, tol in ((O.BF16, O.to_bf16, 2 ** -8), (O.F16, O.to_f16, 2 ** -11)):
        d = O.encode_dev16(p, kind, f)
        np.testing.assert_array_equal(d, narrow(dev))
        back = O.decode_dev16(p, kind, d)
        assert np.max(np.abs(back - f)) <= tol * np.max(np.abs(dev)) + 1e-12
        # the 16-bit step is widen -> deviation-form f32 step -> narrow
        np.testing.assert_array_equal(O.run_dev16(p, kind, d, 1),
                                      narrow(O.step_devf32(p, widen16(kind, d))))


@pytest.mark.parametrize("boundary", [O.BB_REFERENCE, O.BB_HALFWAY])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
def test_deviation_form_is_the_same_step_more_precisely(boundary, edge):
    """The deviation-form step computes f' - w_i rho0 of the same update: against
    the FP64 step it is at least as accurate as the full-form FP32 step."""
    rng = np.random.default_rng(21)
    h, w = 14, 18
    rho = 1 + 0.01 * rng.standard_normal((h, w))
    ux = 0.05 * rng.standard_normal((h, w))
    uy = 0.05 * rng.standard_normal((h, w))
    flags = (rng.random((h, w)) < 0.15).astype(np.uint8)
    flags[rng.random((h, w)) < 0.05] = O.LID
    p = O.params(w, h, tau=0.7, edge=edge, boundary=boundary, lid=(0.01, 0.04))
    f64 = O.equilibrium(p, rho, ux, uy) + 1e-3 * rng.standard_normal((9, h, w))
    f32 = f64.astype(np.float32)
    # zero-filled edges drain mass (the reference's af::convolve2 padding), so
    # keep that case short
    steps = 20 if edge == O.EDGE_PERIODIC else 2
    ref = O.run(p, f32.astype(np.float64), steps, flags)
    full = O.run(p, f32, steps, flags).astype(np.float64)
    dev = O.from_dev(p, O.run_devf32(p, O.to_dev(p, f32), steps, flags)).astype(np.float64)
    base = np.array(O.consts_f64(p).base)[:, None, None]
    scale = np.abs(ref - base).max()
    err_full = np.abs(full - ref).max() / scale
    err_dev = np.abs(dev - ref).max() / scale
    assert err_dev < 1e-5 and err_full < 1e-3
    assert err_dev <= err_full * 1.5
