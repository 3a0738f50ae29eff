"""Parity of the HIP step (through the C ABI) with the CPU oracle.

Bar: FP32 and FP64 results equal the oracle's value for value
(np.testing.assert_array_equal); 16-bit storage equals the oracle's 16-bit
restatement bit for bit.  Full-size runs are checked through conservation,
analytic decay and the FP64 GPU run.
"""
import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

STORE = {"f32": np.float32, "f64": np.float64}


def rand_state(rng, h, w, dtype):
    rho = 1 + 0.02 * rng.standard_normal((h, w))
    ux = 0.03 * rng.standard_normal((h, w))
    uy = 0.03 * rng.standard_normal((h, w))
    p = O.params(w, h)
    f = O.equilibrium(p, rho, ux, uy) + 0.002 * rng.standard_normal((9, h, w))
    return f.astype(dtype)


def random_flags(rng, h, w, lids=True):
    fl = (rng.random((h, w)) < 0.15).astype(np.uint8)
    if lids:
        fl[rng.random((h, w)) < 0.05] = O.LID
    return fl


def gpu_params(op: O.Params) -> Params:
    return Params(tau=op.tau, delta_x=op.delta_x, delta_t=op.delta_t, edge=op.edge, boundary=op.boundary,
                  lid_ux=op.lid_ux, lid_uy=op.lid_uy, inlet_ux=op.inlet_ux, inlet_uy=op.inlet_uy,
                  outlet_rho=op.outlet_rho, rho0=op.rho0, collision=op.collision, tau_minus=op.tau_minus,
                  ks_viscosity=op.ks_viscosity)


@pytest.fixture(params=["auto", "one_step", "vector"])
def kernel(request):
    """The step kernel the lattice takes: the built-in choice (FP32 and FP32
    deviation lattices without geometry: the temporally blocked kernel), the
    one-step kernels' choice by size (small test lattices: the scalar kernel),
    or the vector kernel, forced wherever rows split into 16-byte vectors."""
    if request.param == "vector":
        capi.call("cs_set_cells_per_thread", capi.CS_STEP_VECTOR)
    elif request.param == "one_step":
        capi.call("cs_set_temporal_steps", 0)
    try:
        yield request.param
    finally:
        capi.call("cs_set_cells_per_thread", 0)
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)


@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("with_flags", [False, True])
def test_reference_mode_matches_oracle(ctx, storage, edge, with_flags, kernel):
    rng = np.random.default_rng(11)
    h, w = (21, 136) if kernel == "vector" else (21, 137)
    f = rand_state(rng, h, w, STORE[storage])
    flags = random_flags(rng, h, w) if with_flags else None
    op = O.params(w, h, tau=0.6, edge=edge)
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.upload(f)
        ref = f
        for n in (1, 24):
            lat.step(n)
            ref = O.run(op, ref, n, flags)
            np.testing.assert_array_equal(lat.download(STORE[storage]), ref)
        assert lat.time == 25.0


@pytest.mark.parametrize("cpt", [1, 2, 4])
def test_launch_shapes_agree(ctx, cpt):
    rng = np.random.default_rng(12)
    h, w = 9, 300
    f = rand_state(rng, h, w, np.float32)
    op = O.params(w, h, tau=0.7, edge=O.EDGE_PERIODIC)
    try:
        capi.call("cs_set_cells_per_thread", cpt)
        with Lattice(ctx, w, h, "f32", gpu_params(op)) as lat:
            lat.upload(f)
            lat.step(3)
            np.testing.assert_array_equal(lat.download(), O.run(op, f, 3))
    finally:
        capi.call("cs_set_cells_per_thread", 0)


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_unfused_phases(ctx, storage):
    """stream, bounce_back and collide alone (lbm.rs:716-751) compose to one step."""
    rng = np.random.default_rng(13)
    h, w = 17, 23
    dt = STORE[storage]
    f = rand_state(rng, h, w, dt)
    flags = random_flags(rng, h, w)
    op = O.params(w, h, tau=0.9)
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.stream()
        np.testing.assert_array_equal(lat.download(dt), O.stream(f))
        lat.bounce_back()
        b = O.bounce_back(O.stream(f), flags)
        np.testing.assert_array_equal(lat.download(dt), b)
        lat.collide()
        np.testing.assert_array_equal(lat.download(dt), O.step_unfused(op, f, flags))
        np.testing.assert_array_equal(lat.download(dt), O.step(op, f, flags))


def test_reciprocal_equals_ieee_for_every_input(ctx):
    """The kernels' 1 / rho (hardware reciprocal + one FMA Newton step on
    normal |b| < 2^126, IEEE division elsewhere) against IEEE 1.0f / b over
    all 2^32 bit patterns on this GPU: the exhaustive check that the FP32 and
    deviation steps keep the reference's quotient."""
    assert ctx.check_reciprocal() == 0


def test_cavity_128_halfway_lid_10000_steps(ctx):
    """Lid-driven cavity 128x128, tau 0.6, lid 0.1 on row 0, half-way walls,
    f = f_eq(1, 0), 10,000 steps: f, rho and u equal the oracle's."""
    n, steps = 128, 10000
    flags = W.cavity_flags(n, n)
    op = O.params(n, n, tau=0.6, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.1))
    rho, ux, uy = W.rest_state(n, n)
    f0 = O.equilibrium(op, rho, ux, uy)
    ref = O.run(op, f0, steps, flags)
    with Lattice(ctx, n, n, "f32", gpu_params(op), flags) as lat:
        lat.init_equilibrium(rho, ux, uy)
        np.testing.assert_array_equal(lat.download(), f0)
        lat.step(steps)
        got = lat.download()
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(lat.density(), O.density(ref))
        gx, gy = lat.velocity()
        rx, ry = O.velocity(ref)
        np.testing.assert_array_equal(gx, rx)
        np.testing.assert_array_equal(gy, ry)
    # the flow is developed: the lid drags the top row, a vortex forms
    assert ry[1, n // 2] > 0.03 and ry[n - 10, n // 2] < 0


@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
def test_halfway_random_geometry(ctx, storage, edge, kernel):
    rng = np.random.default_rng(14)
    h, w = (33, 72) if kernel == "vector" else (33, 70)
    dt = STORE[storage]
    f = rand_state(rng, h, w, dt)
    flags = random_flags(rng, h, w)
    op = O.params(w, h, tau=0.75, edge=edge, boundary=O.BB_HALFWAY, lid=(0.02, 0.05))
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.step(40)
        np.testing.assert_array_equal(lat.download(dt), O.run(op, f, 40, flags))


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_zou_he_channel(ctx, storage, kernel):
    ch = W.Channel(w=96, h=40, circles=[(20, 30, 7)], re=30.0, u_in=0.05)
    flags = ch.flags()
    op = O.params(ch.w, ch.h, tau=ch.tau, boundary=O.BB_HALFWAY, inlet=ch.inlet(), outlet_rho=1.0)
    dt = STORE[storage]
    rho, ux, uy = (a.astype(dt) for a in W.rest_state(ch.w, ch.h))
    f0 = O.equilibrium(op, rho, ux, uy)
    with Lattice(ctx, ch.w, ch.h, storage, gpu_params(op), flags) as lat:
        lat.init_equilibrium(rho, ux, uy)
        lat.step(300)
        np.testing.assert_array_equal(lat.download(dt), O.run(op, f0, 300, flags))


def narrow16(kind, g):
    return O.to_bf16(g) if kind == O.BF16 else O.to_f16(g)


@pytest.mark.parametrize("kind", ["bf16", "f16"])
@pytest.mark.parametrize("halfway", [False, True])
def test_dev16_matches_oracle(ctx, kind, halfway):
    n, m = 48, 72
    rho, ux, uy = W.taylor_green(n, 0.05, np.float32, m=m)
    okind = O.BF16 if kind == "bf16" else O.F16
    if halfway:
        flags = W.cavity_flags(m, n)
        op = O.params(m, n, tau=0.8, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.08))
    else:
        flags = None
        op = O.params(m, n, tau=0.8, edge=O.EDGE_PERIODIC)
    d0 = narrow16(okind, O.equilibrium_dev(op, rho, ux, uy))
    with Lattice(ctx, m, n, kind, gpu_params(op), flags) as lat:
        lat.init_equilibrium(rho, ux, uy)
        np.testing.assert_array_equal(lat.download_raw(), d0)
        lat.step(20)
        np.testing.assert_array_equal(lat.download_raw(), O.run_dev16(op, okind, d0, 20, flags))
        np.testing.assert_array_equal(lat.download(), O.decode_dev16(op, okind, lat.download_raw()))
        # full populations uploaded are encoded as deviations
        f = O.equilibrium(op, rho, ux, uy)
        lat.upload(f)
        np.testing.assert_array_equal(lat.download_raw(), O.encode_dev16(op, okind, f))


@pytest.mark.parametrize("kind", ["bf16", "f16"])
def test_dev16_init_rounds_once_to_f32_then_16bit(ctx, kind):
    """Random (rho, u) give ~2^-13 of the values on a rounding tie; the stored
    16-bit value must be the f32 equilibrium rounded to nearest even (a fused
    multiply-convert would round once and differ on those ties)."""
    rng = np.random.default_rng(19)
    h, w = 200, 300
    rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
    ux = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
    op = O.params(w, h, tau=0.7)
    okind = O.BF16 if kind == "bf16" else O.F16
    with Lattice(ctx, w, h, kind, gpu_params(op)) as lat:
        lat.init_equilibrium(rho, ux, uy)
        np.testing.assert_array_equal(lat.download_raw(), narrow16(okind, O.equilibrium_dev(op, rho, ux, uy)))


@pytest.mark.parametrize("kind", ["bf16", "f16"])
@pytest.mark.parametrize("halfway", [False, True])
def test_dev16_stochastic_rounding_matches_oracle(ctx, kind, halfway):
    n, m = 40, 64
    rho, ux, uy = W.taylor_green(n, 0.05, np.float32, m=m)
    okind = O.BF16 if kind == "bf16" else O.F16
    if halfway:
        flags = W.cavity_flags(m, n)
        op = O.params(m, n, tau=0.8, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.08))
    else:
        flags = None
        op = O.params(m, n, tau=0.8, edge=O.EDGE_PERIODIC)
    seed = 0xC0FFEE
    gp = gpu_params(op)
    gp.rounding = capi.CS_ROUND_STOCHASTIC
    gp.seed = seed
    d0 = narrow16(okind, O.equilibrium_dev(op, rho, ux, uy))
    with Lattice(ctx, m, n, kind, gp, flags) as lat:
        lat.init_equilibrium(rho, ux, uy)
        lat.step(15)
        d15 = O.run_dev16(op, okind, d0, 15, flags, rounding=1, seed=seed)
        np.testing.assert_array_equal(lat.download_raw(), d15)
        lat.step(5)
        np.testing.assert_array_equal(lat.download_raw(),
                                      O.run_dev16(op, okind, d15, 5, flags, rounding=1, seed=seed, step0=15))
        # stochastic and nearest rounding differ
        assert not np.array_equal(d15, O.run_dev16(op, okind, d0, 15, flags))


@pytest.mark.parametrize("case", ["ref_zero", "ref_periodic", "halfway", "channel"])
def test_f32dev_matches_oracle(ctx, case, kernel):
    rng = np.random.default_rng(18)
    if case == "channel":
        ch = W.Channel(w=80, h=30, circles=[(15, 25, 5)], re=30.0, u_in=0.05)
        flags = ch.flags()
        op = O.params(ch.w, ch.h, tau=ch.tau, boundary=O.BB_HALFWAY, inlet=ch.inlet(), outlet_rho=1.0)
        h, w = ch.h, ch.w
    else:
        h, w = (27, 44) if kernel == "vector" else (27, 45)
        flags = random_flags(rng, h, w)
        op = O.params(w, h, tau=0.7, edge=O.EDGE_PERIODIC if case == "ref_periodic" else O.EDGE_ZERO,
                      boundary=O.BB_HALFWAY if case == "halfway" else O.BB_REFERENCE, lid=(0.0, 0.06))
    rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
    ux = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    g0 = O.equilibrium_dev(op, rho, ux, uy)
    steps = 3 if case == "ref_zero" else 60
    with Lattice(ctx, w, h, "f32dev", gpu_params(op), flags) as lat:
        lat.init_equilibrium(rho, ux, uy)
        np.testing.assert_array_equal(lat.download_raw(), g0)
        lat.step(steps)
        g = O.run_devf32(op, g0, steps, flags)
        np.testing.assert_array_equal(lat.download_raw(), g)
        np.testing.assert_array_equal(lat.download(), O.from_dev(op, g))
        r, vx, vy = lat.moments_f64()
        k = O.consts_f32(op)
        drho = g.sum(0, dtype=np.float32)
        np.testing.assert_allclose(r, 1.0 + drho.astype(np.float64), rtol=1e-6)


def test_moment_api(ctx):
    rng = np.random.default_rng(15)
    h, w = 19, 31
    f = rand_state(rng, h, w, np.float32)
    op = O.params(w, h, tau=0.6)
    with Lattice(ctx, w, h, "f32", gpu_params(op)) as lat:
        lat.upload(f)
        np.testing.assert_array_equal(lat.density(), O.density(f))
        mx, my = lat.momentum_density()
        rx, ry = O.momentum(f)
        np.testing.assert_array_equal(mx, rx)
        np.testing.assert_array_equal(my, ry)
        ux, uy = lat.velocity()
        vx, vy = O.velocity(f)
        np.testing.assert_array_equal(ux, vx)
        np.testing.assert_array_equal(uy, vy)
        np.testing.assert_array_equal(lat.speed(), np.sqrt(vx * vx + vy * vy))
        cs2 = np.float32(O.consts_f32(op).cs2)
        np.testing.assert_array_equal(lat.pressure(), O.density(f) * cs2)
        feq = O.equilibrium(op, O.density(f), vx, vy)
        np.testing.assert_array_equal(lat.equilibrium(), feq)
        np.testing.assert_array_equal(lat.non_equilibrium(), f - feq)
        assert not lat.is_unstable()
        # |u|^2 > 2/3 makes f_eq[0] negative (lbm.rs:815-818)
        lat.init_equilibrium(np.ones((h, w), np.float32), np.full((h, w), 0.9, np.float32),
                             np.zeros((h, w), np.float32))
        assert lat.is_unstable()


def test_sums_and_residual(ctx):
    rng = np.random.default_rng(16)
    h, w = 40, 50
    f = rand_state(rng, h, w, np.float32)
    flags = random_flags(rng, h, w, lids=False)
    op = O.params(w, h, tau=0.8, boundary=O.BB_HALFWAY)
    with Lattice(ctx, w, h, "f32", gpu_params(op), flags) as lat:
        lat.upload(f)
        np.testing.assert_allclose(lat.sums_f64(), O.sums(f, flags), rtol=1e-12, atol=1e-12)
        lat.snapshot_velocity()
        u0 = O.velocity(f)
        lat.step(5)
        g = lat.download()
        u1 = O.velocity(g)
        np.testing.assert_allclose(lat.residual_f64(), O.residual(*u1, *u0, flags), rtol=1e-12)


def test_geometry_edit_between_steps(ctx):
    """The mask may change at any step boundary (main.rs:77-89)."""
    rng = np.random.default_rng(17)
    h, w = 24, 24
    f = rand_state(rng, h, w, np.float32)
    fl1 = random_flags(rng, h, w, lids=False)
    fl2 = random_flags(rng, h, w, lids=False)
    op = O.params(w, h, tau=0.7)
    with Lattice(ctx, w, h, "f32", gpu_params(op), fl1) as lat:
        lat.upload(f)
        lat.step(3)
        lat.set_flags(fl2)
        lat.step(3)
        lat.set_flags(None)
        lat.step(2)
        ref = O.run(op, O.run(op, O.run(op, f, 3, fl1), 3, fl2), 2, None)
        np.testing.assert_array_equal(lat.download(), ref)


@pytest.mark.parametrize("storage,tol", [("f32dev", 1e-6), ("f32", 2e-5)])
def test_run_to_residual_converges(ctx, storage, tol):
    """Steady lid-driven cavity.  Rounding noise bounds the reachable residual:
    about 6e-6 in the reference's full-population FP32 form, about 4e-7 in
    deviation arithmetic."""
    n = 32
    flags = W.cavity_flags(n, n)
    op = O.params(n, n, tau=0.8, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.05))
    with Lattice(ctx, n, n, storage, gpu_params(op), flags) as lat:
        lat.init_equilibrium(*W.rest_state(n, n))
        steps, res = lat.run_to_residual(tol, 100, 200000)
        assert res <= tol and 0 < steps < 200000
        lat.snapshot_velocity()
        lat.step(100)
        assert lat.residual_f64() < 2 * tol


def test_taylor_green_4096_fp32_vs_fp64(ctx):
    """Full size: TGV 4096^2, tau 0.8, u0 0.05, 1,000 steps; FP32 (reference
    arithmetic) and FP32 deviation arithmetic against the FP64 GPU run, analytic
    decay and FP64 mass conservation."""
    n, steps, tau, u0 = 4096, 1000, 0.8, 0.05
    rho, ux, uy = W.taylor_green(n, u0, np.float64)
    p = Params(tau=tau, edge=capi.CS_EDGE_PERIODIC)
    res = {}
    for storage in ("f64", "f32", "f32dev"):
        with Lattice(ctx, n, n, storage, p) as lat:
            lat.init_equilibrium(*(a if storage == "f64" else a.astype(np.float32) for a in (rho, ux, uy)))
            s0 = lat.sums_f64()
            lat.step(steps)
            s1 = lat.sums_f64()
            res[storage] = (lat.moments_f64(), abs(s1[0] - s0[0]) / s0[0])
    (r64, x64, y64), drift64 = res["f64"]
    assert drift64 < 1e-12
    amp = ((x64 * ux + y64 * uy).sum() / (ux * ux + uy * uy).sum())
    # within 1% of the total decay (1 - 0.99953): f = f_eq initialisation leaves a
    # small initial-layer offset
    dec = W.taylor_green_decay(n, tau, steps)
    assert abs(amp - dec) < 0.01 * (1 - dec)
    den = np.sqrt((x64 ** 2 + y64 ** 2).sum())
    err = {}
    for st in ("f32", "f32dev"):
        (r, x, y), drift = res[st]
        err[st] = np.sqrt(((x - x64) ** 2 + (y - y64) ** 2).sum()) / den
        print(f"{st}: u rel L2 vs f64 {err[st]:.3e}, mass drift {drift:.3e}, max |drho| {np.abs(r - r64).max():.3e}")
    # The reference's FP32 constants (1/cs^2 = 3.0000002 etc.) add ~1e-7 relative
    # momentum per collision, ~1e-4 in u after 1,000 steps, and its full-
    # population rounding drifts the mass by ~1e-5.  Deviation arithmetic with
    # the once-rounded constants keeps both near FP32 rounding.
    assert err["f32"] < 5e-4 and res["f32"][1] < 5e-5
    assert err["f32dev"] < 3e-5 and res["f32dev"][1] < 1e-6
    assert err["f32dev"] < err["f32"]


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_taylor_green_8192_matches_oracle(ctx, storage):
    """The largest benchmark lattice, 8192^2, two steps value for value with
    the oracle: FP32 runs the vector kernel with 32-bit offsets (2.4 GB per
    buffer), FP64 the scalar kernel with 64-bit offsets (4.8 GB per buffer)."""
    n, steps = 8192, 2
    dt = STORE[storage]
    print(f"[8192] {storage}: {9 * n * n * np.dtype(dt).itemsize / 1e9:.1f} GB per buffer", flush=True)
    rho, ux, uy = W.taylor_green(n, 0.05, dt)
    op = O.params(n, n, tau=0.8, edge=O.EDGE_PERIODIC)
    f0 = O.equilibrium(op, rho, ux, uy)
    del rho, ux, uy
    with Lattice(ctx, n, n, storage, gpu_params(op)) as lat:
        lat.upload(f0)
        lat.step(steps)
        got = lat.download(dt)
    if storage == "f32":
        ref, _ = O.run_omp(op, f0, steps)
    else:
        ref = O.run(op, f0, steps)
    del f0
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("storage", ["f32", "f64", "f32dev", "f16", "bf16"])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
def test_init_nonequilibrium_matches_oracle(ctx, storage, edge):
    """Consistent start f_eq + f_neq on random walls and lids: value for value
    (16-bit: the f32 deviation rounded to nearest even)."""
    rng = np.random.default_rng(23)
    h, w = 45, 70
    rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
    ux = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    flags = random_flags(rng, h, w)
    op = O.params(w, h, tau=0.7, edge=edge, boundary=O.BB_HALFWAY, lid=(0.01, -0.02))
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.init_nonequilibrium(rho, ux, uy)
        if storage in ("f32", "f64"):
            dt = STORE[storage]
            ref = O.nonequilibrium_init(op, *(a.astype(dt) for a in (rho, ux, uy)), flags)
            np.testing.assert_array_equal(lat.download(dt), ref)
        else:
            g = O.nonequilibrium_init(op, rho, ux, uy, flags, dev=True)
            if storage == "f32dev":
                np.testing.assert_array_equal(lat.download_raw(), g)
            else:
                np.testing.assert_array_equal(lat.download_raw(), narrow16(O.F16 if storage == "f16" else O.BF16, g))


@pytest.mark.parametrize("storage", ["f32", "f64", "f32dev", "f16", "bf16", "f16_sr", "bf16_sr"])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("shape", [(16, 5), (264, 7), (520, 3), (1040, 9), (2048, 16)])
def test_vector_fluid_step_matches_oracle(ctx, storage, edge, shape):
    """The vector kernel (forced with CS_STEP_VECTOR, and CS_STEP_VECTOR_WIDE)
    on lattices without geometry whose rows split into vectors (16-byte lanes
    for 4- and 8-byte storage; 8- and 16-byte lanes for 16-bit storage; DPP
    lane shifts for the row-wise streaming); partial waves, one- and many-wave
    rows, row counts that do and do not fill a block: value for value with the
    oracle, and equal to the scalar kernel."""
    w, h = shape
    kind = storage.split("_")[0]
    sr = storage.endswith("_sr")
    rng = np.random.default_rng(w * 7 + h)
    rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
    ux = (0.04 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.04 * rng.standard_normal((h, w))).astype(np.float32)
    op = O.params(w, h, tau=0.7, edge=edge)
    gp = gpu_params(op)
    if sr:
        gp.rounding = capi.CS_ROUND_STOCHASTIC
        gp.seed = 99
    steps = 4
    if kind in ("f32", "f64"):
        dt = STORE[kind]
        f0 = O.equilibrium(op, *(a.astype(dt) for a in (rho, ux, uy)))
        ref = O.run(op, f0, steps)
    elif kind == "f32dev":
        f0 = O.equilibrium_dev(op, rho, ux, uy)
        ref = O.run_devf32(op, f0, steps)
    else:
        okind = O.BF16 if kind == "bf16" else O.F16
        f0 = narrow16(okind, O.equilibrium_dev(op, rho, ux, uy))
        ref = O.run_dev16(op, okind, f0, steps, None, rounding=1 if sr else 0, seed=99)
    outs = []
    for cpt in (capi.CS_STEP_VECTOR, capi.CS_STEP_VECTOR_WIDE, 1):
        try:
            capi.call("cs_set_cells_per_thread", cpt)
            with Lattice(ctx, w, h, kind, gp) as lat:
                lat.upload_raw(f0)
                lat.step(steps)
                outs.append(lat.download_raw())
        finally:
            capi.call("cs_set_cells_per_thread", 0)
    for o in outs:
        np.testing.assert_array_equal(o, ref)



@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("case", ["fluid_vector", "fluid_ragged", "reference_mask", "halfway_lid"])
def test_trt_matches_oracle(ctx, storage, edge, case, kernel):
    """TRT (lbm.rs:374-451, with its swap_equilibrium) in the fused step, every
    kernel path: value for value with the oracle."""
    rng = np.random.default_rng(31)
    h, w = (12, 264) if case == "fluid_vector" else (11, 36 if kernel == "vector" and case != "fluid_ragged" else 37)
    dt = STORE[storage]
    f = rand_state(rng, h, w, dt)
    flags = None
    kw = {}
    if case == "reference_mask":
        flags = random_flags(rng, h, w, lids=False)
    elif case == "halfway_lid":
        flags = random_flags(rng, h, w)
        kw = dict(boundary=O.BB_HALFWAY, lid=(0.02, -0.03))
    tp, tm = O.trt_new(0.25, 0.08)
    op = O.params(w, h, edge=edge, trt=(tp, tm), **kw)
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.step(6)
        np.testing.assert_array_equal(lat.download(dt), O.run(op, f, 6, flags))


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_trt_unfused_phases(ctx, storage):
    rng = np.random.default_rng(32)
    h, w = 10, 21
    dt = STORE[storage]
    f = rand_state(rng, h, w, dt)
    flags = random_flags(rng, h, w, lids=False)
    op = O.params(w, h, trt=(0.9, 0.6))
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.stream()
        lat.bounce_back()
        q = O.bounce_back(O.stream(f, O.EDGE_ZERO), flags)
        np.testing.assert_array_equal(lat.download(dt), q)
        lat.collide()
        np.testing.assert_array_equal(lat.download(dt), O.trt(op, q))


def test_trt_new_and_storage_check(ctx):
    from chemsim_amd.lattice import trt_new
    for lam, nu in ((0.25, 10.0), (3.0 / 16, 0.1)):
        assert trt_new(lam, nu) == O.trt_new(lam, nu)
    op = O.params(16, 8, trt=(0.9, 0.6))
    with Lattice(ctx, 16, 8, "f16") as lat:
        with pytest.raises(capi.ChemsimError):
            lat.set_params(gpu_params(op))


OTHER = {"regularized": dict(regularized=True, tau=0.8), "kbc": dict(kbc=0.05)}


@pytest.mark.parametrize("coll", ["regularized", "kbc"])
@pytest.mark.parametrize("storage", ["f32", "f64"])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("case", ["fluid_vector", "fluid_ragged", "reference_mask", "halfway_lid"])
def test_regularized_kbc_match_oracle(ctx, coll, storage, edge, case, kernel):
    """Regularized (lbm.rs:596-666) and KBC (lbm.rs:458-590) in the fused step,
    every kernel path: value for value with the oracle."""
    rng = np.random.default_rng(41)
    h, w = (12, 264) if case == "fluid_vector" else (11, 36 if kernel == "vector" and case != "fluid_ragged" else 37)
    dt = STORE[storage]
    f = rand_state(rng, h, w, dt)
    flags = None
    kw = dict(OTHER[coll])
    if case == "reference_mask":
        flags = random_flags(rng, h, w, lids=False)
    elif case == "halfway_lid":
        flags = random_flags(rng, h, w)
        kw.update(boundary=O.BB_HALFWAY, lid=(0.02, -0.03))
    op = O.params(w, h, edge=edge, **kw)
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.step(6)
        np.testing.assert_array_equal(lat.download(dt), O.run(op, f, 6, flags))


@pytest.mark.parametrize("coll", ["regularized", "kbc"])
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_regularized_kbc_unfused_phases(ctx, coll, storage):
    rng = np.random.default_rng(42)
    h, w = 10, 21
    dt = STORE[storage]
    f = rand_state(rng, h, w, dt)
    flags = random_flags(rng, h, w, lids=False)
    op = O.params(w, h, **OTHER[coll])
    with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.stream()
        lat.bounce_back()
        q = O.bounce_back(O.stream(f, O.EDGE_ZERO), flags)
        np.testing.assert_array_equal(lat.download(dt), q)
        lat.collide()
        want = O.regularized(op, q) if coll == "regularized" else O.kbc(op, q)
        np.testing.assert_array_equal(lat.download(dt), want)


def test_regularized_kbc_parameter_checks(ctx):
    op = O.params(16, 8, regularized=True)
    with Lattice(ctx, 16, 8, "bf16") as lat:
        with pytest.raises(capi.ChemsimError):
            lat.set_params(gpu_params(op))
    with Lattice(ctx, 16, 8, "f32") as lat:
        with pytest.raises(capi.ChemsimError):   # KBC needs a positive viscosity
            lat.set_params(gpu_params(O.params(16, 8, kbc=0.0)))
        lat.set_params(gpu_params(O.params(16, 8, kbc=0.01)))


@pytest.mark.parametrize("coll", ["regularized", "kbc"])
def test_regularized_kbc_cavity_full_size(ctx, coll):
    """A lid-driven cavity 512 x 512 for 200 steps: the FP32 run equals the
    oracle's value for value."""
    n, steps = 512, 200
    flags = W.cavity_flags(n, n)
    kw = dict(OTHER[coll])
    op = O.params(n, n, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.05), **kw)
    rho, ux, uy = W.rest_state(n, n)
    f0 = O.equilibrium(op, rho, ux, uy).astype(np.float32)
    with Lattice(ctx, n, n, "f32", gpu_params(op), flags) as lat:
        lat.upload(f0)
        lat.step(steps)
        got = lat.download(np.float32)
    ref, _ = O.run_omp(op, f0, steps, flags)
    np.testing.assert_array_equal(got, ref)
    assert np.isfinite(got).all()


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (7, 1), (2, 3), (5, 4), (3, 65)])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_degenerate_shapes_match_oracle(ctx, shape, edge, storage):
    """Tiny and one-cell-wide lattices (every neighbour outside or wrapped onto
    the cell itself), with and without a mask: value for value with the oracle."""
    h, w = shape
    rng = np.random.default_rng(h * 100 + w)
    dt = STORE[storage]
    for with_flags in (False, True):
        f = rand_state(rng, h, w, dt)
        flags = (rng.random((h, w)) < 0.4).astype(np.uint8) if with_flags else None
        op = O.params(w, h, tau=0.7, edge=edge)
        with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
            lat.upload(f)
            lat.step(3)
            np.testing.assert_array_equal(lat.download(dt), O.run(op, f, 3, flags),
                                          err_msg=f"{shape} flags={with_flags}")


@pytest.mark.parametrize("storage", ["f32", "f64", "f32dev", "f16", "bf16"])
@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("halfway", [False, True])
@pytest.mark.parametrize("shape", [(4, 5), (8, 3), (260, 9), (72, 33)])
def test_vector_geometry_matches_oracle(ctx, storage, edge, halfway, shape):
    """The vector kernel (forced with CS_STEP_VECTOR) on lattices with
    geometry whose width splits into vectors: value for value the oracle's run,
    and the scalar kernel's, in the reference and half-way modes."""
    w, h = shape
    rng = np.random.default_rng(w * 131 + h)
    flags = random_flags(rng, h, w, lids=halfway)
    rho = (1 + 0.02 * rng.standard_normal((h, w))).astype(np.float32)
    ux = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    op = O.params(w, h, tau=0.7, edge=edge, boundary=O.BB_HALFWAY if halfway else O.BB_REFERENCE,
                  lid=(0.01, 0.04) if halfway else (0.0, 0.0))
    steps = 6
    if storage == "f32dev":
        f0 = O.equilibrium_dev(op, rho, ux, uy)
        ref = O.run_devf32(op, f0, steps, flags)
    elif storage in ("f16", "bf16"):
        okind = O.BF16 if storage == "bf16" else O.F16
        f0 = narrow16(okind, O.equilibrium_dev(op, rho, ux, uy))
        ref = O.run_dev16(op, okind, f0, steps, flags)
    else:
        dt = STORE[storage]
        f0 = O.equilibrium(op, *(a.astype(dt) for a in (rho, ux, uy)))
        ref = O.run(op, f0, steps, flags)
    print(f"[vector geometry] {storage} {w}x{h} edge={edge} halfway={halfway}", flush=True)
    outs = []
    for cpt in (capi.CS_STEP_VECTOR, 1):
        try:
            capi.call("cs_set_cells_per_thread", cpt)
            with Lattice(ctx, w, h, storage, gpu_params(op), flags) as lat:
                lat.upload_raw(f0)
                lat.step(steps)
                outs.append(lat.download_raw())
        finally:
            capi.call("cs_set_cells_per_thread", 0)
    np.testing.assert_array_equal(outs[0], ref)
    np.testing.assert_array_equal(outs[1], ref)


@pytest.mark.parametrize("n", [2048, 4096])
def test_cavity_full_size_matches_oracle(ctx, n):
    """A lid-driven cavity at 2048^2 (the scalar kernel by the size choice) and
    4096^2 (the vector kernel with geometry), 30 steps from a perturbed rest
    state: the FP32 run equals the oracle's value for value."""
    steps = 30
    flags = W.cavity_flags(n, n)
    op = O.params(n, n, tau=0.6, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.1))
    rng = np.random.default_rng(n)
    rho = (1 + 0.001 * rng.standard_normal((n, n))).astype(np.float32)
    ux = (0.001 * rng.standard_normal((n, n))).astype(np.float32)
    uy = (0.001 * rng.standard_normal((n, n))).astype(np.float32)
    f0 = O.equilibrium(op, rho, ux, uy).astype(np.float32)
    del rho, ux, uy
    print(f"[cavity full size] {n}x{n} {steps} steps", flush=True)
    with Lattice(ctx, n, n, "f32", gpu_params(op), flags) as lat:
        lat.upload(f0)
        lat.step(steps)
        got = lat.download(np.float32)
    ref, _ = O.run_omp(op, f0, steps, flags)
    np.testing.assert_array_equal(got, ref)
