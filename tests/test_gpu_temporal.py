"""The temporally blocked steps (K steps per launch: k_step_tbr with 64 x 64
regions in registers, k_step_tb with 64 x 64 or 32 x 32 regions in LDS)
against the CPU oracle, value for value, for FP32 and FP32 deviation
storage: zero and periodic edges, every collision operator, geometry (the
reference's on-node swap, half-way walls with a moving lid, Zou-He inlet and
outlet), lattices smaller than one region and ragged ones, step counts that
leave a remainder for the one-step kernels, graph replay, and a full-size
lattice against the one-step kernels."""
import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def rand_state(rng, h, w):
    rho = 1 + 0.02 * rng.standard_normal((h, w))
    ux = 0.03 * rng.standard_normal((h, w))
    uy = 0.03 * rng.standard_normal((h, w))
    p = O.params(w, h)
    f = O.equilibrium(p, rho, ux, uy) + 0.002 * rng.standard_normal((9, h, w))
    return f.astype(np.float32)


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


@pytest.fixture(params=[4, 8, "small"])
def steps_per_launch(request):
    """4 or 8: forces the 64 x 64 temporally blocked kernels; "small": the
    built-in choice, which takes 8 steps on 32 x 32 regions for lattices of
    at most 1024 of them -- every lattice of these tests (KBC keeps the
    one-step kernels)."""
    k = capi.CS_TEMPORAL_AUTO if request.param == "small" else request.param
    capi.call("cs_set_temporal_steps", k)
    try:
        yield request.param
    finally:
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)


SHAPES = [(1, 1), (37, 29), (200, 130), (96, 48), (13, 300), (300, 7)]


@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("shape", SHAPES, ids=[f"{w}x{h}" for w, h in SHAPES])
def test_bgk_f32_matches_oracle(ctx, edge, shape, steps_per_launch):
    w, h = shape
    rng = np.random.default_rng(41)
    f = rand_state(rng, h, w)
    op = O.params(w, h, tau=0.62, edge=edge)
    with Lattice(ctx, w, h, "f32", gpu_params(op)) as lat:
        lat.upload(f)
        ref = f
        # 8-step launches, then a remainder of 3 one step at a time (19), of 5 as a
        # 4-step launch and one step (13)
        for n in (19, 8, 13):
            lat.step(n)
            ref = O.run(op, ref, n)
            np.testing.assert_array_equal(lat.download(), ref)
        assert lat.time == 40.0


@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("shape", [(37, 29), (200, 130)], ids=["37x29", "200x130"])
def test_bgk_f32dev_matches_oracle(ctx, edge, shape, steps_per_launch):
    w, h = shape
    rng = np.random.default_rng(42)
    op = O.params(w, h, tau=0.7, edge=edge)
    rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
    ux = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.03 * rng.standard_normal((h, w))).astype(np.float32)
    g0 = O.equilibrium_dev(op, rho, ux, uy)
    with Lattice(ctx, w, h, "f32dev", gpu_params(op)) as lat:
        lat.init_equilibrium(rho, ux, uy)
        lat.step(21)
        np.testing.assert_array_equal(lat.download_raw(), O.run_devf32(op, g0, 21))


OPERATORS = {"trt": dict(trt=(0.9, 0.6)), "regularized": dict(regularized=True, tau=0.8),
             "kbc": dict(kbc=0.05, tau=0.8)}


@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("coll", list(OPERATORS))
def test_operators_match_oracle(ctx, edge, coll, steps_per_launch):
    w, h = 70, 45
    rng = np.random.default_rng(43)
    f = rand_state(rng, h, w)
    op = O.params(w, h, edge=edge, **OPERATORS[coll])
    with Lattice(ctx, w, h, "f32", gpu_params(op)) as lat:
        lat.upload(f)
        lat.step(16)
        np.testing.assert_array_equal(lat.download(), O.run(op, f, 16))


def test_graph_replay(ctx, steps_per_launch):
    """Launch-bound lattices replay captured runs of 64 steps, here eight
    K = 8 launches each, then take a remainder."""
    w, h = 128, 96
    rng = np.random.default_rng(44)
    f = rand_state(rng, h, w)
    op = O.params(w, h, tau=0.8, edge=O.EDGE_PERIODIC)
    with Lattice(ctx, w, h, "f32", gpu_params(op)) as lat:
        lat.upload(f)
        lat.step(64 * 3 + 13)
        np.testing.assert_array_equal(lat.download(), O.run(op, f, 64 * 3 + 13))
        assert lat.time == 64 * 3 + 13


def test_switching_between_kernels(ctx):
    """Turning the temporal blocking off and on between calls continues the
    same trajectory (the one-step and K-step launches share the buffers)."""
    w, h = 150, 90
    rng = np.random.default_rng(45)
    f = rand_state(rng, h, w)
    op = O.params(w, h, tau=0.75, edge=O.EDGE_PERIODIC)
    try:
        with Lattice(ctx, w, h, "f32", gpu_params(op)) as lat:
            lat.upload(f)
            total = 0
            for k, n in ((8, 17), (0, 5), (4, 9), (8, 64), (capi.CS_TEMPORAL_AUTO, 11)):
                capi.call("cs_set_temporal_steps", k)
                lat.step(n)
                total += n
            np.testing.assert_array_equal(lat.download(), O.run(op, f, total))
    finally:
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)


def test_full_size_equals_one_step_kernels(ctx):
    """4096^2 periodic Taylor-Green (the benchmark lattice), 16 steps: the
    temporally blocked launches equal the one-step vector kernel bit for bit."""
    n = 4096
    rho, ux, uy = W.taylor_green(n, 0.05, np.float32)
    out = {}
    try:
        for k in (8, 0):
            capi.call("cs_set_temporal_steps", k)
            with Lattice(ctx, n, n, "f32", Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC)) as lat:
                lat.init_equilibrium(rho, ux, uy)
                lat.step(16)
                out[k] = lat.download_raw()
    finally:
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)
    np.testing.assert_array_equal(out[8].view(np.uint32), out[0].view(np.uint32))


GEOM_SHAPES = [(37, 29), (200, 130), (13, 300)]


@pytest.mark.parametrize("edge", [O.EDGE_ZERO, O.EDGE_PERIODIC])
@pytest.mark.parametrize("mode", ["reference", "halfway"])
@pytest.mark.parametrize("shape", GEOM_SHAPES, ids=[f"{w}x{h}" for w, h in GEOM_SHAPES])
def test_geometry_f32_matches_oracle(ctx, edge, mode, shape, steps_per_launch):
    w, h = shape
    rng = np.random.default_rng(46)
    f = rand_state(rng, h, w)
    flags = random_flags(rng, h, w)
    kw = dict(boundary=O.BB_HALFWAY, lid=(0.02, -0.03)) if mode == "halfway" else {}
    op = O.params(w, h, tau=0.7, edge=edge, **kw)
    with Lattice(ctx, w, h, "f32", gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.step(19)
        np.testing.assert_array_equal(lat.download(), O.run(op, f, 19, flags))


@pytest.mark.parametrize("coll", list(OPERATORS))
def test_operators_with_walls_match_oracle(ctx, coll, steps_per_launch):
    w, h = 70, 45
    rng = np.random.default_rng(47)
    f = rand_state(rng, h, w)
    flags = random_flags(rng, h, w)
    op = O.params(w, h, edge=O.EDGE_PERIODIC, boundary=O.BB_HALFWAY, lid=(0.0, 0.04), **OPERATORS[coll])
    with Lattice(ctx, w, h, "f32", gpu_params(op), flags) as lat:
        lat.upload(f)
        lat.step(16)
        np.testing.assert_array_equal(lat.download(), O.run(op, f, 16, flags))


@pytest.mark.parametrize("storage", ["f32", "f32dev"])
def test_zou_he_channel_matches_oracle(ctx, storage, steps_per_launch):
    """Channel with a velocity inlet, a density outlet, half-way walls and an
    obstacle (the warm start's lattices, smaller)."""
    ch = W.Channel(w=150, h=40, circles=[(30, 20, 6)], re=30.0, u_in=0.05)
    flags = ch.flags()
    op = O.params(ch.w, ch.h, tau=ch.tau, boundary=O.BB_HALFWAY, inlet=ch.inlet(), outlet_rho=1.0)
    rho, ux, uy = W.rest_state(ch.w, ch.h)
    with Lattice(ctx, ch.w, ch.h, storage, gpu_params(op), flags) as lat:
        lat.init_equilibrium(rho, ux, uy)
        lat.step(45)
        if storage == "f32":
            f0 = O.equilibrium(op, rho, ux, uy).astype(np.float32)
            np.testing.assert_array_equal(lat.download(), O.run(op, f0, 45, flags))
        else:
            g0 = O.equilibrium_dev(op, rho.astype(np.float32), ux.astype(np.float32), uy.astype(np.float32))
            np.testing.assert_array_equal(lat.download_raw(), O.run_devf32(op, g0, 45, flags))


def test_cavity_graph_replay(ctx, steps_per_launch):
    """128^2 lid-driven cavity (half-way walls, moving lid): replayed 64-step
    graphs of 8-step launches, then a remainder, against the oracle."""
    n = 128
    flags = W.cavity_flags(n, n)
    op = O.params(n, n, tau=0.6, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.1))
    rho, ux, uy = W.rest_state(n, n)
    f0 = O.equilibrium(op, rho, ux, uy).astype(np.float32)
    with Lattice(ctx, n, n, "f32", gpu_params(op), flags) as lat:
        lat.init_equilibrium(rho, ux, uy)
        lat.step(64 * 4 + 7)
        np.testing.assert_array_equal(lat.download(), O.run(op, f0, 64 * 4 + 7, flags))


@pytest.mark.parametrize("shape", [(1000, 700), (512, 512), (520, 520)], ids=["1000x700", "512x512", "520x520"])
def test_built_in_choice(ctx, shape):
    """The built-in choice: 1000 x 700 and 520 x 520 (more than 1024 regions
    of 32 x 32) take 64 x 64 regions in registers, 512 x 512 (1024) takes
    32 x 32 regions in LDS; each equals the one-step kernels bit for bit."""
    w, h = shape
    rng = np.random.default_rng(48)
    f = rand_state(rng, h, w)
    op = O.params(w, h, tau=0.8, edge=O.EDGE_PERIODIC)
    out = {}
    try:
        for k in (capi.CS_TEMPORAL_AUTO, 0):
            capi.call("cs_set_temporal_steps", k)
            with Lattice(ctx, w, h, "f32", gpu_params(op)) as lat:
                lat.upload(f)
                lat.step(24)
                assert lat.temporal_steps() == (8 if k == capi.CS_TEMPORAL_AUTO else 1)
                out[k] = lat.download()
    finally:
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)
    np.testing.assert_array_equal(out[capi.CS_TEMPORAL_AUTO], out[0])
    np.testing.assert_array_equal(out[0], O.run(op, f, 24))


def test_temporal_steps_query(ctx):
    """cs_lattice_temporal_steps reports the kernel cs_lattice_step takes."""
    big = Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC)
    kbc = Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC, collision=capi.CS_COLL_KBC, ks_viscosity=0.1)
    cav = Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_uy=0.1)
    try:
        with Lattice(ctx, 1000, 700, "f32", big) as a, Lattice(ctx, 1000, 700, "f32", kbc) as b, \
                Lattice(ctx, 128, 128, "f32", cav, W.cavity_flags(128, 128)) as c, \
                Lattice(ctx, 1000, 700, "f64", big) as d, Lattice(ctx, 1000, 700, "f32dev", big) as e:
            assert [x.temporal_steps() for x in (a, b, c, d, e)] == [8, 1, 8, 1, 8]
            capi.call("cs_set_temporal_steps", 4)
            assert [x.temporal_steps() for x in (a, b, c, d, e)] == [4, 8, 4, 1, 4]
            capi.call("cs_set_temporal_steps", 0)
            assert [x.temporal_steps() for x in (a, b, c, d, e)] == [1, 1, 1, 1, 1]
            capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)
            capi.call("cs_set_cells_per_thread", 1)
            assert a.temporal_steps() == 1
    finally:
        capi.call("cs_set_cells_per_thread", 0)
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)
