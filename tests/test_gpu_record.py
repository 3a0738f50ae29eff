"""Frame recording (src/display.rs:157-185) through cs_lattice_record: every
frame equals steps + the render pass (RGB::to_rgba, alpha 255) taken one by
one on a twin lattice, byte for byte, whatever the ring depth; the recorded
run's populations equal the oracle's run of the same steps value for value;
the last frame matches the numpy render restatement of the oracle's state
(within one unit on a few pixels, as the render pass itself)."""
import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W
from oracle import pyoracle as O
from oracle import render_oracle as R

pytestmark = pytest.mark.gpu


def cavity(n=64):
    flags = W.cavity_flags(n, n)
    p = Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_uy=0.1)
    op = O.params(n, n, tau=0.6, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.1))
    return n, flags, p, op


@pytest.mark.parametrize("depth", [1, 2, 4])
@pytest.mark.parametrize("mode", [capi.CS_RENDER_DENSITY, capi.CS_RENDER_VELOCITY])
def test_record_equals_step_then_render(ctx, depth, mode):
    n, flags, p, op = cavity()
    frames, spf = 7, 13
    print(f"[record] depth={depth} mode={mode} {n}x{n} frames={frames} spf={spf}", flush=True)
    with Lattice(ctx, n, n, "f32", p, flags) as a, Lattice(ctx, n, n, "f32", p, flags) as b:
        a.init_equilibrium(*W.rest_state(n, n))
        b.init_equilibrium(*W.rest_state(n, n))
        rec, stats = a.record(frames, spf, mode, True, depth)
        assert rec.shape == (frames, n, n, 4) and stats["frames"] == frames and not stats["stopped"]
        for k in range(frames):
            b.step(spf)
            img, _ = b.render(mode, True)
            np.testing.assert_array_equal(rec[k, :, :, :3], img)
            assert (rec[k, :, :, 3] == 255).all()
        assert a.time == b.time == frames * spf
        f = a.download()
        np.testing.assert_array_equal(f, b.download())
    rho, ux, uy = W.rest_state(n, n)
    ref = O.run(op, O.equilibrium(op, rho, ux, uy), frames * spf, flags)
    np.testing.assert_array_equal(f, ref)
    want = R.render(ref, mode, flags, True)
    d = np.abs(rec[-1, :, :, :3].astype(int) - want.astype(int))
    assert d.max() <= 1 and (d > 0).sum() <= max(3, d.size // 1000)


def test_record_callback_order_and_stop(ctx):
    n, flags, p, _ = cavity(48)
    seen = []

    def on_frame(k, rgba):
        seen.append((k, int(rgba[..., :3].astype(np.int64).sum())))
        return k < 2          # stop after frame 2 is delivered

    with Lattice(ctx, n, n, "f32", p, flags) as a:
        a.init_equilibrium(*W.rest_state(n, n))
        out, stats = a.record(10, 5, capi.CS_RENDER_SPEED, True, 3, on_frame)
        assert out is None and stats["stopped"] and stats["frames"] == 2
        assert [k for k, _ in seen] == [0, 1, 2]
        # frames in flight when the stop came were stepped: 3 .. 3 + depth - 1 frames
        assert 15 <= a.time <= 25


def test_record_callback_exception_propagates(ctx):
    n, flags, p, _ = cavity(32)

    def boom(k, rgba):
        raise ValueError("encoder failed")

    with Lattice(ctx, n, n, "f32", p, flags) as a:
        a.init_equilibrium(*W.rest_state(n, n))
        with pytest.raises(ValueError, match="encoder failed"):
            a.record(4, 1, capi.CS_RENDER_DENSITY, True, 2, boom)
        a.step(1)                                   # the lattice is still usable


def test_record_argument_checks(ctx):
    n, flags, p, _ = cavity(16)
    with Lattice(ctx, n, n, "f32", p, flags) as a:
        a.init_equilibrium(*W.rest_state(n, n))
        for kw in (dict(depth=0), dict(depth=capi.CS_RECORD_MAX_DEPTH + 1), dict(mode=7)):
            with pytest.raises(capi.ChemsimError):
                a.record(2, 1, **kw)
        assert a.time == 0
        rec, stats = a.record(0, 1)
        assert rec.shape[0] == 0 and stats["frames"] == 0


def test_record_large_frames_depths_agree(ctx):
    """A 2048 x 2048 Taylor-Green recording: depth 3 equals depth 1 byte for
    byte.  The wall times are printed, not asserted: they are dominated by
    this test's host callback (summing each 16 MB frame), whose speed varies
    from box to box; bench.py's record_4096 leg measures the overlap."""
    n = 2048
    rho, ux, uy = W.taylor_green(n, 0.05, np.float32)
    p = Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC)
    sums = {}
    walls = {}
    for depth in (1, 3):
        with Lattice(ctx, n, n, "f32", p) as a:
            a.init_equilibrium(rho, ux, uy)
            acc = []
            _, st = a.record(6, 4, capi.CS_RENDER_VELOCITY, True, depth,
                             lambda k, rgba: acc.append(int(rgba.astype(np.int64).sum())))
            sums[depth] = acc
            walls[depth] = st["wall_ms"]
    print(f"[record] 2048^2 wall ms by ring depth: {walls}", flush=True)
    assert sums[1] == sums[3]


@pytest.mark.parametrize("w,h", [(5, 3), (1, 7), (33, 1)])
def test_record_ragged_sizes(ctx, w, h):
    """Frames of 4·w·h bytes that are not a multiple of 16 still land in
    aligned device slots; each equals steps + render on a twin lattice."""
    p = Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC)
    rng = np.random.default_rng(3)
    rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
    ux = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
    uy = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
    print(f"[record] {w}x{h} depth=3", flush=True)
    with Lattice(ctx, w, h, "f32", p) as a, Lattice(ctx, w, h, "f32", p) as b:
        a.init_equilibrium(rho, ux, uy)
        b.init_equilibrium(rho, ux, uy)
        rec, stats = a.record(5, 2, capi.CS_RENDER_VELOCITY, True, 3)
        assert rec.shape == (5, h, w, 4) and stats["frames"] == 5
        for k in range(5):
            b.step(2)
            img, _ = b.render(capi.CS_RENDER_VELOCITY, True)
            np.testing.assert_array_equal(rec[k, :, :, :3], img)
            assert (rec[k, :, :, 3] == 255).all()
