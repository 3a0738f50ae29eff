"""Batched persistent launch: every lattice equals the oracle's run of that
lattice alone, and the one-launch batch equals one launch per lattice per step,
bit for bit."""
import numpy as np
import pytest

from chemsim_amd import Params, capi
from chemsim_amd import workloads as W
from chemsim_amd.batch import Batch
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def build_mixed(ctx, storage="f32"):
    """Lattices of uneven sizes and every boundary mode."""
    rng = np.random.default_rng(5)
    specs = []
    for i, (w, h) in enumerate([(64, 64), (100, 37), (33, 80), (128, 128), (72, 40), (40, 72), (96, 24),
                                (17, 19), (150, 64), (64, 150)]):
        kind = i % 5
        if kind == 0:      # lid-driven cavity, half-way
            flags = W.cavity_flags(w, h)
            op = O.params(w, h, tau=0.55 + 0.05 * i, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.02 + 0.01 * i))
        elif kind == 1:    # periodic, no flags
            flags = None
            op = O.params(w, h, tau=0.8, edge=O.EDGE_PERIODIC)
        elif kind == 2:    # reference mode with a random mask, zero edges
            flags = (rng.random((h, w)) < 0.1).astype(np.uint8)
            op = O.params(w, h, tau=0.7)
        elif kind == 3:    # channel with Zou-He
            ch = W.Channel(w=w, h=h, circles=[(h // 2, w // 3, 4)], re=20.0, u_in=0.04)
            flags = ch.flags()
            op = O.params(w, h, tau=ch.tau, boundary=O.BB_HALFWAY, inlet=ch.inlet(), outlet_rho=1.0)
        else:              # periodic half-way with a moving wall strip
            flags = np.zeros((h, w), np.uint8)
            flags[h // 2, :] = O.LID
            op = O.params(w, h, tau=0.9, edge=O.EDGE_PERIODIC, boundary=O.BB_HALFWAY, lid=(0.0, 0.03))
        rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
        ux = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
        uy = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
        specs.append((w, h, op, flags, (rho, ux, uy)))
    b = Batch(ctx, [(s[0], s[1]) for s in specs], storage)
    for i, (w, h, op, flags, init) in enumerate(specs):
        b.set_params(i, Params(tau=op.tau, edge=op.edge, boundary=op.boundary, lid_ux=op.lid_ux, lid_uy=op.lid_uy,
                               inlet_ux=op.inlet_ux, inlet_uy=op.inlet_uy, outlet_rho=op.outlet_rho,
                               rounding=capi.CS_ROUND_STOCHASTIC, seed=77 + i))
        b.set_flags(i, flags)
        b.init_equilibrium(i, *init)
    return b, specs


@pytest.mark.parametrize("storage", ["f32", "f32dev"])
def test_batch_temporal_matches_oracle(ctx, storage):
    """The temporal launch kind (8 steps per launch out of LDS, then tile
    launches for the remainder) over every boundary mode."""
    b, specs = build_mixed(ctx, storage)
    with b:
        b.tuning = dict(block_threads=256, workgroups_per_cu=1, cells_per_thread=1, order=0,
                        launch=capi.CS_LAUNCH_TEMPORAL)
        b.step(19)
        b.step(8)
        for i, (w, h, op, flags, init) in enumerate(specs):
            if storage == "f32":
                ref = O.run(op, O.equilibrium(op, *init), 27, flags)
            else:
                ref = O.from_dev(op, O.run_devf32(op, O.equilibrium_dev(op, *init), 27, flags))
            np.testing.assert_array_equal(b.download(i), ref, err_msg=f"lattice {i} ({w}x{h})")


def test_batch_temporal_needs_32bit_storage(ctx):
    b, _ = build_mixed(ctx, "f16")
    with b:
        with pytest.raises(RuntimeError, match="32-bit|FP32"):
            b.tuning = dict(block_threads=256, workgroups_per_cu=1, cells_per_thread=1, order=0,
                            launch=capi.CS_LAUNCH_TEMPORAL)


@pytest.mark.parametrize("storage", ["f32", "f32dev"])
def test_batch_matches_oracle(ctx, storage):
    b, specs = build_mixed(ctx, storage)
    with b:
        b.step(7)
        b.step(6)
        for i, (w, h, op, flags, init) in enumerate(specs):
            if storage == "f32":
                ref = O.run(op, O.equilibrium(op, *init), 13, flags)
            else:
                ref = O.from_dev(op, O.run_devf32(op, O.equilibrium_dev(op, *init), 13, flags))
            np.testing.assert_array_equal(b.download(i), ref, err_msg=f"lattice {i} ({w}x{h})")


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_batch_equals_per_lattice_launches(ctx, storage):
    tunings = [dict(block_threads=256, workgroups_per_cu=1, cells_per_thread=1, order=0),
               dict(block_threads=1024, workgroups_per_cu=2, cells_per_thread=2, order=1),
               dict(block_threads=256, workgroups_per_cu=1, cells_per_thread=1, order=0, launch=capi.CS_LAUNCH_TILES),
               dict(block_threads=512, workgroups_per_cu=1, cells_per_thread=1, order=0, launch=capi.CS_LAUNCH_TILES),
               dict(block_threads=1024, workgroups_per_cu=1, cells_per_thread=1, order=0, launch=capi.CS_LAUNCH_TILES),
               dict(block_threads=512, workgroups_per_cu=8, cells_per_thread=2, order=2)]
    steps = [3] * len(tunings)
    if storage == "f32":  # 8-step launches plus a remainder, then the others continue
        tunings.insert(3, dict(block_threads=512, workgroups_per_cu=1, cells_per_thread=1, order=0,
                               launch=capi.CS_LAUNCH_TEMPORAL))
        steps.insert(3, 11)
    a, _ = build_mixed(ctx, storage)
    p, _ = build_mixed(ctx, storage)
    with a, p:
        p.step_per_lattice(sum(steps))
        for t, n in zip(tunings, steps):
            a.tuning = t
            a.step(n)
        for i in range(len(a)):
            np.testing.assert_array_equal(a.download_raw(i), p.download_raw(i), err_msg=f"lattice {i}")


def test_batch_16bit_stochastic_matches_oracle(ctx):
    b, specs = build_mixed(ctx, "f16")
    with b:
        b.step(5)
        for i, (w, h, op, flags, init) in enumerate(specs):
            d0 = O.to_f16(O.equilibrium_dev(op, *init))
            ref = O.run_dev16(op, O.F16, d0, 5, flags, rounding=1, seed=77 + i)
            np.testing.assert_array_equal(b.download_raw(i), ref, err_msg=f"lattice {i}")


def test_autotune_keeps_state(ctx):
    a, _ = build_mixed(ctx, "f32")
    p, _ = build_mixed(ctx, "f32")
    with a, p:
        a.step(4)
        p.step(4)
        tuning, ms = a.autotune(3)
        assert ms > 0 and tuning["block_threads"] in (256, 512, 1024)
        a.step(5)
        p.step(5)
        for i in range(len(a)):
            np.testing.assert_array_equal(a.download(i), p.download(i))


@pytest.mark.parametrize("launch", [None, capi.CS_LAUNCH_TEMPORAL, capi.CS_LAUNCH_TEMPORAL_PERSISTENT])
def test_cavity_batch_small(ctx, launch):
    """A reduced version of the 512-cavity workload: 24 cavities, sides 64..160
    (one boundary-mode group), with the default launch and the temporal kinds
    (6 rounds of 8 steps in one persistent launch, then 2 tile launches)."""
    specs = W.batch_cavities(seed=3, count=24, lo=64, hi=160)
    b = Batch(ctx, [(s.side, s.side) for s in specs], "f32")
    with b:
        if launch is not None:
            b.tuning = dict(block_threads=256, workgroups_per_cu=1, cells_per_thread=1, order=0, launch=launch)
        for i, s in enumerate(specs):
            b.set_params(i, Params(tau=s.tau, boundary=capi.CS_BB_HALFWAY, lid_uy=s.u_lid))
            b.set_flags(i, W.cavity_flags(s.side, s.side))
            b.init_equilibrium(i, *W.rest_state(s.side, s.side))
        b.step(50)
        for i in (0, 7, 23):
            s = specs[i]
            op = O.params(s.side, s.side, tau=s.tau, boundary=O.BB_HALFWAY, lid=W.lid_velocity(s.u_lid))
            ref = O.run(op, O.equilibrium(op, *W.rest_state(s.side, s.side)), 50, W.cavity_flags(s.side, s.side))
            np.testing.assert_array_equal(b.download(i), ref)


def test_inactive_lattices_keep_their_state(ctx):
    """A lattice taken out of the batch for an odd number of steps keeps its
    populations and step count; stepped again, it continues where it stopped
    (stochastic rounding included)."""
    for launch in (capi.CS_LAUNCH_PERSISTENT, capi.CS_LAUNCH_TILES):
        a, specs = build_mixed(ctx, "f16")
        with a:
            t = a.tuning
            t["launch"] = launch
            a.tuning = t
            a.step(2)
            a.set_active(3, False)
            frozen = a.download_raw(3)
            a.step(5)
            np.testing.assert_array_equal(a.download_raw(3), frozen)
            a.set_active(3, True)
            a.step(4)
            w, h, op, flags, init = specs[3]
            d0 = O.to_f16(O.equilibrium_dev(op, *init))
            ref = O.run_dev16(op, O.F16, d0, 6, flags, rounding=1, seed=77 + 3)
            np.testing.assert_array_equal(a.download_raw(3), ref)
            for i in (0, 5):
                w, h, op, flags, init = specs[i]
                d0 = O.to_f16(O.equilibrium_dev(op, *init))
                np.testing.assert_array_equal(a.download_raw(i), O.run_dev16(op, O.F16, d0, 11, flags, rounding=1,
                                                                             seed=77 + i))


@pytest.mark.parametrize("launch,threads", [(capi.CS_LAUNCH_TILES, 256), (capi.CS_LAUNCH_TILES, 1024),
                                            (capi.CS_LAUNCH_PERSISTENT, 256), (capi.CS_LAUNCH_TEMPORAL, 256),
                                            (capi.CS_LAUNCH_TEMPORAL_PERSISTENT, 256)])
def test_batch_tiny_lattices_match_oracle(ctx, launch, threads):
    """1x1, one-cell-wide and small ragged lattices beside a normal one, in
    both launch kinds: every lattice equals the oracle's run of it alone."""
    rng = np.random.default_rng(9)
    shapes = [(1, 1), (1, 9), (9, 1), (3, 5), (70, 2), (64, 64)]
    b = Batch(ctx, shapes, "f32")
    specs = []
    with b:
        b.tuning = dict(block_threads=threads, workgroups_per_cu=1, cells_per_thread=1, order=0, launch=launch)
        for i, (w, h) in enumerate(shapes):
            edge = O.EDGE_PERIODIC if i % 2 else O.EDGE_ZERO
            flags = (rng.random((h, w)) < 0.3).astype(np.uint8) if i in (3, 5) else None
            op = O.params(w, h, tau=0.6 + 0.05 * i, edge=edge)
            b.set_params(i, Params(tau=op.tau, edge=edge))
            b.set_flags(i, flags)
            rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
            ux = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
            uy = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
            b.init_equilibrium(i, rho, ux, uy)
            specs.append((op, flags, O.equilibrium(op, rho, ux, uy)))
        b.step(13)
        for i, (op, flags, f0) in enumerate(specs):
            np.testing.assert_array_equal(b.download(i), O.run(op, f0, 13, flags), err_msg=f"lattice {i} {shapes[i]}")


def test_persistent_temporal_long_call(ctx):
    """A call of more than 4096 rounds of 8 steps (32,811 steps): two
    persistent launches, the second continuing the first's round count (the
    buffer parity of each lattice), then 3 tile launches; against the
    temporal kind's one launch per round, bit for bit."""
    shapes = [(20, 20), (9, 31), (1, 1)]
    def make(kind):
        b = Batch(ctx, shapes, "f32")
        for i, (w, h) in enumerate(shapes):
            b.set_params(i, Params(tau=0.7 + 0.05 * i, edge=capi.CS_EDGE_PERIODIC))
            r = np.random.default_rng(300 + i)
            b.init_equilibrium(i, (1 + 0.01 * r.standard_normal((h, w))).astype(np.float32),
                               (0.02 * r.standard_normal((h, w))).astype(np.float32),
                               (0.02 * r.standard_normal((h, w))).astype(np.float32))
        b.tuning = dict(block_threads=512, workgroups_per_cu=1, cells_per_thread=1, order=0, launch=kind)
        return b
    steps = 8 * 4101 + 3
    with make(capi.CS_LAUNCH_TEMPORAL_PERSISTENT) as a, make(capi.CS_LAUNCH_TEMPORAL) as t:
        a.step(steps)
        t.step(steps)
        for i in range(len(shapes)):
            np.testing.assert_array_equal(a.download_raw(i), t.download_raw(i), err_msg=f"lattice {i} {shapes[i]}")

@pytest.mark.parametrize("storage", ["f32", "f32dev"])
def test_persistent_temporal_fluid_batch(ctx, storage):
    """One boundary-mode group without geometry (periodic, the register
    variant's grid kernel): 12 lattices of uneven sizes, 19 steps (2 rounds in
    one persistent launch, then 3 tile launches) and 16 more, against one
    launch per lattice per step, bit for bit."""
    shapes = [(64, 64), (100, 37), (33, 80), (128, 128), (72, 40), (40, 72), (96, 24), (17, 19), (150, 64),
              (64, 150), (1, 1), (300, 7)]
    def make():
        b = Batch(ctx, shapes, storage)
        for i, (w, h) in enumerate(shapes):
            b.set_params(i, Params(tau=0.6 + 0.03 * i, edge=capi.CS_EDGE_PERIODIC))
            r = np.random.default_rng(100 + i)
            b.init_equilibrium(i, (1 + 0.01 * r.standard_normal((h, w))).astype(np.float32),
                               (0.02 * r.standard_normal((h, w))).astype(np.float32),
                               (0.02 * r.standard_normal((h, w))).astype(np.float32))
        return b
    with make() as a, make() as p:
        a.tuning = dict(block_threads=512, workgroups_per_cu=1, cells_per_thread=1, order=0,
                        launch=capi.CS_LAUNCH_TEMPORAL_PERSISTENT)
        a.step(19)
        a.step(16)
        p.step_per_lattice(35)
        for i in range(len(shapes)):
            np.testing.assert_array_equal(a.download_raw(i), p.download_raw(i), err_msg=f"lattice {i} {shapes[i]}")
