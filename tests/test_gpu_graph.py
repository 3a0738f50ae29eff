"""Graph replay of fused steps (launch-bound lattices): runs of 64 steps are
captured once per buffer parity and replayed.  Every replay is checked against
the same lattice stepped launch by launch (graphs switched off), value for
value, with eager work, odd step counts, parameter and geometry changes between
the replays, and against the oracle."""
import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

STORE = {"f32": np.float32, "f64": np.float64, "f32dev": np.float32, "f16": np.float32}


def stepped(lat, n, graphs):
    capi.call("cs_set_step_graphs", 1 if graphs else 0)
    try:
        lat.step(n)
    finally:
        capi.call("cs_set_step_graphs", 1)


def case(kind, n):
    if kind == "cavity":
        flags = W.cavity_flags(n, n)
        p = Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_uy=0.1)
        op = O.params(n, n, tau=0.6, boundary=O.BB_HALFWAY, lid=W.lid_velocity(0.1))
        return flags, p, op, W.rest_state(n, n)
    p = Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC)
    op = O.params(n, n, tau=0.8, edge=O.EDGE_PERIODIC)
    return None, p, op, W.taylor_green(n, 0.05, np.float32)


@pytest.mark.parametrize("storage", ["f32", "f64", "f32dev", "f16"])
@pytest.mark.parametrize("kind", ["cavity", "tgv"])
def test_graph_replay_equals_step_by_step(ctx, storage, kind):
    n = 128 if kind == "cavity" else 256
    flags, p, op, init = case(kind, n)
    dt = STORE[storage]
    print(f"[graph] {kind} {n}x{n} {storage}", flush=True)
    with Lattice(ctx, n, n, storage, p, flags) as a, Lattice(ctx, n, n, storage, p, flags) as b:
        a.init_equilibrium(*init)
        b.init_equilibrium(*init)
        total = 0
        for k, steps in enumerate((200, 65, 128, 64, 1, 130)):
            stepped(a, steps, True)
            stepped(b, steps, False)
            total += steps
            assert a.time == b.time == total
            np.testing.assert_array_equal(a.download_raw(), b.download_raw(), err_msg=f"call {k}")
            a.velocity()  # eager work between replays
        if storage == "f32" and kind == "cavity":
            rho, ux, uy = init
            ref = O.run(op, O.equilibrium(op, rho, ux, uy), total, flags)
            np.testing.assert_array_equal(a.download(dt), ref)
        # new parameters (and geometry) must not replay the old graphs
        q = Params(**{**p.__dict__, "tau": 0.7})
        a.set_params(q)
        b.set_params(q)
        if flags is not None:
            f2 = flags.copy()
            f2[n // 2, n // 4:n // 2] = O.WALL
            a.set_flags(f2)
            b.set_flags(f2)
        stepped(a, 192, True)
        stepped(b, 192, False)
        np.testing.assert_array_equal(a.download_raw(), b.download_raw())


def test_graph_replay_with_cells_per_thread_override(ctx):
    n = 128
    flags, p, op, init = case("cavity", n)
    with Lattice(ctx, n, n, "f32", p, flags) as a:
        a.init_equilibrium(*init)
        stepped(a, 128, True)
        capi.call("cs_set_cells_per_thread", 2)
        try:
            stepped(a, 128, True)  # the override changes the launch shape: graphs are recaptured
        finally:
            capi.call("cs_set_cells_per_thread", 0)
        stepped(a, 64, True)
        rho, ux, uy = init
        np.testing.assert_array_equal(a.download(), O.run(op, O.equilibrium(op, rho, ux, uy), 320, flags))
