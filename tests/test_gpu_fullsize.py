"""Full-size claims as GPU tests.

1. 16-bit deviation storage on the periodic Taylor-Green vortex at 4096^2
   (1,000 steps) and 8192^2 (200 steps), tau 0.8, u0 0.05: relative L2 error
   in u and in rho - 1 against the FP64 run of the same kernel, and the mass
   drift.  Bounds (written here): FP16 with stochastic rounding u <= 1e-3 (the
   target; measured 3.6e-4 / 3.3e-4), BF16 with stochastic rounding u <= 4e-3
   (measured 2.8e-3 / 2.6e-3; the 1e-3 target is not met by BF16: its 8
   significant bits bound every stored deviation's relative error at 2^-9),
   FP32 deviation arithmetic u <= 5e-5, mass drift <= 1e-6 for all three.  The
   FP64 run itself decays as exp(-2 nu k^2 t) to 1e-5 and conserves mass to
   1e-12.
2. The 512-cavity batch (sides 64-512 seeded, 54.3 M cells, tau and lid
   seeded): after 3 steps the persistent launch, the tile launch and one
   launch per lattice per step are equal bit for bit on every lattice, and the
   smallest, median and largest lattices equal the oracle's run of them alone.
"""
import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W
from chemsim_amd.batch import Batch
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

TAU, U0 = 0.8, 0.05


def tgv(ctx, n, storage, rounding=capi.CS_ROUND_STOCHASTIC):
    rho, ux, uy = W.taylor_green(n, U0, np.float64 if storage == "f64" else np.float32)
    lat = Lattice(ctx, n, n, storage, Params(tau=TAU, edge=capi.CS_EDGE_PERIODIC, rounding=rounding, seed=2024))
    lat.init_equilibrium(rho, ux, uy)
    return lat


def errors_vs(ref, lat):
    r64, x64, y64 = ref
    r, x, y = lat.moments_f64()
    e_u = np.sqrt(((x - x64) ** 2 + (y - y64) ** 2).sum() / (x64 ** 2 + y64 ** 2).sum())
    e_r = np.sqrt(((r - r64) ** 2).sum() / ((r64 - 1.0) ** 2).sum())
    return float(e_u), float(e_r)


@pytest.mark.parametrize("n,steps", [(4096, 1000), (8192, 200)])
def test_16bit_storage_error_vs_fp64(ctx, n, steps):
    print(f"[fullsize] TGV {n}^2 FP64, {steps} steps", flush=True)
    with tgv(ctx, n, "f64") as ref:
        m0 = ref.sums_f64()[0]
        ref.step(steps)
        m1 = ref.sums_f64()[0]
        moments = ref.moments_f64()
    assert abs(m1 - m0) / m0 <= 1e-12
    _, ux0, uy0 = W.taylor_green(n, U0)
    amp = float((moments[1] * ux0 + moments[2] * uy0).sum() / (ux0 * ux0 + uy0 * uy0).sum())
    del ux0, uy0
    assert abs(amp / W.taylor_green_decay(n, TAU, steps) - 1) <= 1e-5
    bounds = {"f16": (1e-3, 1.5e-2), "bf16": (4e-3, 1e-1), "f32dev": (5e-5, 2e-4)}
    for storage, (bu, br) in bounds.items():
        print(f"[fullsize] TGV {n}^2 {storage}, {steps} steps", flush=True)
        with tgv(ctx, n, storage) as lat:
            s0 = lat.sums_f64()[0]
            lat.step(steps)
            s1 = lat.sums_f64()[0]
            e_u, e_r = errors_vs(moments, lat)
        print(f"[fullsize] {storage}: u {e_u:.3e} rho-1 {e_r:.3e} mass drift {abs(s1 - s0) / s0:.2e}", flush=True)
        assert e_u <= bu, (storage, e_u)
        assert e_r <= br, (storage, e_r)
        assert abs(s1 - s0) / s0 <= 1e-6, (storage, s0, s1)


def cavity_batch(ctx, specs, tuning=None):
    b = Batch(ctx, [(s.side, s.side) for s in specs], "f32")
    if tuning is not None:
        b.tuning = tuning
    for i, s in enumerate(specs):
        b.set_params(i, Params(tau=s.tau, boundary=capi.CS_BB_HALFWAY, lid_uy=s.u_lid))
        b.set_flags(i, W.cavity_flags(s.side, s.side))
        b.init_equilibrium(i, *W.rest_state(s.side, s.side))
    return b


def test_512_cavity_batch_full_size(ctx):
    specs = W.batch_cavities(seed=0, count=512, lo=64, hi=512)
    cells = sum(s.side * s.side for s in specs)
    print(f"[fullsize] 512 cavities, {cells} cells", flush=True)
    assert cells > 50_000_000
    steps = 11  # the temporal kind: one 8-step launch, then 3 tile launches
    persistent = dict(block_threads=256, workgroups_per_cu=2, cells_per_thread=1, order=capi.CS_ORDER_LARGEST_FIRST,
                      launch=capi.CS_LAUNCH_PERSISTENT)
    tiles = dict(block_threads=256, workgroups_per_cu=1, cells_per_thread=1, order=capi.CS_ORDER_LARGEST_FIRST,
                 launch=capi.CS_LAUNCH_TILES)
    temporal = dict(tiles, launch=capi.CS_LAUNCH_TEMPORAL)
    with cavity_batch(ctx, specs, persistent) as a, cavity_batch(ctx, specs, tiles) as b:
        a.step(steps)
        b.step(steps)
        with cavity_batch(ctx, specs) as c:
            c.step_per_lattice(steps)
            for i in range(len(specs)):
                ra = a.download_raw(i)
                np.testing.assert_array_equal(ra, b.download_raw(i), err_msg=f"tiles, lattice {i}")
                np.testing.assert_array_equal(ra, c.download_raw(i), err_msg=f"per-lattice, lattice {i}")
        for kind in (capi.CS_LAUNCH_TEMPORAL, capi.CS_LAUNCH_TEMPORAL_PERSISTENT):
            print(f"[fullsize] 512 cavities, launch kind {kind}", flush=True)
            with cavity_batch(ctx, specs, dict(temporal, launch=kind)) as t:
                t.step(steps)
                for i in range(len(specs)):
                    np.testing.assert_array_equal(a.download_raw(i), t.download_raw(i),
                                                  err_msg=f"launch kind {kind}, lattice {i}")
        order = np.argsort([s.side for s in specs], kind="stable")
        for i in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
            s = specs[i]
            print(f"[fullsize] cavity {i}: side {s.side} vs the oracle", flush=True)
            op = O.params(s.side, s.side, tau=s.tau, boundary=O.BB_HALFWAY, lid=W.lid_velocity(s.u_lid))
            f0 = O.equilibrium(op, *W.rest_state(s.side, s.side))
            ref, _ = O.run_omp(op, f0, steps, W.cavity_flags(s.side, s.side))
            np.testing.assert_array_equal(a.download(i), ref, err_msg=f"lattice {i} side {s.side}")
