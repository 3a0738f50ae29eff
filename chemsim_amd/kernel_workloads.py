"""Short runs of each kernel of the library on its measurement workload, for
kernel traces and HBM byte counters (rocprofv3 --kernel-trace / --pmc).

    python3 -m chemsim_amd.kernel_workloads [names ...] [--steps N]

Every workload is preceded by three launches of the copy kernel `k_copy`
(1 GiB read + 1 GiB written), whose counter readings calibrate FETCH_SIZE and
WRITE_SIZE to bytes.  Each workload prints one JSON line on stdout naming the
kernel(s) it launches and their algorithmic bytes per launch.

Workloads:
  collisions  Taylor-Green 4096^2 FP32, periodic: TRT (lbm.rs:374-451,
              TRT::new(1/4, nu of tau 0.8)), Regularized (lbm.rs:596-666) and
              KBC (lbm.rs:458-590); 72 B per cell update each
  f64         Taylor-Green 4096^2 FP64 storage and arithmetic; 144 B/cell
  dev16_8192  Taylor-Green 8192^2, BF16 and FP16 deviations with stochastic
              rounding; 36 B/cell
  cavity4096  lid-driven cavity 4096^2 FP32, half-way walls, lid 0.1, from
              rest, with geometry (node word per cell): the temporally
              blocked kernel (8 steps per launch; regions read once, cells
              written once per launch + 4 B/cell of node words) and the
              one-step geometry kernel (72 B/cell + 4 B/cell of node words)
  cavity128   lid-driven cavity 128^2 FP32 (16,384 cells), launches one by one
              (step graphs off): the temporally blocked kernel on 32 x 32
              regions (8 steps per launch, 64 regions) and the one-step kernel
  render      render pass of the 4096^2 Taylor-Green state in the four display
              modes (render.rs:7-178): field pass (36 B/cell read, 4 or 8 B/cell
              written), statistics, pixel pass (4 or 8 B/cell read, 3 B/pixel)
  batch512    512 lid-driven cavities, sides U{64..512} (seed 0), tile-list
              and persistent launches (72 B per cell update) and the temporal
              launch kind (8 steps per launch; every 64 x 64 region read and
              every cell written once per launch)
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import Context, Lattice, Params, capi
from . import workloads as W
from .lattice import trt_new

N4 = 4096
TAU, U0 = 0.8, 0.05
COPY_BYTES = 1 << 30


def _emit(rec: dict) -> None:
    print(json.dumps(rec), flush=True)


def _calibrate(ctx: Context) -> None:
    ctx.bench_copy(COPY_BYTES, 3)


def _tgv(ctx, n, storage, rounding=capi.CS_ROUND_STOCHASTIC, **kw):
    rho, ux, uy = W.taylor_green(n, U0, np.float64 if storage == "f64" else np.float32)
    p = Params(tau=TAU, edge=capi.CS_EDGE_PERIODIC, rounding=rounding, seed=2024)
    for k, v in kw.items():
        setattr(p, k, v)
    lat = Lattice(ctx, n, n, storage, p)
    lat.init_equilibrium(rho, ux, uy)
    return lat


def _timed(lat, steps):
    lat.step(max(2, lat.temporal_steps()))  # warm-up: every kernel of the timed steps launched once
    ms = lat.step_timed(steps)
    return ms / steps * 1e3


def collisions(ctx, steps):
    cells = N4 * N4
    tp, tm = trt_new(0.25, (TAU - 0.5) / 3.0)
    runs = (("trt", dict(collision=capi.CS_COLL_TRT, tau=tp, tau_minus=tm)),
            ("regularized", dict(collision=capi.CS_COLL_REGULARIZED, ks_viscosity=(TAU - 0.5) / 3.0)),
            ("kbc", dict(collision=capi.CS_COLL_KBC, ks_viscosity=(TAU - 0.5) / 3.0)))
    for name, kw in runs:
        _calibrate(ctx)
        with _tgv(ctx, N4, "f32", **kw) as lat:
            us = _timed(lat, steps)
        _emit({"workload": f"collisions/{name}", "kernel": "k_step_vec", "cells": cells,
               "alg_bytes_per_launch": 72 * cells, "event_us_per_step": round(us, 2)})


def f64(ctx, steps):
    cells = N4 * N4
    _calibrate(ctx)
    with _tgv(ctx, N4, "f64") as lat:
        us = _timed(lat, steps)
    _emit({"workload": "f64", "kernel": "k_step_vec", "cells": cells, "alg_bytes_per_launch": 144 * cells,
           "event_us_per_step": round(us, 2)})


def dev16_8192(ctx, steps):
    n = 8192
    cells = n * n
    for st in ("bf16", "f16"):
        _calibrate(ctx)
        with _tgv(ctx, n, st) as lat:
            us = _timed(lat, steps)
        _emit({"workload": f"dev16_8192/{st}", "kernel": "k_step", "cells": cells,
               "alg_bytes_per_launch": 36 * cells, "event_us_per_step": round(us, 2)})


def _cavity(ctx, n):
    p = Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_ux=0.0, lid_uy=0.1)
    lat = Lattice(ctx, n, n, "f32", p, W.cavity_flags(n, n))
    lat.init_equilibrium(*W.rest_state(n, n))
    return lat


def tb_bytes_per_launch(w: int, h: int, k: int = 8, edge: int = 64) -> int:
    """Bytes one launch of the temporally blocked kernel (k steps) must move
    for a w x h FP32 lattice: every edge x edge region read once, every cell
    written once, 36 B per cell."""
    inner = edge - 2 * k
    return (-(-w // inner) * -(-h // inner) * edge * edge + w * h) * 36


def cavity4096(ctx, steps):
    cells = N4 * N4
    steps8 = max(8, steps // 8 * 8)
    for k in (capi.CS_TEMPORAL_AUTO, 0):
        capi.call("cs_set_temporal_steps", k)
        try:
            _calibrate(ctx)
            with _cavity(ctx, N4) as lat:
                tb = lat.temporal_steps()
                us = _timed(lat, steps8)
        finally:
            capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)
        if tb > 1:
            _emit({"workload": "cavity4096/temporal", "kernel": "k_step_tb", "cells": cells, "steps_per_launch": tb,
                   "alg_bytes_per_launch": tb_bytes_per_launch(N4, N4, tb), "node_bytes_per_launch": 4 * cells,
                   "event_us_per_step": round(us, 2)})
        else:
            _emit({"workload": "cavity4096/one_step", "kernel": "k_step_vec", "cells": cells,
                   "alg_bytes_per_launch": 72 * cells, "node_bytes_per_launch": 4 * cells,
                   "event_us_per_step": round(us, 2)})


def cavity128(ctx, steps):
    n = 128
    capi.call("cs_set_step_graphs", 0)
    try:
        _calibrate(ctx)
        with _cavity(ctx, n) as lat:
            k = lat.temporal_steps()
            us_tb = _timed(lat, max(k, steps // k * k))
        capi.call("cs_set_temporal_steps", 0)
        with _cavity(ctx, n) as lat:
            us = _timed(lat, steps)
    finally:
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)
        capi.call("cs_set_step_graphs", 1)
    _emit({"workload": "cavity128/temporal", "kernel": "k_step_tb (32 x 32 regions)", "cells": n * n,
           "alg_bytes_per_launch": tb_bytes_per_launch(n, n, k, 32), "steps_per_launch": k,
           "event_us_per_step": round(us_tb, 2)})
    _emit({"workload": "cavity128/one_step", "kernel": "k_step", "cells": n * n, "alg_bytes_per_launch": 72 * n * n,
           "event_us_per_step": round(us, 2)})


def render(ctx, steps):
    cells = N4 * N4
    _calibrate(ctx)
    modes = (("density", capi.CS_RENDER_DENSITY, False), ("speed", capi.CS_RENDER_SPEED, False),
             ("velocity", capi.CS_RENDER_VELOCITY, True), ("momentum", capi.CS_RENDER_MOMENTUM, True))
    with _tgv(ctx, N4, "f32") as lat:
        lat.step(2)
        for name, mode, phase in modes:
            lat.render(mode)
            ms = [lat.render(mode)[1] for _ in range(max(1, min(steps, 5)))]
            side = 8 if phase else 4    # sval (+ phase) per cell
            _emit({"workload": f"render/{name}", "kernel": "k_render_field + k_render_stats + k_render_pixels",
                   "cells": cells, "alg_bytes_field": (36 + side) * cells, "alg_bytes_pixels": (side + 3) * cells,
                   "event_ms_all_passes": round(min(ms), 4)})


def batch512(ctx, steps):
    from .batch import Batch
    specs = W.batch_cavities(seed=0, count=512)
    _calibrate(ctx)
    with Batch(ctx, [(s.side, s.side) for s in specs], "f32") as b:
        for i, s in enumerate(specs):
            b.set_params(i, Params(tau=s.tau, boundary=capi.CS_BB_HALFWAY, lid_uy=s.u_lid))
            b.set_flags(i, W.cavity_flags(s.side, s.side))
            b.init_equilibrium(i, *W.rest_state(s.side, s.side))
        cells = b.cells
        tb_bytes = sum(tb_bytes_per_launch(s.side, s.side) for s in specs)
        for kind, tuning in (("tiles", dict(block_threads=256, workgroups_per_cu=1, cells_per_thread=1, order=0,
                                            launch=capi.CS_LAUNCH_TILES)),
                             ("persistent", dict(block_threads=1024, workgroups_per_cu=2, cells_per_thread=1,
                                                 order=0, launch=capi.CS_LAUNCH_PERSISTENT)),
                             ("temporal", dict(block_threads=512, workgroups_per_cu=1, cells_per_thread=1, order=0,
                                               launch=capi.CS_LAUNCH_TEMPORAL)),
                             ("temporal_persistent", dict(block_threads=512, workgroups_per_cu=1, cells_per_thread=1,
                                                          order=0, launch=capi.CS_LAUNCH_TEMPORAL_PERSISTENT))):
            b.tuning = tuning
            tb = kind.startswith("temporal")
            n = max(8, steps // 8 * 8) if tb else steps
            b.step(8 if tb else 2)
            ms = b.step(n)
            kernel = {"tiles": "k_batch_tiles", "persistent": "k_batch", "temporal": "k_batch_tb",
                      "temporal_persistent": "k_batch_tb_grid"}[kind]
            _emit({"workload": f"batch512/{kind}", "kernel": kernel, "lattices": len(b), "cells": cells,
                   "alg_bytes_per_launch": tb_bytes if tb else 72 * cells,
                   "steps_per_launch": n if kind == "temporal_persistent" else (8 if tb else 1),
                   "event_us_per_step": round(ms / n * 1e3, 2)})


WORKLOADS = {"collisions": collisions, "f64": f64, "dev16_8192": dev16_8192, "cavity4096": cavity4096,
             "cavity128": cavity128, "render": render, "batch512": batch512}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("names", nargs="*", help="workloads to run (default: all): " + ", ".join(WORKLOADS))
    ap.add_argument("--steps", type=int, default=20)
    args = ap.parse_args(argv)
    names = args.names or list(WORKLOADS)
    unknown = [n for n in names if n not in WORKLOADS]
    if unknown:
        ap.error(f"unknown workload(s) {unknown}; choose from {list(WORKLOADS)}")
    ctx = Context(0)
    try:
        for name in names:
            print(f"[kernel_workloads] {name}", file=sys.stderr, flush=True)
            WORKLOADS[name](ctx, args.steps)
    finally:
        ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
