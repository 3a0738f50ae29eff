"""Benchmark of the D2Q9 BGK hot path on one MI355X.

Default workload: periodic Taylor-Green vortex, 4096 x 4096, FP32, tau = 0.8,
u0 = 0.05, wavenumber 2 pi / 4096; K timed steps after W warm-up steps, with
the populations resident in HBM.  A step is one full update (stream +
bounce-back + BGK collide) of every cell.  FP32 lattices take the temporally
blocked kernel (8 steps per launch with the region in registers; a remainder
of 4 or more steps one 4-step launch, the rest one step per launch); the
one-step kernel of the same workload is timed beside it in `extras`.  Before
the warm-up, each kernel the timed steps use is launched once on a scratch
lattice of the same shape, so first-launch costs stay out of the timed region
(the measured lattice takes exactly W + K steps), and right before every timed
region a scratch lattice's steps keep the GPU busy for 0.05 s (wake_gpu);
`extras.idle_start` times the same region after one second of idle GPU.

Prints ONE JSON line.  `value` is lattice updates per second (GLUPS) over the
timed steps (wall clock around a synchronised region); the roofline object
prices the step kernel's average duration (HIP events on the lattice's own
stream) against the bytes that kernel must move: for the temporally blocked
kernel each 64 x 64 region (interior and halo) read once and each interior
cell written once per launch, for the one-step kernels 72 bytes per cell
update.  The CPU baseline is the oracle's fused FP32 step (OpenMP over rows)
on a bounded sample of the same lattice.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "D2Q9 GLUPS & % HBM roofline; time-to-steady-state, warm-start vs cold"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E peak
BYTES_PER_CELL = {"f32": 72, "f32dev": 72, "f64": 144, "bf16": 36, "f16": 36}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def hbm_frac(bytes_per_cell: float, cells: int, kernel_ms: float) -> float:
    """Algorithmic bytes of one step launch over its duration, as a fraction
    of the MI355X HBM peak."""
    return round(bytes_per_cell * cells / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)


TB_EDGE, TB_STEPS = 64, 8  # region side and steps per launch of the temporally blocked kernel


def tb_region_edge(w: int, h: int) -> int:
    """Region side the library takes for a w x h FP32 lattice: 32 while it has
    at most 1024 regions of 32 x 32 (interior 16 x 16 at 8 steps), else 64."""
    return 32 if -(-w // 16) * -(-h // 16) <= 1024 else TB_EDGE


def tb_bytes_per_step(w: int, h: int, esize: int = 4, k: int = TB_STEPS, edge: int | None = None) -> float:
    """Bytes one launch of the temporally blocked kernel must move, per step:
    every region (interior of side edge - 2k plus halo) read once, every
    lattice cell written once, 9 populations of `esize` bytes.  `edge`: the
    region side (default: the single-lattice choice; batches use 64)."""
    edge = tb_region_edge(w, h) if edge is None else edge
    inner = edge - 2 * k
    regions = -(-w // inner) * -(-h // inner)
    return (regions * edge * edge + w * h) * 9 * esize / k


def step_bytes_per_cell(storage: str, n: int, k: int) -> float:
    """Algorithmic bytes per cell update of the kernel an n x n lattice takes
    (k: steps per launch, Lattice.temporal_steps(); 1 = a one-step kernel)."""
    if k > 1:
        return tb_bytes_per_step(n, n, k=k) / (n * n)
    return BYTES_PER_CELL[storage]


def run_bytes_per_cell(storage: str, n: int, k: int, steps: int) -> float:
    """Algorithmic bytes per cell update of a run of `steps` BGK steps on an
    n x n lattice: the k-step launches, then (k = 8) one 4-step launch for a
    remainder of 4 or more steps, then one-step launches."""
    if k <= 1:
        return BYTES_PER_CELL[storage]
    nk, r = divmod(steps, k)
    n4 = 1 if k == 8 and r >= 4 else 0
    n1 = r - 4 * n4
    total = nk * k * tb_bytes_per_step(n, n, k=k) + n4 * 4 * tb_bytes_per_step(n, n, k=4) \
        + n1 * n * n * BYTES_PER_CELL[storage]
    return total / (n * n * steps)


def warm_kernels(ctx, n, storage, tau, u0):
    """One launch of each step kernel a run on this lattice shape uses (the
    8-step, 4-step and one-step launches of 13 steps), on a scratch lattice."""
    with tgv_lattice(ctx, n, storage, tau, u0) as scratch:
        scratch.step(13)
        scratch.sync()


def tgv_lattice(ctx, n, storage, tau, u0, rounding="sr", trt=None, collision=None, ks_viscosity=0.0):
    from chemsim_amd import Lattice, Params, capi
    from chemsim_amd import workloads as W
    rho, ux, uy = W.taylor_green(n, u0, np.float64 if storage == "f64" else np.float32)
    rnd = capi.CS_ROUND_STOCHASTIC if rounding == "sr" else capi.CS_ROUND_NEAREST
    p = Params(tau=tau, edge=capi.CS_EDGE_PERIODIC, rounding=rnd, seed=2024)
    if trt is not None:
        p.collision, p.tau, p.tau_minus = capi.CS_COLL_TRT, trt[0], trt[1]
    if collision is not None:
        p.collision, p.ks_viscosity = collision, ks_viscosity
    lat = Lattice(ctx, n, n, storage, p)
    lat.init_equilibrium(rho, ux, uy)
    return lat


def warmstart_leg(ctx, count=16):
    """Time to steady state (velocity residual 1e-6 over 100 steps) of the
    first `count` held-out channels (512 x 128, Zou-He inlet/outlet, 1-4
    obstacles, Re 20-100; the evaluation seed of the shipped network), from the
    cold start and from the network's prediction with the Chapman-Enskog
    non-equilibrium part.  Warm wall time includes the network's forward pass
    and the feature computation on the host."""
    from chemsim_amd import warmstart as WS
    from chemsim_amd.cnn import Net
    p, meta = WS.load_params(WS.WEIGHTS_PATH)
    chans = WS.make_channels(meta.get("seed_test", 2), count)
    out = {"channels": count, "weights": os.path.basename(WS.WEIGHTS_PATH), "tol": 1e-6}
    with Net(ctx, 128, 512, base=meta["base"], batch=8) as net:
        net.set_params(p)
        cold_s = cold_t = warm_s = warm_t = 0.0
        errs_u, errs_r = [], []
        for ch in chans:
            s, dt, _, ref = WS.time_to_steady(ctx, ch, WS.cold_state(ch))
            cold_s += s
            cold_t += dt
            t0 = time.perf_counter()
            pred = WS.predict(net, ch)
            t_net = time.perf_counter() - t0
            s, dt, _, _ = WS.time_to_steady(ctx, ch, pred, neq=True)
            warm_s += s
            warm_t += dt + t_net
            e_r, e_u = WS.relative_error(ch, pred, ref)
            errs_u.append(e_u)
            errs_r.append(e_r)
    out.update({"cold_steps_total": int(cold_s), "warm_steps_total": int(warm_s),
                "cold_seconds": round(cold_t, 3), "warm_seconds_incl_network": round(warm_t, 3),
                "step_ratio": round(cold_s / warm_s, 3), "time_ratio": round(cold_t / warm_t, 3),
                "error_u_median": round(float(np.median(errs_u)), 4),
                "error_rho_median": round(float(np.median(errs_r)), 4)})
    return out


_wake_lattices = {}


def wake_gpu(ctx, seconds=0.05):
    """About `seconds` of FP32 steps on a scratch TGV 2048^2 lattice (one per
    context, kept for the run), right before a timed region.  A GPU left idle
    runs its first milliseconds of work slower: 20 steps of TGV 4096^2 FP32
    timed after one second of idle host-side set-up (building a lattice)
    measure 190-202 GLUPS, 212-215 after 10 ms idle, 227-232 right after
    0.05-1 s of other lattices' steps (1,000 steps: 244).  0.05 s of 256 MiB
    copies instead brings the wall figure to 212-215 but leaves the kernels'
    time where it was (77 against 78 us per step; after steps: 71.5)."""
    scratch = _wake_lattices.get(id(ctx))
    if scratch is None:
        scratch = _wake_lattices[id(ctx)] = tgv_lattice(ctx, 2048, "f32", 0.8, 0.05)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        scratch.step(512)
        scratch.sync()


def close_wake_lattices():
    for scratch in _wake_lattices.values():
        scratch.close()
    _wake_lattices.clear()


def time_steps(lat, steps, wake=True):
    """Wall seconds and HIP-event milliseconds of `steps` steps of `lat`, on a
    GPU kept busy up to the start (wake_gpu) unless `wake` is false."""
    lat.sync()
    if wake:
        wake_gpu(lat.ctx)
    t0 = time.perf_counter()
    ms = lat.step_timed(steps)          # HIP events on the lattice's stream; synchronises
    lat.sync()
    return time.perf_counter() - t0, ms


def cpu_baseline(n, tau, u0, budget_s=12.0):
    """Oracle fused FP32 step (OpenMP over rows) on the same TGV lattice, a
    bounded number of steps; returns (GLUPS, threads, steps)."""
    from chemsim_amd import workloads as W
    from oracle import pyoracle as O
    threads = min(16, os.cpu_count() or 1)
    p = O.params(n, n, tau=tau, edge=O.EDGE_PERIODIC)
    rho, ux, uy = W.taylor_green(n, u0, np.float32)
    f = O.equilibrium(p, rho, ux, uy)
    t0 = time.perf_counter()
    f, used = O.run_omp(p, f, 1, threads=threads)
    one = time.perf_counter() - t0
    steps = int(max(1, min(200, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    O.run_omp(p, f, steps, threads=threads)
    dt = time.perf_counter() - t0
    return n * n * steps / dt / 1e9, used, steps


def cpu_unfused_sample(n, tau, u0, budget_s=6.0):
    """Oracle unfused FP32 step (one pass per reference operation, as ArrayFire
    runs them), one thread, on a 1024 x 1024 crop-sized TGV; returns GLUPS."""
    from chemsim_amd import workloads as W
    from oracle import pyoracle as O
    m = min(n, 1024)
    p = O.params(m, m, tau=tau, edge=O.EDGE_PERIODIC)
    rho, ux, uy = W.taylor_green(m, u0, np.float32)
    f = O.equilibrium(p, rho, ux, uy)
    t0 = time.perf_counter()
    f = O.run_unfused(p, f, 1)
    one = time.perf_counter() - t0
    steps = int(max(1, min(100, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    O.run_unfused(p, f, steps)
    return m * m * steps / (time.perf_counter() - t0) / 1e9, m, steps


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def tgv_8192_run(ctx, tau, u0, steps=200):
    """Taylor-Green 8192 x 8192 (67 M cells): FP32, BF16 and FP16 deviation
    storage (stochastic rounding) against the FP64 run of the same steps."""
    n = 8192
    cells = n * n
    ref = tgv_lattice(ctx, n, "f64", tau, u0)
    w64, _ = time_steps(ref, steps)
    r64, x64, y64 = ref.moments_f64()
    ref.close()
    den = np.sqrt((x64 ** 2 + y64 ** 2).sum())
    rden = np.sqrt(((r64 - 1.0) ** 2).sum())
    out = {"steps": steps, "f64_GLUPS": round(cells * steps / w64 / 1e9, 3)}
    for st, rnd in (("f32", "rn"), ("bf16", "sr"), ("f16", "sr")):
        lat = tgv_lattice(ctx, n, st, tau, u0, rnd)
        k = lat.temporal_steps()
        lat.step(5)
        w, ms = time_steps(lat, steps)
        lat.close()
        lat = tgv_lattice(ctx, n, st, tau, u0, rnd)
        lat.step(steps)
        r, x, y = lat.moments_f64()
        lat.close()
        out[st] = {"GLUPS": round(cells * steps / w / 1e9, 3), "kernel_us": round(ms / steps * 1e3, 1),
                   "hbm_frac": hbm_frac(step_bytes_per_cell(st, n, k), cells, ms / steps), "steps_per_launch": k,
                   "u_rel_L2_vs_f64": float(np.sqrt(((x - x64) ** 2 + (y - y64) ** 2).sum()) / den),
                   "rho_dev_rel_L2_vs_f64": float(np.sqrt(((r - r64) ** 2).sum()) / rden)}
        del r, x, y
    # the one-step FP32 kernel (72 B per cell update) on the same lattice
    from chemsim_amd import capi
    lat = tgv_lattice(ctx, n, "f32", tau, u0, "rn")
    capi.call("cs_set_temporal_steps", 0)
    try:
        lat.step(5)
        w, ms = time_steps(lat, steps)
    finally:
        capi.call("cs_set_temporal_steps", capi.CS_TEMPORAL_AUTO)
        lat.close()
    out["f32_one_step"] = {"GLUPS": round(cells * steps / w / 1e9, 3), "kernel_us": round(ms / steps * 1e3, 1),
                           "hbm_frac": hbm_frac(72, cells, ms / steps)}
    out["bf16_over_f32"] = round(out["bf16"]["GLUPS"] / out["f32"]["GLUPS"], 3)
    out["f16_over_f32"] = round(out["f16"]["GLUPS"] / out["f32"]["GLUPS"], 3)
    out["f16_over_f32_one_step"] = round(out["f16"]["GLUPS"] / out["f32_one_step"]["GLUPS"], 3)
    return out


def cavity_128_run(ctx, steps=10000):
    """Lid-driven cavity 128 x 128, tau 0.6, lid 0.1, half-way walls, from rest,
    10,000 steps: 16 K cells per step, 8 steps per launch on 32 x 32 regions
    (64 blocks), replayed from HIP graphs."""
    from chemsim_amd import Lattice, Params, capi
    from chemsim_amd import workloads as W
    n = 128
    p = Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_ux=0.0, lid_uy=0.1)
    with Lattice(ctx, n, n, "f32", p, W.cavity_flags(n, n)) as lat:
        lat.init_equilibrium(*W.rest_state(n, n))
        lat.step(10)
        w, ms = time_steps(lat, steps)
        ux, uy = lat.velocity()
        k = lat.temporal_steps()
    return {"steps": steps, "seconds": round(w, 4), "GLUPS": round(n * n * steps / w / 1e9, 3),
            "us_per_step": round(w / steps * 1e6, 2), "kernel_us_per_step_events": round(ms / steps * 1e3, 2),
            "steps_per_launch": k, "max_speed": round(float(np.sqrt(ux ** 2 + uy ** 2).max()), 5)}


def batch_cavities_run(ctx, steps=5000, per_lattice_steps=20, other_steps=1000):
    """512 independent lid-driven cavities (sides U{64..512}, tau U[0.55,1],
    lid U[0.02,0.1], seed 0), FP32, half-way walls: the autotuned batch launch
    for `steps` steps, the other launch kinds (persistent, tile list,
    temporal, persistent temporal) for `other_steps`, and one launch per
    lattice per step."""
    from chemsim_amd import Params, capi
    from chemsim_amd import workloads as W
    from chemsim_amd.batch import Batch, tune
    specs = W.batch_cavities(seed=0, count=512)
    b = Batch(ctx, [(s.side, s.side) for s in specs], "f32")
    for i, s in enumerate(specs):
        b.set_params(i, Params(tau=s.tau, boundary=capi.CS_BB_HALFWAY, lid_uy=s.u_lid))
        b.set_flags(i, W.cavity_flags(s.side, s.side))
        b.init_equilibrium(i, *W.rest_state(s.side, s.side))
    cache = os.environ.get("CHEMSIM_TUNING_CACHE",
                           os.path.join(os.path.expanduser("~"), ".cache", "chemsim_amd", "batch_tuning.json"))
    rec = tune(ctx, b, cache, trial_steps=24)
    ms = b.step(steps)
    best = b.tuning
    kinds = {capi.CS_LAUNCH_PERSISTENT: "persistent", capi.CS_LAUNCH_TILES: "tiles",
             capi.CS_LAUNCH_TEMPORAL: "temporal", capi.CS_LAUNCH_TEMPORAL_PERSISTENT: "temporal_persistent"}
    others = {}
    for kind in kinds:
        if kind == best.get("launch", 0):
            continue
        t = dict(best, launch=kind)
        if kind == capi.CS_LAUNCH_PERSISTENT:
            t.update(block_threads=1024, workgroups_per_cu=2, cells_per_thread=1)
        b.tuning = t
        others["GLUPS_" + kinds[kind]] = round(b.cells * other_steps / b.step(other_steps) / 1e6, 3)
    b.tuning = best
    ms_pl = b.step_per_lattice(per_lattice_steps)
    cells = b.cells
    # bytes the launch kind must move per cell update: 72 for one-step launches; for the
    # temporal kind every 64 x 64 region read and every cell written once per 8 steps
    bpc = 72.0
    if best.get("launch", 0) in (capi.CS_LAUNCH_TEMPORAL, capi.CS_LAUNCH_TEMPORAL_PERSISTENT):
        bpc = sum(tb_bytes_per_step(s.side, s.side, edge=TB_EDGE) for s in specs) / cells
    out = {"lattices": len(b), "cells": cells, "steps": steps,
           "GLUPS_batch": round(cells * steps / ms / 1e6, 3), "hbm_frac": hbm_frac(bpc, cells, ms / steps),
           "bytes_per_cell_update": round(bpc, 3), "launch": kinds[best.get("launch", 0)], **others,
           "GLUPS_per_lattice_launches": round(cells * per_lattice_steps / ms_pl / 1e6, 3),
           "seconds": round(ms / 1e3, 3), "tuning": rec["tuning"]}
    b.close()
    return out


def load_profile_traffic(kernel_key: str):
    """HBM bytes per launch of the step kernel measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE) and scaled by a copy kernel of known size; the
    summary lives in profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data.get(kernel_key)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--storage", default="f32", choices=["f32", "f64", "bf16", "f16", "f32dev"])
    ap.add_argument("--tau", type=float, default=0.8)
    ap.add_argument("--u0", type=float, default=0.05)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="skip the FP64 / 16-bit comparison runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-cavity batched run")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the library has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    from chemsim_amd import Context
    ctx = Context(0)
    n, storage = args.n, args.storage
    cells = n * n
    log(f"[bench] device {ctx.device_name()}; TGV {n}^2 {storage}; warmup {args.warmup}, steps {args.steps}")

    warm_kernels(ctx, n, storage, args.tau, args.u0)
    lat = tgv_lattice(ctx, n, storage, args.tau, args.u0)
    if args.warmup:
        lat.step(args.warmup)
    wall, ms = time_steps(lat, args.steps)
    s_final = lat.sums_f64()
    glups = cells * args.steps / wall / 1e9
    kernel_ms = ms / args.steps
    k_launch = lat.temporal_steps()
    temporal = k_launch > 1
    bpc = run_bytes_per_cell(storage, n, k_launch, args.steps)
    achieved = bpc * cells / (kernel_ms * 1e-3) / 1e9
    log(f"[bench] {glups:.2f} GLUPS wall, {kernel_ms * 1e3:.1f} us/step (events), {achieved:.0f} GB/s algorithmic")
    # four more timed regions of the same length, for the spread (value stays the first)
    reps = [glups]
    for _ in range(4):
        w_r, _ = time_steps(lat, args.steps)
        reps.append(cells * args.steps / w_r / 1e9)
    # the same region once more, started after one second of idle GPU (no wake-up)
    time.sleep(1.0)
    w_i, ms_i = time_steps(lat, args.steps, wake=False)

    copy_bytes = 1 << 30
    copy_ms = ctx.bench_copy(copy_bytes, 20)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"[bench] copy kernel {copy_gbs:.0f} GB/s (read + write)")

    extras = {"copy_kernel_GBs": round(copy_gbs, 1),
              "frac_of_copy": round(achieved / copy_gbs, 4),
              "frac_at_144B_per_cell": round(144 * cells / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
              if storage == "f32" else None,
              "mass_final": float(s_final[0]),
              "repeats_GLUPS": [round(r, 3) for r in reps],
              "median_GLUPS": round(float(np.median(reps)), 3),
              "idle_start": {"GLUPS": round(cells * args.steps / w_i / 1e9, 3),
                             "kernel_us": round(ms_i / args.steps * 1e3, 2),
                             "note": "the same timed region after 1 s of idle GPU, without the wake-up"}}
    if temporal:
        # the one-step kernel (vector lanes, 72 B per cell update) on the same lattice and steps
        from chemsim_amd import capi as _cap
        _cap.call("cs_set_temporal_steps", 0)
        try:
            w1, ms1 = time_steps(lat, args.steps)
        finally:
            _cap.call("cs_set_temporal_steps", _cap.CS_TEMPORAL_AUTO)
        extras["one_step_kernel"] = {"GLUPS": round(cells * args.steps / w1 / 1e9, 3),
                                     "kernel_us": round(ms1 / args.steps * 1e3, 2),
                                     "hbm_frac": hbm_frac(BYTES_PER_CELL[storage], cells, ms1 / args.steps),
                                     "bytes_per_cell": BYTES_PER_CELL[storage]}
        extras["temporal_blocking"] = {"steps_per_launch": k_launch, "region": [tb_region_edge(n, n)] * 2,
                                       "bytes_per_cell_update": round(bpc, 3),
                                       "speedup_vs_one_step": round(extras["one_step_kernel"]["kernel_us"]
                                                                    / (kernel_ms * 1e3), 3)}
        log(f"[bench] one-step kernel: {extras['one_step_kernel']}")

    if not args.no_extras and storage == "f32":
        # FP64 oracle run and 16-bit storage on the same workload, same steps
        from chemsim_amd import workloads as W
        lat.close()
        ref = tgv_lattice(ctx, n, "f64", args.tau, args.u0)
        m0 = ref.sums_f64()[0]
        w64, ms64 = time_steps(ref, args.steps)
        r64, x64, y64 = ref.moments_f64()
        m1 = ref.sums_f64()[0]
        ref.close()
        extras["f64"] = {"GLUPS": round(cells * args.steps / w64 / 1e9, 3),
                         "hbm_frac": hbm_frac(144, cells, ms64 / args.steps),
                         "mass_drift_rel": abs(m1 - m0) / m0}
        amp = W.taylor_green_decay(n, args.tau, args.steps)
        rho0, ux0, uy0 = W.taylor_green(n, args.u0)
        a64 = float((x64 * ux0 + y64 * uy0).sum() / (ux0 * ux0 + uy0 * uy0).sum())
        extras["f64"]["decay_vs_analytic_rel"] = abs(a64 / amp - 1)
        del rho0, ux0, uy0
        den = np.sqrt((x64 ** 2 + y64 ** 2).sum())
        rden = np.sqrt(((r64 - 1.0) ** 2).sum())
        for st, rnd in (("f32", "rn"), ("f32dev", "rn"), ("bf16", "sr"), ("f16", "sr"), ("bf16", "rn"),
                        ("f16", "rn")):
            l2 = tgv_lattice(ctx, n, st, args.tau, args.u0, rnd)
            k2 = l2.temporal_steps()
            m0 = l2.sums_f64()[0]
            w2, ms2 = time_steps(l2, args.steps)
            r, x, y = l2.moments_f64()
            m1 = l2.sums_f64()[0]
            l2.close()
            key = st if st in ("f32", "f32dev") else f"{st}_{rnd}"
            extras[key] = {"GLUPS": round(cells * args.steps / w2 / 1e9, 3),
                          "hbm_frac": hbm_frac(step_bytes_per_cell(st, n, k2), cells, ms2 / args.steps),
                          "steps_per_launch": k2,
                          "u_rel_L2_vs_f64": float(np.sqrt(((x - x64) ** 2 + (y - y64) ** 2).sum()) / den),
                          "rho_dev_rel_L2_vs_f64": float(np.sqrt(((r - r64) ** 2).sum()) / rden),
                          "mass_drift_rel": abs(m1 - m0) / m0}
            log(f"[bench] {key}: {extras[key]}")
        # TRT operator (lbm.rs:374-451) on the same workload: TRT::new(lambda 1/4, nu of tau 0.8)
        from chemsim_amd.lattice import trt_new
        tp, tm = trt_new(0.25, (args.tau - 0.5) / 3.0)
        lt = tgv_lattice(ctx, n, "f32", args.tau, args.u0, trt=(tp, tm))
        kt = lt.temporal_steps()
        wt, mst = time_steps(lt, args.steps)
        lt.close()
        extras["trt_f32"] = {"GLUPS": round(cells * args.steps / wt / 1e9, 3),
                             "hbm_frac": hbm_frac(step_bytes_per_cell("f32", n, kt), cells, mst / args.steps),
                             "steps_per_launch": kt,
                             "kernel_us": round(mst / args.steps * 1e3, 2), "tau_plus": tp, "tau_minus": tm}
        log(f"[bench] trt: {extras['trt_f32']}")
        # Regularized (lbm.rs:596-666) and KBC (lbm.rs:458-590, nu of tau 0.8) on the same workload
        from chemsim_amd import capi as _cc
        for key, coll in (("regularized_f32", _cc.CS_COLL_REGULARIZED), ("kbc_f32", _cc.CS_COLL_KBC)):
            lo = tgv_lattice(ctx, n, "f32", args.tau, args.u0, collision=coll,
                             ks_viscosity=(args.tau - 0.5) / 3.0)
            ko = lo.temporal_steps()
            wo, mso = time_steps(lo, args.steps)
            lo.close()
            extras[key] = {"GLUPS": round(cells * args.steps / wo / 1e9, 3),
                           "hbm_frac": hbm_frac(step_bytes_per_cell("f32", n, ko), cells, mso / args.steps),
                           "steps_per_launch": ko,
                           "kernel_us": round(mso / args.steps * 1e3, 2)}
            log(f"[bench] {key}: {extras[key]}")
        # render pass (render.rs) of the 4096^2 state, each display mode; bytes moved per
        # render: 36 B/cell populations + 16 B/cell field scratch + 3 B/pixel
        from chemsim_amd import capi as _c
        lr = tgv_lattice(ctx, n, "f32", args.tau, args.u0)
        rend = {}
        for name, mode in (("density", _c.CS_RENDER_DENSITY), ("speed", _c.CS_RENDER_SPEED),
                           ("velocity", _c.CS_RENDER_VELOCITY), ("momentum", _c.CS_RENDER_MOMENTUM)):
            lr.render(mode)
            t = min(lr.render(mode)[1] for _ in range(3))
            rend[name + "_ms"] = round(t, 3)
        rend["GBs_density"] = round(55 * cells / (rend["density_ms"] * 1e-3) / 1e9, 1)
        lr.close()
        extras["render_4096"] = rend
        log(f"[bench] render: {rend}")
        # frame recording (display.rs:157-185): 20 frames of 10 steps each, velocity mode, RGBA
        # frames (64 MiB) to pinned host memory; depth 1 = the reference's order (no overlap)
        recd = {}
        for depth in (1, 3):
            lrec = tgv_lattice(ctx, n, "f32", args.tau, args.u0)
            lrec.record(2, 1, _c.CS_RENDER_VELOCITY, True, depth, lambda k, a: True)
            _, st = lrec.record(20, 10, _c.CS_RENDER_VELOCITY, True, depth, lambda k, a: True)
            lrec.close()
            recd[f"depth{depth}"] = {"frames_per_s": round(20 / (st["wall_ms"] / 1e3), 1),
                                     "wall_ms": round(st["wall_ms"], 2), "gpu_ms": round(st["gpu_ms"], 2),
                                     "host_wait_ms": round(st["wait_ms"], 2)}
        extras["record_4096"] = recd
        log(f"[bench] record: {recd}")
        del x64, y64, r64
        extras["tgv_8192"] = tgv_8192_run(ctx, args.tau, args.u0)
        log(f"[bench] 8192^2: {extras['tgv_8192']}")
        extras["cavity_128"] = cavity_128_run(ctx)
        log(f"[bench] cavity 128^2: {extras['cavity_128']}")

    if not args.no_batch and not args.no_extras:
        extras["batch_512_cavities"] = batch_cavities_run(ctx)
        log(f"[bench] batch: {extras['batch_512_cavities']}")

    if not args.no_extras:
        # warm-start network: one training step of 8 samples of 128 x 512
        from chemsim_amd.warmstart_run import time_step
        extras["cnn_train_step"] = time_step(ctx, iters=10)
        log(f"[bench] cnn: {extras['cnn_train_step']}")
        extras["warmstart_16"] = warmstart_leg(ctx)
        log(f"[bench] warm start: {extras['warmstart_16']}")

    cpu = None
    if not args.no_cpu:
        cg, threads, csteps = cpu_baseline(n, args.tau, args.u0)
        cpu = {"value": round(cg, 5), "unit": "GLUPS", "cores": threads, "kind": "port",
               "sample": f"oracle fused FP32 step, OpenMP over rows, TGV {n}x{n}, {csteps} steps"}
        ug, um, usteps = cpu_unfused_sample(n, args.tau, args.u0)
        cpu["unfused_1thread"] = {"GLUPS": round(ug, 5),
                                  "sample": f"oracle unfused FP32 step (stream, bounce-back, equilibrium, BGK "
                                            f"passes), 1 thread, TGV {um}x{um}, {usteps} steps"}
        cpu["host"] = host_cpu()
        log(f"[bench] cpu baseline {cg:.4f} GLUPS on {threads} threads ({csteps} steps)")

    traffic = load_profile_traffic(f"k_step_tb_{storage}_periodic_fluid_{n}" if temporal
                                   else f"k_step_{storage}_periodic_fluid_{n}")
    line = {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "f32dev": "f32", "f64": "f64", "bf16": "bf16-storage/f32-math",
                  "f16": "f16-storage/f32-math"}[storage],
        "data": "synthetic (analytic Taylor-Green initial state)",
        "config": {"workload": f"D2Q9 BGK periodic Taylor-Green {n}x{n}, tau={args.tau}, u0={args.u0}, "
                               f"k=2pi/{n}, {storage} storage",
                   "lattice": [n, n], "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_cell": round(bpc, 3), "kernel_us": round(kernel_ms * 1e3, 2),
                     "kernel": (f"temporally blocked, {k_launch} steps per launch (a remainder of 4 or more steps: "
                                "one 4-step launch)") if temporal else "k_step / k_step_vec",
                     "per": "launch" if temporal else "step"},
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)
    close_wake_lattices()
    ctx.close()


if __name__ == "__main__":
    main()
