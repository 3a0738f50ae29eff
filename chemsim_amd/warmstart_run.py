"""Train and evaluate the learned warm start for steady channel flows.

    python -m chemsim_amd.warmstart_run all --train 2000 --test 200 --out DIR

1. Steady states of `--train` + `--test` seeded channels (512x128, Zou-He
   inlet/outlet, 1-4 circular obstacles of radius 6-20, Re 20-100) from the
   cold start, in one batch of lattices (persistent launches), to the
   velocity residual `--data-tol`.
2. Adam on the training states (cosine learning rate), BF16 MFMA network.
3. For every held-out channel: steps and wall time of one lattice to the
   residual `--tol` from the cold start, from the network's prediction
   (equilibrium start, and equilibrium + Chapman-Enskog non-equilibrium
   start), and from the converged field itself (the bound of a perfect
   prediction).

Writes DIR/warmstart_c{base}.npz (bf16 weights + metadata) and
DIR/warmstart_eval.json.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import warmstart as WS
from .cnn import Net
from .lattice import Context


def log(msg: str) -> None:
    print(msg, flush=True)


relative_error = WS.relative_error


def predict(net: Net, chans):
    """Network prediction (rho, ux, uy) per channel and the forward time per channel."""
    x = np.stack([WS.features(c) for c in chans])
    t0 = time.perf_counter()
    y = net.forward(x)
    dt = (time.perf_counter() - t0) / len(chans)
    return [WS.decode_prediction(c, y[i], x[i]) for i, c in enumerate(chans)], dt


def evaluate(ctx: Context, net: Net, chans, storage: str = "f32dev", tol: float = 1e-6, interval: int = 100,
             max_steps: int = 400000, bound: bool = True):
    """Per-channel steps/seconds to the residual tol from the cold start and
    the warm starts; see the module docstring."""
    preds, t_inf = predict(net, chans)
    rows = []
    for i, (ch, pred) in enumerate(zip(chans, preds)):
        row = {"re": ch.re, "obstacles": len(ch.circles), "inference_s": t_inf}
        s, dt, res, ref = WS.time_to_steady(ctx, ch, WS.cold_state(ch), tol, interval, max_steps, storage)
        row["cold"] = {"steps": s, "seconds": dt, "residual": res}
        row["error_rho"], row["error_u"] = relative_error(ch, pred, ref)
        for name, neq in (("warm", False), ("warm_neq", True)):
            s, dt, res, _ = WS.time_to_steady(ctx, ch, pred, tol, interval, max_steps, storage, neq=neq)
            row[name] = {"steps": s, "seconds": dt + t_inf, "residual": res}
        if bound:
            exact = tuple(np.asarray(a, np.float32) for a in ref)
            for name, neq in (("exact", False), ("exact_neq", True)):
                s, dt, res, _ = WS.time_to_steady(ctx, ch, exact, tol, interval, max_steps, storage, neq=neq)
                row[name] = {"steps": s, "seconds": dt, "residual": res}
        rows.append(row)
        log(f"[eval {i + 1}/{len(chans)}] Re {ch.re:5.1f} err u {row['error_u']:.2e} rho {row['error_rho']:.2e} "
            f"steps cold {row['cold']['steps']} warm {row['warm']['steps']} warm+neq {row['warm_neq']['steps']}"
            + (f" exact {row['exact']['steps']} exact+neq {row['exact_neq']['steps']}" if bound else ""))
    return rows


def error_gain_curve(ctx: Context, net: Net, chans, alphas, storage: str = "f32dev", tol: float = 1e-6,
                     max_steps: int = 400000):
    """Steps to the residual tol from starts between the converged field and
    the network's prediction: start = exact + alpha (prediction - exact), with
    the Chapman-Enskog non-equilibrium part, for every alpha.  This keeps the
    spatial structure of the network's own error and scales its size, so it
    shows how the gain over the cold start grows as the prediction improves."""
    preds, _ = predict(net, chans)
    per = []
    for i, (ch, pred) in enumerate(zip(chans, preds)):
        s_cold, _, _, ref = WS.time_to_steady(ctx, ch, WS.cold_state(ch), tol, 100, max_steps, storage)
        exact = [np.asarray(a, np.float64) for a in ref]
        p = [np.asarray(a, np.float64) for a in pred]
        row = {"re": ch.re, "cold": s_cold, "error_u": relative_error(ch, pred, ref)[1],
               "error_rho": relative_error(ch, pred, ref)[0], "steps": {}}
        for a in alphas:
            start = tuple((e + a * (q - e)).astype(np.float32) for e, q in zip(exact, p))
            s, _, _, _ = WS.time_to_steady(ctx, ch, start, tol, 100, max_steps, storage, neq=True)
            row["steps"][str(a)] = s
        per.append(row)
        log(f"[curve {i + 1}/{len(chans)}] Re {ch.re:5.1f} cold {s_cold} " +
            " ".join(f"a={a}:{row['steps'][str(a)]}" for a in alphas))
    cold = sum(r["cold"] for r in per)
    summary = {"cases": len(per), "cold_steps_total": cold, "error_u_median": float(np.median([r["error_u"] for r in per])),
               "error_rho_median": float(np.median([r["error_rho"] for r in per])), "curve": []}
    for a in alphas:
        tot = sum(r["steps"][str(a)] for r in per)
        summary["curve"].append({"alpha": a, "error_u_median": a * summary["error_u_median"],
                                 "error_rho_median": a * summary["error_rho_median"],
                                 "steps_total": tot, "step_ratio": cold / max(tot, 1)})
    return {"summary": summary, "cases": per}


def summarize(rows, max_steps: int):
    """Ratios cold / warm of total steps and total seconds over the cases
    where both runs reached the tolerance, plus medians of per-case ratios."""
    out = {"cases": len(rows)}
    for name in ("warm", "warm_neq", "exact", "exact_neq"):
        if name not in rows[0]:
            continue
        ok = [r for r in rows if r["cold"]["steps"] < max_steps and r[name]["steps"] < max_steps]
        if not ok:
            continue
        cs = sum(r["cold"]["steps"] for r in ok)
        ws = sum(r[name]["steps"] for r in ok)
        ct = sum(r["cold"]["seconds"] for r in ok)
        wt = sum(r[name]["seconds"] for r in ok)
        out[name] = {
            "converged_cases": len(ok),
            "step_ratio_total": cs / max(ws, 1),
            "time_ratio_total": ct / wt,
            "step_ratio_median": float(np.median([r["cold"]["steps"] / max(r[name]["steps"], 1) for r in ok])),
            "time_ratio_median": float(np.median([r["cold"]["seconds"] / r[name]["seconds"] for r in ok])),
            "mean_steps": float(np.mean([r[name]["steps"] for r in ok])),
        }
    out["cold_mean_steps"] = float(np.mean([r["cold"]["steps"] for r in rows]))
    out["cold_unconverged"] = int(sum(r["cold"]["steps"] >= max_steps for r in rows))
    out["error_u_median"] = float(np.median([r["error_u"] for r in rows]))
    out["error_rho_median"] = float(np.median([r["error_rho"] for r in rows]))
    return out


def time_step(ctx: Context, base: int = 32, batch: int = 8, iters: int = 10) -> dict:
    """One training step (forward + loss + backward, no optimizer) of the
    network on `batch` random 128x512 inputs: milliseconds and the GEMM
    throughput against the BF16 dense MFMA peak (2.5 PFLOP/s)."""
    with Net(ctx, 128, 512, base=base, batch=batch) as net:
        x = np.random.default_rng(0).standard_normal((batch, 128, 512, 8)).astype(np.float32)
        net.forward(x)
        ms, flops = net.time_step(iters)
    tflops = flops / ms / 1e9
    return {"batch": batch, "ms": round(ms, 3), "gemm_tflops": round(tflops, 1),
            "frac_of_bf16_peak": round(tflops / 2500.0, 4), "gemm_gflop_per_step": round(flops / 1e9, 1)}


def dataset(ctx: Context, chans, args):
    """Training arrays from the steady states of `chans`; channels that did not
    reach the residual within --data-max-steps are left out (their fields are
    not steady states and would teach the network transients)."""
    fields, steps, res = WS.steady_states(ctx, chans, tol=args.data_tol, interval=100, chunk=args.chunk,
                                          max_steps=args.data_max_steps, storage=args.storage, log=log)
    keep = [i for i in range(len(chans)) if steps[i] >= 0]
    if len(keep) < len(chans):
        log(f"[data] {len(chans) - len(keep)} unconverged channels left out")
    chans = [chans[i] for i in keep]
    fields = [fields[i] for i in keep]
    n = len(chans) * (2 if args.mirror else 1)
    h, w = chans[0].h, chans[0].w
    x = np.empty((n, h, w, 8), np.float32)
    t = np.empty((n, h, w, 3), np.float32)
    wt = np.empty((n, h, w), np.float32)
    for i, (c, f) in enumerate(zip(chans, fields)):
        x[i] = WS.features(c)
        t[i], wt[i] = WS.encode_targets(c, *f, feats=x[i])
        if args.mirror:
            # reflection across the centre line: rows reversed, ux negated
            j = len(chans) + i
            x[j] = WS.features(WS.mirrored(c))
            t[j] = t[i, ::-1]
            t[j, ..., 1] *= -1
            wt[j] = wt[i, ::-1]
    return x, t, wt, steps, res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("cmd", choices=["all", "evaluate", "time", "curve"])
    ap.add_argument("--alphas", default="1,0.5,0.25,0.1,0.03,0.01,0",
                    help="error scales of the curve command (start = exact + alpha (network - exact))")
    ap.add_argument("--iters", type=int, default=10, help="forward+backward passes to time (cmd time)")
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--seed-train", type=int, default=1)
    ap.add_argument("--seed-test", type=int, default=2)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--train-seconds", type=float, default=1e9, help="stop training after this long")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--mirror", type=int, default=1, help="add the mirrored channels to the training set")
    ap.add_argument("--storage", default="f32dev")
    ap.add_argument("--eval-storage", default="f32dev")
    ap.add_argument("--data-tol", type=float, default=1e-6)
    ap.add_argument("--data-max-steps", type=int, default=200000)
    ap.add_argument("--chunk", type=int, default=4000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-steps", type=int, default=400000)
    ap.add_argument("--eval-count", type=int, default=None, help="evaluate only the first N held-out channels")
    ap.add_argument("--no-bound", action="store_true", help="skip the exact-field starts")
    ap.add_argument("--weights", default=None, help="weights to evaluate (cmd evaluate)")
    ap.add_argument("--out", default="warmstart_out")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    ctx = Context(0)
    log(f"[warmstart] device {ctx.device_name()}")
    if args.cmd == "time":
        print(json.dumps(time_step(ctx, args.base, args.batch, args.iters)), flush=True)
        ctx.close()
        return 0
    test = WS.make_channels(args.seed_test, args.test)
    if args.eval_count is not None:
        test = test[:args.eval_count]
    net = Net(ctx, 128, 512, base=args.base, batch=args.batch, seed=args.seed_train)
    meta = {"base": args.base, "height": 128, "width": 512}
    if args.cmd == "all":
        t0 = time.perf_counter()
        train = WS.make_channels(args.seed_train, args.train)
        x, t, w, steps, res = dataset(ctx, train, args)
        log(f"[data] {len(train)} steady states in {time.perf_counter() - t0:.0f}s; steps mean {steps.mean():.0f} "
            f"max {steps.max()}, unconverged {(steps < 0).sum()}")
        t1 = time.perf_counter()
        hist = []
        lr_final = args.lr * 0.05
        rng = np.random.default_rng(args.seed_train)
        net.set_dataset(x, t, w)
        n_samples = x.shape[0]
        del x, t, w
        for e in range(args.epochs):
            frac = e / max(1, args.epochs - 1)
            lr_e = lr_final + 0.5 * (args.lr - lr_final) * (1 + np.cos(np.pi * frac))
            loss = net.train_epoch(rng.permutation(n_samples).astype(np.int32), lr_e)
            hist.append(loss)
            log(f"[train] epoch {e + 1}/{args.epochs} lr {lr_e:.2e} loss {loss:.6f} "
                f"({time.perf_counter() - t1:.0f}s)")
            if time.perf_counter() - t1 > args.train_seconds:
                log("[train] time budget reached")
                break
        ms, flops = net.time_step(5)
        log(f"[train] forward+backward of {args.batch} samples: {ms:.2f} ms, {flops / ms / 1e9:.1f} TFLOP/s (GEMM)")
        meta.update({"step_ms": ms, "step_tflops": flops / ms / 1e9})
        meta.update({"train": args.train, "seed_train": args.seed_train, "epochs": len(hist),
                     "loss": hist, "data_steps_mean": float(steps.mean()), "train_seconds":
                     time.perf_counter() - t1})
        WS.save_params(os.path.join(args.out, f"warmstart_c{args.base}.npz"), net.get_params(), meta)
    else:
        params, meta = WS.load_params(args.weights or WS.WEIGHTS_PATH)
        net.set_params(params)
    if args.cmd == "curve":
        alphas = [float(a) for a in args.alphas.split(",")]
        out = error_gain_curve(ctx, net, test, alphas, args.eval_storage, args.tol, args.max_steps)
        with open(os.path.join(args.out, "warmstart_curve.json"), "w") as fh:
            json.dump(out, fh, indent=1)
        log(json.dumps(out["summary"]))
        net.close()
        ctx.close()
        return 0
    rows = evaluate(ctx, net, test, args.eval_storage, args.tol, 100, args.max_steps, bound=not args.no_bound)
    summary = summarize(rows, args.max_steps)
    summary.update({"storage": args.eval_storage, "tol": args.tol, "seed_test": args.seed_test})
    with open(os.path.join(args.out, "warmstart_eval.json"), "w") as fh:
        json.dump({"summary": summary, "meta": meta, "cases": rows}, fh, indent=1)
    log(json.dumps(summary))
    net.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
