"""Learned warm start for steady channel flows.

A network (chemsim_amd/csrc/cnn.hip) predicts the steady (rho, u) of a
channel flow past circular obstacles from the obstacle mask, the Reynolds
number and position channels.  The lattice is initialised at the predicted
equilibrium (compute_equilibrium, src/lbm.rs:43-71, as src/main.rs:257-267
initialises a lattice) and the plain BGK step runs to the residual
tolerance, so the LBM stays the judge of the prediction.  Training data are
steady states computed by the GPU LBM itself (a batch of lattices in
deviation arithmetic).
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import asdict

import numpy as np

from . import capi
from .batch import Batch
from .lattice import Context, Lattice, Params
from . import workloads as W

RHO_SCALE = 0.01          # target channel 0 = (rho - 1) / RHO_SCALE
RE_MID, RE_HALF = 60.0, 40.0


def features(ch: W.Channel) -> np.ndarray:
    """[h][w][8] network input: solid mask, normalised Re, row and column
    coordinates in [-1, 1], the Poiseuille shape across the channel, inlet
    and outlet columns, and zero."""
    h, w = ch.h, ch.w
    f = np.zeros((h, w, 8), np.float32)
    flags = ch.flags()
    solid = (flags == W.WALL).astype(np.float32)
    f[..., 0] = solid
    f[..., 1] = (ch.re - RE_MID) / RE_HALF
    r = (np.arange(h, dtype=np.float32) + 0.5) / h * 2 - 1
    c = (np.arange(w, dtype=np.float32) + 0.5) / w * 2 - 1
    f[..., 2] = r[:, None]
    f[..., 3] = c[None, :]
    s = np.clip((np.arange(h, dtype=np.float32) - 0.5) / (h - 2), 0, 1)
    f[..., 4] = (4 * s * (1 - s))[:, None]
    f[..., 5] = (flags == W.INLET)
    f[..., 6] = (flags == W.OUTLET)
    return f


def encode_targets(ch: W.Channel, rho, ux, uy) -> tuple[np.ndarray, np.ndarray]:
    t = np.stack([(rho - 1.0) / RHO_SCALE, ux / ch.u_in, uy / ch.u_in], axis=-1).astype(np.float32)
    wgt = (ch.flags() != W.WALL).astype(np.float32)
    t[wgt == 0] = 0.0
    return t, wgt


def decode_prediction(ch: W.Channel, y: np.ndarray):
    solid = ch.flags() == W.WALL
    rho = (1.0 + RHO_SCALE * y[..., 0]).astype(np.float32)
    ux = (ch.u_in * y[..., 1]).astype(np.float32)
    uy = (ch.u_in * y[..., 2]).astype(np.float32)
    rho[solid] = 1.0
    ux[solid] = 0.0
    uy[solid] = 0.0
    return rho, ux, uy


def cold_state(ch: W.Channel):
    """Cold start: rho = 1 and the inlet velocity in every fluid cell."""
    solid = ch.flags() == W.WALL
    rho = np.ones((ch.h, ch.w), np.float32)
    ux = np.zeros((ch.h, ch.w), np.float32)
    uy = np.full((ch.h, ch.w), ch.u_in, np.float32)
    uy[solid] = 0.0
    return rho, ux, uy


def lbm_params(ch: W.Channel) -> Params:
    return Params(tau=ch.tau, boundary=capi.CS_BB_HALFWAY, inlet_ux=0.0, inlet_uy=ch.u_in, outlet_rho=1.0)


def make_channels(seed: int, count: int, w: int = 512, h: int = 128, u_in: float = 0.05):
    rng = np.random.default_rng(seed)
    return [W.random_channel(rng, w=w, h=h, u_in=u_in) for _ in range(count)]


def steady_states(ctx: Context, chans, tol: float = 1e-6, interval: int = 100, chunk: int = 4000,
                  max_steps: int = 600000, storage: str = "f32dev", log=print):
    """Steady states of many channels in one batch (persistent launches),
    from the cold start; a lattice leaves the batch once its residual over
    `interval` steps is <= tol.  Returns (fields, steps, residuals)."""
    n = len(chans)
    b = Batch(ctx, [(c.w, c.h) for c in chans], storage)
    for i, c in enumerate(chans):
        b.set_params(i, lbm_params(c))
        b.set_flags(i, c.flags())
        b.init_equilibrium(i, *cold_state(c))
    done = np.zeros(n, bool)
    steps = np.full(n, -1, np.int64)
    res = np.full(n, np.inf)
    t0 = time.perf_counter()
    taken = 0
    while taken < max_steps and not done.all():
        b.step(chunk - interval)
        b.snapshot_velocity()
        b.step(interval)
        taken += chunk
        r = b.residuals()
        newly = (~done) & (r <= tol)
        res[~done] = r[~done]
        for i in np.nonzero(newly)[0]:
            b.set_active(int(i), False)
            steps[i] = taken
        done |= newly
        if (taken // chunk) % 10 == 0:
            log(f"[steady] {taken} steps: {done.sum()}/{n} converged, max residual "
                f"{np.max(np.where(done, 0, r)):.2e}, {time.perf_counter() - t0:.0f}s")
    fields = [b.moments(i) for i in range(n)]
    b.close()
    return fields, steps, res


def train(net, x, t, wgt, epochs: int, lr: float = 1e-3, lr_final: float = 1e-4, seed: int = 0, log=print):
    """Adam over the device-resident training set, cosine-decayed learning rate."""
    net.set_dataset(x, t, wgt)
    rng = np.random.default_rng(seed)
    hist = []
    for e in range(epochs):
        frac = e / max(1, epochs - 1)
        lr_e = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
        order = rng.permutation(x.shape[0]).astype(np.int32)
        loss = net.train_epoch(order, lr_e)
        hist.append(loss)
        log(f"[train] epoch {e + 1}/{epochs} lr {lr_e:.2e} loss {loss:.5f}")
    return hist


def time_to_steady(ctx: Context, ch: W.Channel, init, tol: float = 1e-6, interval: int = 100,
                   max_steps: int = 600000, storage: str = "f32dev", neq: bool = False):
    """Steps and wall seconds for one lattice from `init` (rho, ux, uy) to
    residual <= tol; neq adds the Chapman-Enskog non-equilibrium part."""
    with Lattice(ctx, ch.w, ch.h, storage, lbm_params(ch), ch.flags()) as lat:
        if neq:
            lat.init_nonequilibrium(*init)
        else:
            lat.init_equilibrium(*init)
        lat.sync()
        t0 = time.perf_counter()
        steps, res = lat.run_to_residual(tol, interval, max_steps)
        dt = time.perf_counter() - t0
        fields = lat.moments_f64()
    return steps, dt, res, fields


def save_params(path: str, params: np.ndarray, meta: dict) -> None:
    """Parameters as bfloat16 bit patterns (the precision the kernels use) + metadata."""
    f = np.asarray(params, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    bits = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, params_bf16=bits, meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


def load_params(path: str):
    d = np.load(path)
    p = (d["params_bf16"].astype(np.uint32) << 16).view(np.float32)
    meta = json.loads(bytes(d["meta"]).decode())
    return p, meta
