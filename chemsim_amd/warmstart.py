"""Learned warm start for steady channel flows.

A network (chemsim_amd/csrc/cnn.hip) predicts the steady (rho, u) of a
channel flow past circular obstacles from the obstacle mask, the Reynolds
number, position channels, a lubrication-theory estimate of the flow and the
distance to the nearest solid.  The lattice is initialised at the predicted
equilibrium (compute_equilibrium, src/lbm.rs:43-71, as src/main.rs:257-267
initialises a lattice), optionally plus the Chapman-Enskog non-equilibrium
part of the predicted velocity gradient, and the plain BGK step runs to the
residual tolerance, so the LBM itself decides when the flow is steady.  Training
data are steady states computed by the GPU LBM itself (a batch of lattices
in deviation arithmetic).
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

RE_MID, RE_HALF = 60.0, 40.0
# the shipped network weights (bf16) and their metadata
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart_c32.npz")
DIST_SCALE = 16.0         # signed distance feature = distance / DIST_SCALE, clipped to [-1, 1]


def lubrication(ch: W.Channel):
    """Lubrication-theory estimate of the steady flow, column by column: the
    inlet flux Q = u_in (h - 2) splits over the open gaps of a column in
    proportion to gap^3 (parallel Poiseuille resistances), each gap carries
    a Poiseuille profile, and the pressure falls by 12 nu Q / sum(gap^3) per
    column towards the outlet (p = 0 there; rho - 1 = 3 p).  Returns
    (rho - 1, uy) as h x w arrays; it ignores inertia and the entrance length,
    the network corrects it."""
    h, w = ch.h, ch.w
    solid = ch.flags() == W.WALL
    fluid = ~solid
    q = ch.u_in * (h - 2)
    rows = np.arange(h)[:, None] * np.ones((1, w), np.int64)
    # gaps: runs of fluid rows in a column, labelled uniquely over the lattice
    starts = fluid & np.vstack([np.ones((1, w), bool), ~fluid[:-1]])
    label = (np.cumsum(starts, axis=0) + (h + 1) * np.arange(w)[None, :]) * fluid
    n_lab = (h + 1) * w + h + 1
    g = np.bincount(label[fluid], minlength=n_lab).astype(np.float64)
    first = np.full(n_lab, h, np.int64)
    np.minimum.at(first, label[fluid], rows[fluid])
    g3 = np.bincount(label[fluid] // (h + 1), weights=g[label[fluid]] ** 2, minlength=w)  # sum of gap^3
    uy = np.zeros((h, w), np.float64)
    gl = g[label[fluid]]
    qk = q * gl ** 3 / g3[label[fluid] // (h + 1)]
    x = (rows[fluid] - first[label[fluid]] + 0.5) / gl
    uy[fluid] = 6.0 * (qk / gl) * x * (1 - x)
    dpdy = np.where(g3 > 0, 12.0 * ch.nu * q / np.maximum(g3, 1e-30), 0.0)
    p = np.concatenate([np.cumsum(dpdy[::-1])[::-1][1:], [0.0]])
    return 3.0 * np.repeat(p[None, :], h, 0) * fluid, uy


def rho_scale(ch: W.Channel) -> float:
    """rho - 1 of the inlet of the empty channel by the lubrication estimate
    (3 * 12 nu u_in L / (h - 2)^2): the density target is (rho - 1) / this."""
    return 36.0 * ch.nu * ch.u_in * (ch.w - 1) / (ch.h - 2) ** 2


def signed_distance(ch: W.Channel) -> np.ndarray:
    """Distance to the nearest wall or obstacle surface (walls half-way
    between the first fluid row and the wall row), negative inside solids."""
    h, w = ch.h, ch.w
    rr = np.arange(h, dtype=np.float64)[:, None] * np.ones((1, w))
    cc = np.ones((h, 1)) * np.arange(w, dtype=np.float64)[None, :]
    d = np.minimum(rr - 0.5, (h - 1.5) - rr)
    for (r0, c0, rad) in ch.circles:
        d = np.minimum(d, np.sqrt((rr - r0) ** 2 + (cc - c0) ** 2) - rad)
    return d


def features(ch: W.Channel) -> np.ndarray:
    """[h][w][8] network input: solid mask, normalised Re, row and column
    coordinates in [-1, 1], the lubrication estimate of uy / u_in and of
    (rho - 1) / rho_scale, the signed distance to the nearest solid, and the
    inlet (+1) / outlet (-1) columns."""
    h, w = ch.h, ch.w
    f = np.zeros((h, w, 8), np.float32)
    flags = ch.flags()
    f[..., 0] = flags == W.WALL
    f[..., 1] = (ch.re - RE_MID) / RE_HALF
    f[..., 2] = ((np.arange(h, dtype=np.float32) + 0.5) / h * 2 - 1)[:, None]
    f[..., 3] = ((np.arange(w, dtype=np.float32) + 0.5) / w * 2 - 1)[None, :]
    drho, uy = lubrication(ch)
    f[..., 4] = uy / ch.u_in
    f[..., 5] = drho / rho_scale(ch)
    f[..., 6] = np.clip(signed_distance(ch) / DIST_SCALE, -1, 1)
    f[..., 7] = (flags == W.INLET).astype(np.float32) - (flags == W.OUTLET)
    return f


def baseline(ch: W.Channel, feats: np.ndarray | None = None) -> np.ndarray:
    """The lubrication estimate in target units, [h][w][3]: the network
    predicts the difference to it (residual learning)."""
    f = features(ch) if feats is None else feats
    b = np.zeros((ch.h, ch.w, 3), np.float32)
    b[..., 0] = f[..., 5]
    b[..., 2] = f[..., 4]
    return b


def encode_targets(ch: W.Channel, rho, ux, uy, feats: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Targets ((rho - 1) / rho_scale, ux / u_in, uy / u_in) minus the
    lubrication baseline, and the loss weight (1 at fluid cells, 0 at solid
    ones)."""
    t = np.stack([(rho - 1.0) / rho_scale(ch), ux / ch.u_in, uy / ch.u_in], axis=-1).astype(np.float32)
    t -= baseline(ch, feats)
    wgt = (ch.flags() != W.WALL).astype(np.float32)
    t[wgt == 0] = 0.0
    return t, wgt


def decode_prediction(ch: W.Channel, y: np.ndarray, feats: np.ndarray | None = None):
    """(rho, ux, uy) from the network output (baseline added back)."""
    y = y[..., :3] + baseline(ch, feats)
    solid = ch.flags() == W.WALL
    rho = (1.0 + rho_scale(ch) * y[..., 0]).astype(np.float32)
    ux = (ch.u_in * y[..., 1]).astype(np.float32)
    uy = (ch.u_in * y[..., 2]).astype(np.float32)
    rho[solid] = 1.0
    ux[solid] = 0.0
    uy[solid] = 0.0
    return rho, ux, uy


def mirrored(ch: W.Channel) -> W.Channel:
    """The channel reflected across its centre line (same flow, ux negated)."""
    return W.Channel(w=ch.w, h=ch.h, circles=[(ch.h - 1 - r, c, rad) for (r, c, rad) in ch.circles], re=ch.re,
                     u_in=ch.u_in)


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
    t_set = time.perf_counter()
    for i, c in enumerate(chans):
        b.set_params(i, lbm_params(c))
        b.set_flags(i, c.flags())
        b.init_equilibrium(i, *cold_state(c))
        if (i + 1) % 500 == 0:
            log(f"[steady] {i + 1}/{n} lattices set up ({time.perf_counter() - t_set:.0f}s)")
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


def predict(net, ch: W.Channel):
    """The network's (rho, ux, uy) for one channel (features on the host,
    forward pass on the GPU)."""
    x = features(ch)
    y = net.forward(x[None])
    return decode_prediction(ch, y[0], x)


def relative_error(ch: W.Channel, pred, ref):
    """L2 errors of a prediction over fluid cells: (rho - 1) and |u|, relative."""
    fluid = ch.flags() != W.WALL
    r_p, ux_p, uy_p = (np.asarray(a, np.float64)[fluid] for a in pred)
    r_t, ux_t, uy_t = (np.asarray(a, np.float64)[fluid] for a in ref)
    e_rho = np.linalg.norm(r_p - r_t) / max(np.linalg.norm(r_t - 1.0), 1e-30)
    e_u = np.sqrt(((ux_p - ux_t) ** 2 + (uy_p - uy_t) ** 2).sum() / (ux_t ** 2 + uy_t ** 2).sum())
    return float(e_rho), float(e_u)
