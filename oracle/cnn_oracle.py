"""CPU oracle of the warm-start network (float64 numpy).

TEST INFRASTRUCTURE ONLY: the parity reference of chemsim_amd/csrc/cnn.hip; imported by
tests/ only.  The network has no counterpart in the reference (it is a
warm-start extension); this restatement is pinned by central finite
differences and by torch CPU autograd (tests/test_cnn_oracle.py).

With bf16=True the forward and backward passes round to bfloat16 (nearest
even) at the points where the GPU stores bf16 (weights, inputs, every
activation, the gradients fed to the GEMMs); everything else is float64, so
the GPU differs only by its FP32 accumulation order.
"""
from __future__ import annotations

import numpy as np


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_spec(c: int):
    """(cin, cout, relu) of the 15 convolutions in forward order."""
    return [(8, c, 1), (c, c, 1), (c, 2 * c, 1), (2 * c, 2 * c, 1), (2 * c, 4 * c, 1), (4 * c, 4 * c, 1),
            (4 * c, 8 * c, 1), (8 * c, 8 * c, 1), (12 * c, 4 * c, 1), (4 * c, 4 * c, 1), (6 * c, 2 * c, 1),
            (2 * c, 2 * c, 1), (3 * c, c, 1), (c, c, 1), (c, 8, 0)]


# Dilation of each convolution (the 3x3 taps sit d cells apart): the last
# encoder level and the bottleneck widen the receptive field to the length of
# the channel.
DILATION = [1, 1, 1, 1, 1, 2, 4, 8, 2, 1, 1, 1, 1, 1, 1]


def unflatten(params: np.ndarray, c: int):
    """Flat parameter vector -> list of (W [cout][9][cin], b [cout])."""
    out, off = [], 0
    for cin, cout, _ in layer_spec(c):
        nw = cout * 9 * cin
        W = params[off:off + nw].reshape(cout, 9, cin)
        off += nw
        b = params[off:off + cout]
        off += cout
        out.append((W, b))
    assert off == len(params)
    return out


def flatten(grads) -> np.ndarray:
    return np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])


def im2col(x: np.ndarray, sign: int = 1, dil: int = 1) -> np.ndarray:
    """x [N][H][W][C] -> [N*H*W][9][C]; tap t = 3(dy+1) + (dx+1) reads
    x[y + s*d*dy][x + s*d*dx] (zero outside)."""
    N, H, W, C = x.shape
    d = dil
    xp = np.pad(x, ((0, 0), (d, d), (d, d), (0, 0)))
    cols = []
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            sy, sx = sign * dy * d, sign * dx * d
            cols.append(xp[:, d + sy:d + sy + H, d + sx:d + sx + W, :])
    return np.stack(cols, axis=3).reshape(N * H * W, 9, C)


def conv(x, W, b, dil=1):
    N, H, Wd, C = x.shape
    col = im2col(x, 1, dil).reshape(N * H * Wd, -1)
    y = col @ W.reshape(W.shape[0], -1).T + b
    return y.reshape(N, H, Wd, -1)


def conv_back(x, W, dz, dil=1):
    """Gradients of y = conv(x, W, b, dil) for upstream dz: (dx, dW, db)."""
    N, H, Wd, C = x.shape
    cout = W.shape[0]
    col = im2col(x, 1, dil).reshape(N * H * Wd, -1)
    d2 = dz.reshape(-1, cout)
    dW = (d2.T @ col).reshape(W.shape)
    db = d2.sum(0)
    # dx[p][ci] = sum_{t,co} dz[p - d(t)][co] W[co][t][ci]
    colT = im2col(dz, -1, dil).reshape(N * H * Wd, 9 * cout)       # [p][t][co]
    Wd_ = W.transpose(2, 1, 0).reshape(C, 9 * cout)                  # [ci][t][co]
    dx = (colT @ Wd_.T).reshape(N, H, Wd, C)
    return dx, dW, db


def pool(x):
    N, H, W, C = x.shape
    return 0.25 * ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2]))


def pool_back(dy):
    return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(dy):
    N, H, W, C = dy.shape
    return dy.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


class Net:
    """Forward/backward of the encoder-decoder, same layer order as cnn.hip."""

    def __init__(self, params: np.ndarray, c: int, bf16_points: bool = False):
        self.c = c
        self.q = bf16 if bf16_points else (lambda a: a)
        self.layers = [(self.q(W), b) for W, b in unflatten(np.asarray(params, np.float64), c)]
        self.relu = [r for _, _, r in layer_spec(c)]

    def _conv(self, i, x):
        W, b = self.layers[i]
        z = conv(x, W, b, DILATION[i])
        if self.relu[i]:
            z = np.maximum(z, 0.0)
            return self.q(z)
        return z

    def forward(self, x):
        q = self.q
        a = {}
        x = q(np.asarray(x, np.float64))
        a["in0"] = x
        e1a = self._conv(0, x); a["in1"] = e1a
        e1 = self._conv(1, e1a); a["e1"] = e1
        p1 = q(pool(e1)); a["in2"] = p1
        e2a = self._conv(2, p1); a["in3"] = e2a
        e2 = self._conv(3, e2a); a["e2"] = e2
        p2 = q(pool(e2)); a["in4"] = p2
        e3a = self._conv(4, p2); a["in5"] = e3a
        e3 = self._conv(5, e3a); a["e3"] = e3
        p3 = q(pool(e3)); a["in6"] = p3
        ba = self._conv(6, p3); a["in7"] = ba
        bb = self._conv(7, ba); a["bb"] = bb
        c3 = np.concatenate([up(bb), e3], axis=3); a["in8"] = c3
        d3a = self._conv(8, c3); a["in9"] = d3a
        d3 = self._conv(9, d3a); a["d3"] = d3
        c2 = np.concatenate([up(d3), e2], axis=3); a["in10"] = c2
        d2a = self._conv(10, c2); a["in11"] = d2a
        d2 = self._conv(11, d2a); a["d2"] = d2
        c1 = np.concatenate([up(d2), e1], axis=3); a["in12"] = c1
        d1a = self._conv(12, c1); a["in13"] = d1a
        d1 = self._conv(13, d1a); a["in14"] = d1
        y = self._conv(14, d1)
        self.acts = a
        self.outs = {0: e1a, 1: e1, 2: e2a, 3: e2, 4: e3a, 5: e3, 6: ba, 7: bb, 8: d3a, 9: d3, 10: d2a, 11: d2,
                     12: d1a, 13: d1, 14: y}
        return y

    def loss_grad(self, x, t, w):
        """loss = sum w (y[..., :3] - t)^2 / (3 N H W) and all parameter gradients."""
        y = self.forward(x)
        N, H, W, _ = y.shape
        norm = 3.0 * N * H * W
        d = y[..., :3] - t
        loss = float((w[..., None] * d * d).sum() / norm)
        dy = np.zeros_like(y)
        dy[..., :3] = 2.0 * w[..., None] * d / norm
        grads = [None] * 15
        q = self.q
        a = self.acts

        def back(i, g):
            W, _ = self.layers[i]
            z = self.outs[i]
            dz = q(g * (z > 0)) if self.relu[i] else q(g)
            dx, dW, db = conv_back(a[f"in{i}"], W, dz, DILATION[i])
            grads[i] = (dW, db)
            return dx

        g = back(14, dy)
        g = back(13, g)
        g = back(12, g)                                  # d cat1
        g = q(g)
        c11 = self.outs[11].shape[3]
        dskip1 = g[..., c11:]
        g = up_back(g[..., :c11])
        g = back(11, g)
        g = back(10, g)                                  # d cat2
        g = q(g)
        c9 = self.outs[9].shape[3]
        dskip2 = g[..., c9:]
        g = up_back(g[..., :c9])
        g = back(9, g)
        g = back(8, g)                                   # d cat3
        g = q(g)
        c7 = self.outs[7].shape[3]
        dskip3 = g[..., c7:]
        g = up_back(g[..., :c7])
        g = back(7, g)
        g = back(6, g)                                   # d p3
        g = dskip3 + pool_back(q(g))
        g = back(5, g)
        g = back(4, g)                                   # d p2
        g = dskip2 + pool_back(q(g))
        g = back(3, g)
        g = back(2, g)                                   # d p1
        g = dskip1 + pool_back(q(g))
        g = back(1, g)
        back(0, g)
        return loss, flatten(grads)


def init_params(c: int, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    out = []
    for cin, cout, relu in layer_spec(c):
        W = rng.standard_normal((cout, 9, cin)) * np.sqrt(2.0 / (9 * cin)) * (1.0 if relu else 0.5)
        b = 0.05 * rng.standard_normal(cout)
        out.append(np.concatenate([W.ravel(), b]))
    return np.concatenate(out)
