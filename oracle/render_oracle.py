"""CPU restatement of the render pass (src/render.rs:7-178 as src/main.rs:157-176
drives it), numpy float32.

TEST INFRASTRUCTURE ONLY: the parity reference of chemsim_amd/csrc/render.hip; imported
by tests/ only.

Parity note: af::mean_all / af::stdev_all are taken as the population mean and
standard deviation (division by N) accumulated in double, and af::hsv2rgb as
the published sextant formula (C = V S, X = C (1 - |(6H mod 2) - 1|),
m = V - C).  ArrayFire's own sources are not in the reference tree, so the
bytes are pinned by the closed-form cases in tests/test_render_oracle.py and
the GPU is compared with this restatement to within one unit of the last
byte (exp/atan2 may differ by an ulp between libraries).
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as O

DENSITY, SPEED, VELOCITY, MOMENTUM = range(4)
F32 = np.float32


def field(f, mode):
    """The display mode's field: (scalar value, phase or None), float32."""
    f = np.asarray(f, np.float32)
    if mode == DENSITY:
        return O.density(f), None
    vx, vy = O.velocity(f) if mode in (SPEED, VELOCITY) else O.momentum(f)
    if mode == SPEED:
        return np.sqrt(vx * vx + vy * vy), None
    return vx * vx + vy * vy, np.arctan2(vy, vx).astype(np.float32)


def stats(s):
    """(mean as f32, 1 / std as f32) from the population moments in f64."""
    s64 = s.astype(np.float64)
    n = s64.size
    mean = s64.sum() / n
    var = max((s64 * s64).sum() / n - mean * mean, 0.0)
    with np.errstate(divide="ignore"):
        return F32(mean), F32(1.0) / F32(np.sqrt(var))


def clamp01(v):
    return np.fmin(np.fmax(v, F32(0)), F32(1))


def hsv_to_rgb(hue, sat, val):
    c = val * sat
    h6 = hue * F32(6)
    x = c * (F32(1) - np.abs(np.fmod(h6, F32(2)) - F32(1)))
    m = val - c
    z = np.zeros_like(c)
    sel = [h6 < 1, h6 < 2, h6 < 3, h6 < 4, h6 < 5]
    r1 = np.select(sel, [c, x, z, z, x], c)
    g1 = np.select(sel, [x, c, c, x, z], z)
    b1 = np.select(sel, [z, z, x, c, c], x)
    return r1 + m, g1 + m, b1 + m


def to_byte(v):
    y = (F32(256) * v).astype(np.float64)
    return np.clip(np.floor(y + 0.5), 0, 255).astype(np.uint8)   # round half away from zero (y >= 0)


def render(f, mode, flags=None, overlay=True):
    s, phase = field(f, mode)
    avg, inv = stats(s)
    with np.errstate(invalid="ignore", over="ignore"):
        z = (s - avg) * inv
        val = clamp01(F32(1) / (F32(1) + np.exp(-z)))
    if mode in (VELOCITY, MOMENTUM):
        pi = F32(np.pi)
        hue = clamp01((phase + pi) * (F32(1.0 / np.pi) * F32(0.5)))
        sat = np.full_like(val, F32(0.8))
    else:
        hue = np.zeros_like(val)
        sat = np.ones_like(val)
    r, g, b = hsv_to_rgb(hue.astype(np.float32), sat.astype(np.float32), val.astype(np.float32))
    img = np.stack([to_byte(r), to_byte(g), to_byte(b)], axis=-1)
    if overlay and flags is not None:
        solid = (flags == O.WALL) | (flags == O.LID)
        img[solid] = (0, 0, 255)
    return img
