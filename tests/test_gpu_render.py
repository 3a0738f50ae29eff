"""Device render pass (src/render.rs:7-178) against the numpy restatement:
every display mode, with and without the geometry overlay.  Bytes agree to
within one unit (exp / atan2 may differ by an ulp between libraries), on all
but a few pixels exactly."""
import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W
from oracle import pyoracle as O
from oracle import render_oracle as R

pytestmark = pytest.mark.gpu


def cases():
    n, m = 48, 80
    rho, ux, uy = W.taylor_green(n, 0.05, np.float32, m=m)
    yield "tgv", m, n, Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC), None, (rho, ux, uy), 30
    flags = W.cavity_flags(64, 64)
    yield "cavity", 64, 64, Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_uy=0.1), flags, W.rest_state(64, 64), 300


@pytest.mark.parametrize("mode", [capi.CS_RENDER_DENSITY, capi.CS_RENDER_SPEED, capi.CS_RENDER_VELOCITY,
                                  capi.CS_RENDER_MOMENTUM])
def test_render_matches_oracle(ctx, mode):
    for name, w, h, p, flags, init, steps in cases():
        with Lattice(ctx, w, h, "f32", p, flags) as lat:
            lat.init_equilibrium(*init)
            lat.step(steps)
            f = lat.download()
            for overlay in (True, False):
                img, ms = lat.render(mode, overlay)
                ref = R.render(f, mode, flags, overlay)
                d = np.abs(img.astype(int) - ref.astype(int))
                assert d.max() <= 1, (name, mode, overlay, d.max())
                assert (d > 0).sum() <= max(3, d.size // 1000), (name, mode, overlay, (d > 0).sum())
                assert ms > 0


@pytest.mark.parametrize("storage", ["f64", "f32dev", "f16"])
def test_render_other_storages(ctx, storage):
    """The same image (to one unit on a few pixels) from the other storages of the same state."""
    n = 64
    rho, ux, uy = W.taylor_green(n, 0.05, np.float32)
    p = Params(tau=0.8, edge=capi.CS_EDGE_PERIODIC)
    with Lattice(ctx, n, n, "f32", p) as a, Lattice(ctx, n, n, storage, p) as b:
        a.init_equilibrium(rho, ux, uy)
        b.init_equilibrium(rho, ux, uy)
        ia, _ = a.render(capi.CS_RENDER_VELOCITY)
        ib, _ = b.render(capi.CS_RENDER_VELOCITY)
        d = np.abs(ia.astype(int) - ib.astype(int))
        assert d.max() <= (1 if storage in ("f64", "f32dev") else 3), d.max()


@pytest.mark.parametrize("mode", [capi.CS_RENDER_DENSITY, capi.CS_RENDER_SPEED, capi.CS_RENDER_VELOCITY,
                                  capi.CS_RENDER_MOMENTUM])
def test_render_ragged_sizes(ctx, mode):
    """Lattices whose cell count is not a multiple of 4 (the pixel pass's
    tail), one cell wide or high (overlay rows wrapping inside a thread's
    pixels) and 1x1 (zero spread: black, as the restatement gives)."""
    rng = np.random.default_rng(17)
    for w, h in [(1, 1), (1, 6), (6, 1), (3, 5), (5, 3), (7, 9), (2, 3), (13, 11)]:
        flags = rng.choice(np.array([O.FLUID, O.WALL, O.LID], np.uint8), size=(h, w), p=[0.7, 0.2, 0.1])
        p = Params(tau=0.7, boundary=capi.CS_BB_HALFWAY, lid_uy=0.05)
        rho = (1 + 0.01 * rng.standard_normal((h, w))).astype(np.float32)
        ux = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
        uy = (0.02 * rng.standard_normal((h, w))).astype(np.float32)
        print(f"[render] {w}x{h} mode={mode}", flush=True)
        with Lattice(ctx, w, h, "f32", p, flags) as lat:
            lat.init_equilibrium(rho, ux, uy)
            lat.step(3)
            f = lat.download()
            for overlay in (True, False):
                img, _ = lat.render(mode, overlay)
                ref = R.render(f, mode, flags, overlay)
                d = np.abs(img.astype(int) - ref.astype(int))
                assert d.max() <= 1, (w, h, mode, overlay, d.max())
                assert (d > 0).sum() <= 3, (w, h, mode, overlay, (d > 0).sum())
