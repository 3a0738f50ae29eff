"""Learned warm start on held-out channels: the trained network's
prediction, started with the Chapman-Enskog non-equilibrium part, reaches the
residual 1e-6 in fewer LBM steps than the cold start (same lattice, same
tolerance and check interval)."""
import os

import numpy as np
import pytest

from chemsim_amd import warmstart as WS
from chemsim_amd.cnn import Net

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(WS.__file__), "weights", "warmstart_c32.npz")


@pytest.fixture(scope="module")
def net(ctx):
    params, meta = WS.load_params(WEIGHTS)
    n = Net(ctx, meta["height"], meta["width"], base=meta["base"], batch=4)
    n.set_params(params)
    yield n
    n.close()


def test_warm_start_beats_cold_start(ctx, net):
    chans = WS.make_channels(2, 6)          # the held-out seed
    x = np.stack([WS.features(c) for c in chans])
    y = net.forward(x)
    cold, warm = [], []
    for i, ch in enumerate(chans):
        pred = WS.decode_prediction(ch, y[i], x[i])
        s_cold, _, r_cold, _ = WS.time_to_steady(ctx, ch, WS.cold_state(ch), max_steps=300000)
        s_warm, _, r_warm, _ = WS.time_to_steady(ctx, ch, pred, max_steps=300000, neq=True)
        print(f"[warm start] Re {ch.re:.1f}: cold {s_cold} steps, warm {s_warm} steps", flush=True)
        assert r_cold <= 1e-6 and r_warm <= 1e-6
        cold.append(s_cold)
        warm.append(s_warm)
    assert all(w < c for w, c in zip(warm, cold))
    assert sum(cold) / sum(warm) > 1.2
