"""Learned warm start on held-out channels: the trained network's
prediction, started with the Chapman-Enskog non-equilibrium part, reaches the
residual 1e-6 in fewer LBM steps than the cold start (same lattice, same
tolerance and check interval).

Measured (first 6 held-out channels): 325,300 cold steps against 234,600
warm, 1.39x; on all 200 held-out channels 1.28x.  The target of a >= 3x
lower time to steady state is NOT met.  The second test shows why, on the
network's own error: scaling that error down (start = exact + alpha
(prediction - exact)) gains little per decade -- the measured curve on 16
channels is 1.34x at alpha = 1 (u error 7.7%), 1.93x at 0.1, 3.0x at 0.01
and 10.7x for the converged field itself -- because the channel's slowest
modes decay no faster from a small start error than from a large one."""
import numpy as np
import pytest

from chemsim_amd import warmstart as WS
from chemsim_amd.cnn import Net

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(ctx):
    params, meta = WS.load_params(WS.WEIGHTS_PATH)
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
    # measured 1.39x on these channels (3x is the unmet target)
    assert sum(cold) / sum(warm) >= 1.3


def test_gain_grows_slowly_as_the_error_shrinks(ctx, net):
    """Starts between the converged field and the prediction: fewer steps as
    alpha falls, but 10x less error still leaves most of the cold-start steps
    (measured on these two channels: 16,700 / 12,400 / 3,700 and
    63,400 / 47,800 / 7,600 steps at alpha 1 / 0.1 / 0)."""
    for ch in WS.make_channels(2, 2):
        s_cold, _, _, ref = WS.time_to_steady(ctx, ch, WS.cold_state(ch), max_steps=300000)
        pred = WS.predict(net, ch)
        steps = {}
        for a in (1.0, 0.1, 0.0):
            start = tuple((np.asarray(e, np.float64) + a * (np.asarray(q, np.float64) - e)).astype(np.float32)
                          for e, q in zip(ref, pred))
            steps[a] = WS.time_to_steady(ctx, ch, start, max_steps=300000, neq=True)[0]
        print(f"[warm start] Re {ch.re:.1f}: cold {s_cold}, alpha 1 / 0.1 / 0: {steps}", flush=True)
        assert steps[0.0] < steps[0.1] < steps[1.0] < s_cold
        assert s_cold / steps[0.1] < 3.0          # 10x lower error is still far from 3x fewer steps
        assert s_cold / steps[0.0] > 3.0          # the converged field itself clears it
