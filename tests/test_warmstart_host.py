"""Host-side logic of the learned warm start (no GPU): input features,
target encoding, mirror augmentation, the lubrication estimate, the training
set's assembly and the summary of warm-against-cold runs."""
import numpy as np

from chemsim_amd import warmstart as WS
from chemsim_amd import workloads as W


def test_lubrication_empty_channel_is_poiseuille():
    ch = W.Channel(w=64, h=34, circles=[], re=40.0, u_in=0.04)
    drho, uy = WS.lubrication(ch)
    H = ch.h - 2
    x = (np.arange(H) + 0.5) / H
    np.testing.assert_allclose(uy[1:-1, 5], 6 * ch.u_in * x * (1 - x), rtol=1e-12)
    assert np.all(uy[[0, -1]] == 0)
    # pressure gradient 12 nu Q / H^3 per column, zero at the outlet column
    np.testing.assert_allclose(drho[10, :], 36 * ch.nu * ch.u_in * H * (ch.w - 1 - np.arange(ch.w)) / H ** 3,
                               rtol=1e-12)
    np.testing.assert_allclose(drho[10, 0], WS.rho_scale(ch), rtol=1e-12)


def test_lubrication_flux_split():
    """An obstacle splits a column into gaps; the flux through the column is
    the inlet flux and the gaps share it as gap^3."""
    ch = W.Channel(w=80, h=66, circles=[(24, 40, 10)], re=30.0, u_in=0.05)
    drho, uy = WS.lubrication(ch)
    q = ch.u_in * (ch.h - 2)
    # sum of the discrete Poiseuille profile (midpoint rule) is q (1 - 1/(g^2)) per gap
    np.testing.assert_allclose(uy[:, 40].sum(), q, rtol=2e-3)
    assert drho[5, 0] > drho[5, -20] > 0 and np.all(np.diff(drho[5]) <= 1e-15)
    solid = ch.flags() == W.WALL
    assert np.all(uy[solid] == 0) and np.all(drho[solid] == 0)


def test_features_and_targets_mirror():
    ch = WS.make_channels(11, 1)[0]
    m = WS.mirrored(ch)
    np.testing.assert_array_equal(m.flags(), ch.flags()[::-1])
    f, fm = WS.features(ch), WS.features(m)
    for k in (0, 1, 3, 4, 5, 6, 7):
        np.testing.assert_allclose(fm[::-1, :, k], f[..., k], atol=1e-6)
    np.testing.assert_allclose(fm[::-1, :, 2], -f[..., 2], atol=1e-6)
    rng = np.random.default_rng(0)
    rho = 1 + 0.01 * rng.standard_normal((ch.h, ch.w))
    ux, uy = 0.05 * rng.standard_normal((2, ch.h, ch.w))
    t, wgt = WS.encode_targets(ch, rho, ux, uy)
    r2, x2, y2 = WS.decode_prediction(ch, t)
    # the mirrored channel's targets are the reflected ones with ux negated
    tm, _ = WS.encode_targets(m, rho[::-1], -ux[::-1], uy[::-1])
    np.testing.assert_allclose(tm[::-1] * [1, -1, 1], t, atol=1e-5)
    fluid = wgt > 0
    np.testing.assert_allclose(r2[fluid], rho[fluid], rtol=1e-6)
    np.testing.assert_allclose(x2[fluid], ux[fluid], rtol=1e-5, atol=1e-8)
    np.testing.assert_allclose(y2[fluid], uy[fluid], rtol=1e-5, atol=1e-8)
    assert np.all(r2[~fluid] == 1) and np.all(x2[~fluid] == 0)


def test_signed_distance():
    ch = W.Channel(w=40, h=30, circles=[(15, 20, 5)], re=30.0)
    d = WS.signed_distance(ch)
    assert abs(d[15, 20] + 5) < 1e-12 and abs(d[15, 28] - 3) < 1e-12
    assert abs(d[1, 2] - 0.5) < 1e-12


def test_dataset_leaves_out_unconverged_channels(monkeypatch):
    """Channels whose steady-state run did not reach the residual (steps < 0)
    give no training sample; mirroring doubles the rest, the mirrored half
    being the first half reflected with ux negated."""
    import types
    from chemsim_amd import warmstart_run as R
    chans = [W.Channel(w=48, h=20, circles=[(9, 20 + 4 * k, 3)], re=30.0, u_in=0.05) for k in range(3)]
    rng = np.random.default_rng(1)
    fields = [(1 + 0.01 * rng.standard_normal((c.h, c.w)), *(0.05 * rng.standard_normal((2, c.h, c.w))))
              for c in chans]
    steps = np.array([400, -1, 700])
    monkeypatch.setattr(R.WS, "steady_states", lambda *a, **k: (fields, steps, np.zeros(3)))
    args = types.SimpleNamespace(data_tol=1e-6, chunk=100, data_max_steps=1000, storage="f32dev", mirror=True)
    x, t, wt, s, _ = R.dataset(None, chans, args)
    assert x.shape == (4, 20, 48, 8) and t.shape == (4, 20, 48, 3) and wt.shape == (4, 20, 48)
    for i, c in zip((0, 1), (0, 2)):
        t_ref, w_ref = WS.encode_targets(chans[c], *fields[c], feats=WS.features(chans[c]))
        np.testing.assert_array_equal(t[i], t_ref)
        np.testing.assert_array_equal(wt[i], w_ref)
        np.testing.assert_array_equal(x[i], WS.features(chans[c]))
        np.testing.assert_array_equal(t[2 + i], t_ref[::-1] * np.float32([1, -1, 1]))
        np.testing.assert_array_equal(x[2 + i], WS.features(WS.mirrored(chans[c])))
    assert list(s) == [400, -1, 700]


def test_summarize_counts_only_cases_both_runs_finished():
    from chemsim_amd import warmstart_run as R
    run = lambda steps, sec: {"steps": steps, "seconds": sec}  # noqa: E731
    rows = [{"cold": run(900, 0.9), "warm": run(300, 0.4), "error_u": 0.1, "error_rho": 0.2},
            {"cold": run(600, 0.6), "warm": run(200, 0.2), "error_u": 0.3, "error_rho": 0.1},
            {"cold": run(1000, 1.0), "warm": run(100, 0.1), "error_u": 0.2, "error_rho": 0.3}]
    out = R.summarize(rows, max_steps=1000)
    w = out["warm"]
    assert out["cases"] == 3 and w["converged_cases"] == 2 and out["cold_unconverged"] == 1
    assert w["step_ratio_total"] == 1500 / 500 and abs(w["time_ratio_total"] - 1.5 / 0.6) < 1e-12
    assert w["step_ratio_median"] == 3.0 and w["mean_steps"] == 250.0
    assert out["error_u_median"] == 0.2 and out["error_rho_median"] == 0.2
