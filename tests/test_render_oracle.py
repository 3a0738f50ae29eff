"""Pins of the render restatement (oracle/render_oracle.py, src/render.rs):
the z-score / logistic value, the hue orientation of the vector modes in
closed form, and the geometry overlay."""
import numpy as np

from oracle import pyoracle as O
from oracle import render_oracle as R


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def byte(v):
    return int(np.floor(256.0 * v + 0.5))


def state(rho, ux, uy):
    h, w = rho.shape
    return O.equilibrium(O.params(w, h), rho.astype(np.float32), ux.astype(np.float32), uy.astype(np.float32))


def test_scalar_checkerboard_density():
    """rho = 1 +- a: mean 1, std a, so z = +-1: red = round(256 sigmoid(+-1))."""
    h, w = 8, 10
    chk = ((np.arange(h)[:, None] + np.arange(w)[None, :]) % 2) * 2 - 1
    img = R.render(state(1 + 0.01 * chk, np.zeros((h, w)), np.zeros((h, w))), R.DENSITY)
    assert np.all(img[..., 1:] == 0)
    assert np.all(img[..., 0][chk > 0] == byte(sigmoid(1.0)))   # 187
    assert np.all(img[..., 0][chk < 0] == byte(sigmoid(-1.0)))  # 69


def test_vector_hue_orientation():
    """Directions +x, +y, -x, -y (phase 0, pi/2, pi, -pi/2; hue 1/2, 3/4, 1,
    1/4) with speeds U and 2U on a checkerboard: z = +-1 for |v|^2, V =
    sigmoid(z), saturation 0.8, so with m = 0.2 V the colours are
    (m, V, V), (0.6V, m, V), (V, m, m) and (0.6V, V, m)."""
    h, w, U = 8, 8, 0.02
    rr, cc = np.mgrid[0:h, 0:w]
    speed = np.where((rr + cc) % 2 == 0, U, 2 * U)
    quad = (rr >= h // 2) * 2 + (cc >= w // 2)
    dirs = {0: (1, 0), 1: (0, 1), 2: (-1, 0), 3: (0, -1)}
    ux = np.zeros((h, w))
    uy = np.zeros((h, w))
    for q, (dx, dy) in dirs.items():
        ux[quad == q] = dx * speed[quad == q]
        uy[quad == q] = dy * speed[quad == q]
    img = R.render(state(np.ones((h, w)), ux, uy), R.VELOCITY)
    for q in range(4):
        for big, z in ((False, -1.0), (True, 1.0)):
            V = sigmoid(z)
            m = 0.2 * V
            want = {0: (m, V, V), 1: (0.6 * V, m, V), 2: (V, m, m), 3: (0.6 * V, V, m)}[q]
            sel = (quad == q) & ((speed > 1.5 * U) == big)
            for ch in range(3):
                assert np.all(np.abs(img[sel][:, ch].astype(int) - byte(want[ch])) <= 1), (q, big, ch)


def test_geometry_overlay_and_speed():
    rng = np.random.default_rng(2)
    h, w = 12, 16
    f = state(1 + 0.01 * rng.standard_normal((h, w)), 0.02 * rng.standard_normal((h, w)),
              0.02 * rng.standard_normal((h, w)))
    flags = (rng.random((h, w)) < 0.2).astype(np.uint8)
    flags[0, 0] = O.LID
    img = R.render(f, R.SPEED, flags)
    solid = flags > 0
    assert np.all(img[solid] == (0, 0, 255))
    plain = R.render(f, R.SPEED, flags, overlay=False)
    assert np.array_equal(img[~solid], plain[~solid]) and not np.array_equal(img, plain)
    # speed: sqrt(vx^2 + vy^2) of the velocity, red channel only
    assert np.all(plain[..., 1:] == 0)


def test_hsv_matches_colorsys():
    """The sextant HSV -> RGB formula against Python's colorsys (an independent
    implementation of the published conversion) on random colours, including
    the sextant boundaries."""
    import colorsys
    rng = np.random.default_rng(3)
    h = np.concatenate([rng.random(4000), np.arange(7) / 6.0 * 0.999999]).astype(np.float32)
    s = np.concatenate([rng.random(4000), np.full(7, 0.8)]).astype(np.float32)
    v = np.concatenate([rng.random(4000), np.full(7, 0.6)]).astype(np.float32)
    r, g, b = R.hsv_to_rgb(h, s, v)
    want = np.array([colorsys.hsv_to_rgb(float(a), float(c), float(d)) for a, c, d in zip(h, s, v)])
    np.testing.assert_allclose(np.stack([r, g, b], 1), want, rtol=0, atol=2e-6)


def test_stats_are_population_mean_and_std():
    rng = np.random.default_rng(4)
    x = (1 + 0.01 * rng.standard_normal(10000)).astype(np.float32)
    mean, inv = R.stats(x)
    assert abs(float(mean) - float(np.mean(x.astype(np.float64)))) < 1e-7
    assert abs(1.0 / float(inv) / float(np.std(x.astype(np.float64), ddof=0)) - 1) < 1e-5
