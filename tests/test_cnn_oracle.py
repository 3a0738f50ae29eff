"""Pins of the float64 network oracle (oracle/cnn_oracle.py): central finite
differences and torch CPU autograd of the same architecture."""
import numpy as np
import pytest

from oracle import cnn_oracle as K


def small_case(c=8, n=2, h=16, w=32, seed=1):
    rng = np.random.default_rng(seed)
    p = K.init_params(c, seed)
    x = rng.standard_normal((n, h, w, 8))
    t = rng.standard_normal((n, h, w, 3))
    wt = (rng.random((n, h, w)) < 0.8).astype(np.float64)
    return p, x, t, wt


def test_gradients_match_finite_differences():
    c = 8
    p, x, t, wt = small_case(c)
    loss, g = K.Net(p, c).loss_grad(x, t, wt)
    rng = np.random.default_rng(7)
    idx = rng.choice(len(p), 40, replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (K.Net(pp, c).loss_grad(x, t, wt)[0] - K.Net(pm, c).loss_grad(x, t, wt)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    c = 8
    p, x, t, wt = small_case(c, seed=3)
    loss_np, g_np = K.Net(p, c).loss_grad(x, t, wt)
    tensors = []
    for (W, b) in K.unflatten(p, c):
        cout, _, cin = W.shape
        tw = torch.tensor(W.reshape(cout, 3, 3, cin).transpose(0, 3, 1, 2).copy(), dtype=torch.float64,
                          requires_grad=True)
        tb = torch.tensor(b, dtype=torch.float64, requires_grad=True)
        tensors.append((tw, tb))
    X = torch.tensor(x.transpose(0, 3, 1, 2).copy(), dtype=torch.float64)

    def cv(i, z, relu=True):
        d = K.DILATION[i]
        y = F.conv2d(z, tensors[i][0], tensors[i][1], padding=d, dilation=d)
        return torch.relu(y) if relu else y

    pool = lambda z: F.avg_pool2d(z, 2)
    up = lambda z: F.interpolate(z, scale_factor=2, mode="nearest")
    e1 = cv(1, cv(0, X))
    e2 = cv(3, cv(2, pool(e1)))
    e3 = cv(5, cv(4, pool(e2)))
    bb = cv(7, cv(6, pool(e3)))
    d3 = cv(9, cv(8, torch.cat([up(bb), e3], 1)))
    d2 = cv(11, cv(10, torch.cat([up(d3), e2], 1)))
    d1 = cv(13, cv(12, torch.cat([up(d2), e1], 1)))
    y = cv(14, d1, relu=False)
    T = torch.tensor(t.transpose(0, 3, 1, 2).copy())
    Wt = torch.tensor(wt)[:, None]
    loss = (Wt * (y[:, :3] - T) ** 2).sum() / (3 * x.shape[0] * x.shape[1] * x.shape[2])
    loss.backward()
    assert abs(loss.item() - loss_np) < 1e-12 * max(1, abs(loss_np))
    g_t = np.concatenate([np.concatenate([tw.grad.numpy().transpose(0, 2, 3, 1).ravel(), tb.grad.numpy()])
                          for tw, tb in tensors])
    np.testing.assert_allclose(g_np, g_t, rtol=1e-9, atol=1e-12)


def test_bf16_points_stay_close_to_f64():
    c = 8
    p, x, t, wt = small_case(c, seed=5)
    l64, g64 = K.Net(p, c).loss_grad(x, t, wt)
    lbf, gbf = K.Net(p, c, bf16_points=True).loss_grad(x, t, wt)
    assert abs(lbf - l64) < 0.05 * l64
    assert np.linalg.norm(gbf - g64) < 0.05 * np.linalg.norm(g64)
    assert K.bf16(np.float64(1.0 + 2 ** -9)) == 1.0 and K.bf16(np.float64(1.0 + 3 * 2 ** -9)) == 1.0 + 2 ** -7


@pytest.mark.parametrize("dil", [1, 2, 3])
def test_dilated_conv_and_backward_match_torch(dil):
    """One 3x3 convolution with dilation d on a map larger than its reach:
    forward, weight/bias and input gradients against torch."""
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    rng = np.random.default_rng(dil)
    x = rng.standard_normal((2, 11, 14, 8))
    W = rng.standard_normal((5, 9, 8))
    b = rng.standard_normal(5)
    dz = rng.standard_normal((2, 11, 14, 5))
    y = K.conv(x, W, b, dil)
    dx, dW, db = K.conv_back(x, W, dz, dil)
    X = torch.tensor(x.transpose(0, 3, 1, 2).copy(), requires_grad=True)
    TW = torch.tensor(W.reshape(5, 3, 3, 8).transpose(0, 3, 1, 2).copy(), requires_grad=True)
    TB = torch.tensor(b, requires_grad=True)
    Y = F.conv2d(X, TW, TB, padding=dil, dilation=dil)
    np.testing.assert_allclose(y, Y.detach().numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-12)
    (Y * torch.tensor(dz.transpose(0, 3, 1, 2).copy())).sum().backward()
    np.testing.assert_allclose(dx, X.grad.numpy().transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dW, TW.grad.numpy().transpose(0, 2, 3, 1).reshape(5, 9, 8), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(db, TB.grad.numpy(), rtol=1e-12, atol=1e-12)
