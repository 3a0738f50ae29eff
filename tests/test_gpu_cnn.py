"""Warm-start network kernels against the float64 oracle (oracle/cnn_oracle.py):
the MFMA GEMM tile, the forward pass and every parameter gradient."""
import numpy as np
import pytest

from chemsim_amd.cnn import Net, gemm_bf16
from oracle import cnn_oracle as K

pytestmark = pytest.mark.gpu


def to_bits(x):
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (256, 96, 288), (1024, 8, 864), (128, 200, 64)])
def test_gemm_tile(ctx, M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = to_bits(rng.standard_normal((M, K)))
    B = to_bits(rng.standard_normal((N, K)))
    Cg, ms = gemm_bf16(ctx, A, B)
    ref = from_bits(A) @ from_bits(B).T
    np.testing.assert_allclose(Cg, ref, rtol=1e-5, atol=1e-4 * np.sqrt(K))


def small(c=8, n=2, h=16, w=32, seed=11):
    rng = np.random.default_rng(seed)
    p = K.init_params(c, seed).astype(np.float32)
    x = rng.standard_normal((n, h, w, 8)).astype(np.float32)
    t = rng.standard_normal((n, h, w, 3)).astype(np.float32)
    wt = (rng.random((n, h, w)) < 0.8).astype(np.float32)
    return p, x, t, wt


@pytest.mark.parametrize("h,w", [(16, 32), (64, 128)])
def test_forward_matches_oracle(ctx, h, w):
    """(64, 128): the dilated bottleneck (taps 4 and 8 apart on an 8 x 16
    map) reads real neighbours, not only the zero border."""
    c, (p, x, t, wt) = 8, small(h=h, w=w)
    with Net(ctx, h, w, base=c, batch=2) as net:
        net.set_params(p)
        np.testing.assert_array_equal(net.get_params(), p)
        y = net.forward(x)
    ref = K.Net(p, c, bf16_points=True).forward(x)[..., :3]
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    assert err < 1e-2, err


@pytest.mark.parametrize("h,w,c", [(16, 32, 8), (64, 128, 8), (64, 128, 32)])
def test_gradients_match_oracle(ctx, h, w, c):
    """(16, 32): every convolution on the GEMM path; (64, 128): the LDS-tile
    kernels at the levels whose width is a multiple of 64 (forward, data and
    weight gradient), the GEMM below; base 32 adds output-channel blocks of
    32 and 64, several 32-channel chunks per tap and the virtual concatenations
    of 96 and 192 channels on the tile kernels."""
    _, (p, x, t, wt) = c, small(c=c, h=h, w=w, seed=12)
    with Net(ctx, h, w, base=c, batch=2) as net:
        net.set_params(p)
        loss = net.loss_grad(x, t, wt)
        g = net.get_grads()
    lref, gref = K.Net(p, c, bf16_points=True).loss_grad(x, t, wt)
    assert abs(loss - lref) < 1e-3 * lref
    off = 0
    for i, (cin, cout, _) in enumerate(K.layer_spec(c)):
        n = cout * 9 * cin + cout
        gi, ri = g[off:off + n], gref[off:off + n]
        err = np.linalg.norm(gi - ri) / max(np.linalg.norm(ri), 1e-30)
        assert err < 3e-2, (i, err)
        off += n


def test_training_reduces_loss(ctx):
    c, (p, x, t, wt) = 8, small(seed=13)
    with Net(ctx, 16, 32, base=c, batch=2, seed=5) as net:
        l0 = net.loss_grad(x, t, wt)
        for _ in range(30):
            net.adam_step(3e-3)
            l1 = net.loss_grad(x, t, wt)
        assert l1 < 0.8 * l0
