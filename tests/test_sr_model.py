"""Stochastic rounding: the oracle's bit model reproduces gfx950's
v_cvt_sr_bf16_f32 / v_cvt_sr_f16_f32 on vectors recorded from the hardware
(tests/golden/sr_vectors.npz, made by tests/golden/sr_probe.hip), and the
rounding is unbiased."""
import os

import numpy as np

from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sr_vectors.npz")


def test_sr_model_matches_hardware_vectors():
    d = np.load(GOLD)
    bf = np.array([O.lib.orc_f32_to_bf16_sr(float(x), int(r)) for x, r in zip(d["x"], d["r"])], np.uint16)
    hf = np.array([O.lib.orc_f32_to_f16_sr(float(x), int(r)) for x, r in zip(d["x"], d["r"])], np.uint16)
    assert len(d["x"]) > 10000
    np.testing.assert_array_equal(bf, d["bf16"])
    np.testing.assert_array_equal(hf, d["f16"])


def test_sr_is_unbiased_and_brackets():
    rng = np.random.default_rng(3)
    for x in (0.0123456, -3.3e-5, 1.7e-7, 0.4444445):
        rs = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64)
        hb = np.array([O.lib.orc_f32_to_f16_sr(x, int(r)) for r in rs], np.uint16)
        vals = hb.view(np.float16).astype(np.float64)
        lo, hi = vals.min(), vals.max()
        assert lo <= x <= hi and len(np.unique(vals)) <= 2
        assert abs(vals.mean() - np.float32(x)) < 0.02 * max(hi - lo, 1e-30)
        bb = np.array([O.lib.orc_f32_to_bf16_sr(x, int(r)) for r in rs], np.uint16)
        bv = (bb.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        assert abs(bv.mean() - np.float32(x)) < 0.02 * max(bv.max() - bv.min(), 1e-30)


def test_sr_bits_schedule():
    key = O.lib.orc_sr_key(7, 11)
    R = O.sr_bits(key, 12345)
    # two directions share one 32-bit word: the word and the word << 16
    assert all(R[2 * j + 1] == (int(R[2 * j]) << 16) & 0xffffffff for j in range(4))
    assert len(set(int(v) >> 16 for v in R)) == 9
