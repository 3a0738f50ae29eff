"""Writes tests/golden/step_fixtures.npz: seeded random D2Q9 states (FP32,
near equilibrium, with and without a solid mask, zero and periodic edges) and
the populations after 3 reference steps computed by the oracle.  The fixtures
anchor the oracle, a numpy transcription of src/lbm.rs and the HIP kernels to
the same bytes.  Run from the repository root:
    python tests/golden/make_step_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as O  # noqa: E402

CASES = [  # name, h, w, tau, edge, with mask
    ("r8", 8, 8, 0.6, O.EDGE_ZERO, False),
    ("r17x31_mask", 17, 31, 0.75, O.EDGE_ZERO, True),
    ("r40_periodic", 40, 40, 0.8, O.EDGE_PERIODIC, False),
    ("r33x20_mask_periodic", 33, 20, 1.3, O.EDGE_PERIODIC, True),
]
STEPS = 3


def state(rng, h, w):
    rho = 1 + 0.02 * rng.standard_normal((h, w))
    ux = 0.04 * rng.standard_normal((h, w))
    uy = 0.04 * rng.standard_normal((h, w))
    p = O.params(w, h)
    f = O.equilibrium(p, rho, ux, uy) + 0.003 * rng.standard_normal((9, h, w))
    return f.astype(np.float32)


def main():
    rng = np.random.default_rng(20260101)
    out = {}
    for name, h, w, tau, edge, mask in CASES:
        f0 = state(rng, h, w)
        flags = (rng.random((h, w)) < 0.2).astype(np.uint8) if mask else np.zeros((h, w), np.uint8)
        op = O.params(w, h, tau=tau, edge=edge)
        out[name + "_in"] = f0
        out[name + "_flags"] = flags
        out[name + "_meta"] = np.array([tau, edge, STEPS], np.float64)
        out[name + "_out"] = O.run(op, f0, STEPS, flags if mask else None)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "step_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
