"""bench.py's parts that need no GPU: more than one GPU is refused with a
message before anything touches a device, and the algorithmic bytes behind
`roofline.achieved` equal an explicit region-by-region count."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def _tiled_bytes_per_step(w, h, edge, k, esize=4):
    """Regions laid out one by one: origins every edge - 2k cells from 0, each
    reading its edge x edge cells (halo included, also beyond the lattice's
    edge); every lattice cell written once; 9 populations.  The interiors must
    cover each cell exactly once."""
    import numpy as np
    inner = edge - 2 * k
    cover = np.zeros((h, w), np.int32)
    regions = 0
    for y in range(0, h, inner):
        for x in range(0, w, inner):
            cover[y:y + inner, x:x + inner] += 1
            regions += 1
    assert (cover == 1).all()
    return (regions * edge * edge + w * h) * 9 * esize / k


def test_temporal_bytes_match_an_explicit_tiling():
    sys.path.insert(0, ROOT)
    import bench
    for w, h, edge, k in ((4096, 4096, 64, 8), (4096, 4096, 64, 4), (1000, 700, 64, 8),
                          (128, 128, 32, 8), (200, 130, 32, 8), (200, 130, 32, 4), (8, 8, 32, 8)):
        assert bench.tb_bytes_per_step(w, h, k=k, edge=edge) == _tiled_bytes_per_step(w, h, edge, k)
    # the 4096^2 figure the profiles quote: 1,694.6 MB per 8-step launch, 12.63 B per cell update
    per8 = _tiled_bytes_per_step(4096, 4096, 64, 8)
    assert round(8 * per8 / 1e6, 1) == 1694.6
    assert round(per8 / 4096 ** 2, 2) == 12.63


def test_run_bytes_follow_the_launch_split():
    """20 steps = two 8-step launches + one 4-step launch; 19 = two 8-step
    launches + three one-step launches (72 B per cell); 13 = 8 + 4 + 1."""
    sys.path.insert(0, ROOT)
    import bench
    n = 4096
    cells = n * n
    b8 = _tiled_bytes_per_step(n, n, 64, 8)
    b4 = _tiled_bytes_per_step(n, n, 64, 4)
    for steps, total in ((20, 16 * b8 + 4 * b4), (19, 16 * b8 + 3 * 72 * cells),
                         (13, 8 * b8 + 4 * b4 + 72 * cells), (1000, 1000 * b8)):
        assert abs(bench.run_bytes_per_cell("f32", n, 8, steps) - total / (cells * steps)) < 1e-9
    assert bench.run_bytes_per_cell("f16", n, 1, 20) == 36
