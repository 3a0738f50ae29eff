"""GIF recording (display::record with its encoder): the decoded GIF frames
equal the RGBA frames of cs_lattice_record exactly for the single-hue density
mode (<= 256 colours), and closely for the velocity mode (quantised)."""
import numpy as np
import pytest

from chemsim_amd import Lattice, Params, capi
from chemsim_amd import workloads as W
from chemsim_amd.gif import record_gif

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [capi.CS_RENDER_DENSITY, capi.CS_RENDER_VELOCITY])
def test_gif_frames_match_recorded_frames(ctx, tmp_path, mode):
    from PIL import Image, ImageSequence
    n, frames, spf = 64, 6, 20
    flags = W.cavity_flags(n, n)
    p = Params(tau=0.6, boundary=capi.CS_BB_HALFWAY, lid_uy=0.1)
    path = str(tmp_path / "cavity.gif")
    with Lattice(ctx, n, n, "f32", p, flags) as a, Lattice(ctx, n, n, "f32", p, flags) as b:
        a.init_equilibrium(*W.rest_state(n, n))
        b.init_equilibrium(*W.rest_state(n, n))
        stats = record_gif(a, path, frames, spf, mode, True, 3, frame_ms=40)
        rec, _ = b.record(frames, spf, mode, True, 3)
    assert stats["gif_frames"] == frames
    with Image.open(path) as im:
        got = [np.asarray(fr.convert("RGB")) for fr in ImageSequence.Iterator(im)]
    assert len(got) == frames
    for k in range(frames):
        want = rec[k, ..., :3]
        if mode == capi.CS_RENDER_DENSITY:
            assert len(np.unique(want.reshape(-1, 3), axis=0)) <= 256
            np.testing.assert_array_equal(got[k], want)
        else:
            # 256-colour quantisation of a many-colour frame: small on average,
            # larger on rare colours
            d = np.abs(got[k].astype(int) - want.astype(int))
            assert d.mean() < 3 and np.percentile(d, 99) < 32, (k, d.mean(), np.percentile(d, 99), d.max())
