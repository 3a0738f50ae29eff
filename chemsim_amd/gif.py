"""Animated-GIF recording of a lattice: display::record (src/display.rs:157-185)
with its encoder.

The reference encodes each RGBA frame with the `gif` crate (0.10.1,
`Frame::from_rgba`: NeuQuant colour quantisation to 256 colours, then LZW).
That crate is not available here; this writer quantises each frame with
Pillow (median cut, no dithering) and writes GIF89a with Pillow's encoder.
The frames themselves come from cs_lattice_record (steps and render pass on
the GPU, frames copied out while the next ones are computed), and the
quantisation runs on the host inside the frame callback, overlapped with the
GPU work of the following frames.  A frame with at most 256 colours (every
scalar display mode: density and speed render in one hue) round-trips
exactly.
"""
from __future__ import annotations

from . import capi


def record_gif(lattice, path: str, frames: int, steps_per_frame: int = 1, mode: int = capi.CS_RENDER_DENSITY,
               overlay: bool = True, depth: int = 3, frame_ms: int = 0) -> dict:
    """record(initial, steps, encoder) for a Lattice: `frames` times
    `steps_per_frame` steps, the render pass, and one GIF frame each, written
    to `path` (infinite loop, `frame_ms` per frame).  Returns the recording
    statistics of Lattice.record."""
    from PIL import Image

    images = []

    def on_frame(k, rgba):
        img = Image.fromarray(rgba[..., :3].copy(), "RGB")
        images.append(img.quantize(colors=256, method=Image.Quantize.MEDIANCUT, dither=Image.Dither.NONE))
        return True

    _, stats = lattice.record(frames, steps_per_frame, mode, overlay, depth, on_frame)
    if images:
        images[0].save(path, save_all=True, append_images=images[1:], loop=0, duration=max(int(frame_ms), 0),
                       optimize=False, disposal=1)
    stats["gif_frames"] = len(images)
    return stats
