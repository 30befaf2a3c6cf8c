"""Items rendered together in packed frames equal the items rendered alone.

An item is rendered alone at the same position relative to the 16-pixel
tile grid as in its frame (moved by its cell origin modulo 16), since the
split into tile-local fills depends on it. Moving an item's points by its
cell's origin still changes their float32 rounding, so the frames agree to a
tolerance: every channel within 1/255 and at least 99.9% of the values
identical.
"""
import numpy as np
import pytest

from pathfinder_amd import (ColorF, DestFramebuffer, RectF, Renderer, RendererOptions, Scene,
                            batch, synth)

pytestmark = pytest.mark.gpu


def _render(scene, w, h, bg):
    r = Renderer(RendererOptions(DestFramebuffer.full_window(w, h), bg))
    try:
        scene.build_and_render(r)
        return r.read_pixels().astype(np.int32)
    finally:
        r.close()


def _alone(item, bg, phase):
    """The item alone, moved by phase = (x, y) mod 16 of its cell, cropped."""
    px, py = phase
    s = Scene()
    s.set_view_box(RectF(0, 0, item.width + px, item.height + py))
    from pathfinder_amd.scene import ColorU, PathArrays
    a = item.arrays
    paints = np.asarray([s.push_paint(ColorU(*c)) for c in item.colors], np.uint32)
    moved = np.asarray(a.points, np.float32) + np.asarray([px, py], np.float32)
    s.push_draw_paths(PathArrays(moved, a.flags, a.contour_ends, a.path_contour_ends, paints,
                                 a.fill_rules))
    out = _render(s, item.width + px, item.height + py, bg)
    return out[py:py + item.height, px:px + item.width]


@pytest.mark.parametrize("kind", ["canvases", "glyphs"])
def test_packed_items_match_items_alone(kind):
    if kind == "canvases":
        items = synth.canvases(count=10, size_lo=48, size_hi=300, paths_lo=5, paths_hi=60)
        bg, fw, fh, align = ColorF(1, 1, 1, 1), 1024, 1024, 16
    else:
        items = synth.glyphs(count=300)
        bg, fw, fh, align = ColorF(0, 0, 0, 0), 512, 512, 1
    frames = batch.build_frames(items, fw, fh, 10 ** 9, align=align)
    print("case batch", kind, len(items), "items in", len(frames), "frames", flush=True)
    diffs = []
    for f in frames:
        px = _render(f.scene, f.width, f.height, bg)
        for i, cell in f.cells.items():
            if kind == "glyphs" and i % 10:
                continue  # every tenth glyph alone keeps the test short
            d = np.abs(batch.crop(px, cell) - _alone(items[i], bg, (cell[0] % 16, cell[1] % 16)))
            diffs.append(d.ravel())
    d = np.concatenate(diffs)
    print("max %d, identical %.5f" % (d.max(), np.mean(d == 0)), flush=True)
    assert d.max() <= 1
    assert np.mean(d == 0) >= 0.999
