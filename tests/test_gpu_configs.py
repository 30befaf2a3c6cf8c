"""The batched workloads through the oracle.

Glyph atlas pages (50,000 glyphs at 8-72 px on 4096 x 4096 pages) and
multi-canvas frames (512 canvases of 64^2-1024^2 on 8192-wide frames) are
built exactly as the benchmark builds them (pathfinder_amd.batch), rendered
on the GPU and compared with the oracle bit for bit: every glyph page, and
the canvas frames of the 12 largest and of the 179 smallest canvases. Every
frame of both workloads is also checked item by item against items rendered
alone.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, compare_frame, oracle_frame, product_frame
from pathfinder_amd import (ColorF, DestFramebuffer, RectF, Renderer, RendererOptions, Scene,
                            batch, synth)
from pathfinder_amd.scene import ColorU, PathArrays

pytestmark = pytest.mark.gpu

GLYPHS = dict(fw=4096, fh=4096, budget=6_000_000, align=1, bg=ColorF(0, 0, 0, 0))
CANVASES = dict(fw=8192, fh=8192, budget=14_000_000, align=16, bg=ColorF(1, 1, 1, 1))


def _say(*a):
    print(*a, flush=True)


_frames = {}


def frames(kind):
    if kind not in _frames:
        p = GLYPHS if kind == "glyphs" else CANVASES
        items = synth.glyphs() if kind == "glyphs" else synth.canvases()
        _frames[kind] = (items, batch.build_frames(items, p["fw"], p["fh"], p["budget"],
                                                   align=p["align"]))
    return _frames[kind]


@pytest.mark.parametrize("kind,index", [("glyphs", i) for i in range(5)] +
                         [("canvases", 0), ("canvases", 11)])
def test_batched_frame_bit_exact(kind, index, renderer_factory):
    items, fr = frames(kind)
    assert len(fr) == (5 if kind == "glyphs" else 12)
    f = fr[index]
    bg = (GLYPHS if kind == "glyphs" else CANVASES)["bg"]
    _say("case", kind, "frame", index, f.width, f.height, len(f.cells), "items")
    cmds, o = oracle_frame(f.scene, (f.width, f.height), bg, threads=16)
    r = renderer_factory(f.width, f.height, bg)
    product_frame(r, f.scene)
    report = compare_frame(r, o, cmds)
    o.close()
    _say(report)
    assert_bit_exact(report)


def _render(scene, w, h, bg):
    r = Renderer(RendererOptions(DestFramebuffer.full_window(w, h), bg))
    try:
        scene.build_and_render(r)
        return r.read_pixels().astype(np.int32)
    finally:
        r.close()


def _alone(item, bg, cell):
    """The item alone at its cell's position (points moved by the same origin,
    so the same float32 values as in the packed frame), cropped to the cell."""
    x, y, w, h = cell
    s = Scene()
    s.set_view_box(RectF(0, 0, x + w, y + h))
    a = item.arrays
    paints = np.asarray([s.push_paint(ColorU(*c)) for c in item.colors], np.uint32)
    moved = np.asarray(a.points, np.float32) + np.asarray([x, y], np.float32)
    s.push_draw_paths(PathArrays(moved, a.flags, a.contour_ends, a.path_contour_ends, paints,
                                 a.fill_rules))
    return batch.crop(_render(s, x + w, y + h, bg), cell)


@pytest.mark.parametrize("kind", ["glyphs", "canvases"])
def test_every_frame_matches_items_alone(kind):
    """Every frame of the workload, with sampled items (8 per frame) rendered
    alone at their own position: the crops must be identical (items do not
    interfere). An item moved to another position is not compared: the new
    float32 rounding of its points can flip dice's segment count of a curve
    (a 0.25 px flattening difference, up to 65/255 on a pixel)."""
    items, fr = frames(kind)
    bg = (GLYPHS if kind == "glyphs" else CANVASES)["bg"]
    rng = np.random.default_rng(5)
    compared = 0
    for k, f in enumerate(fr):
        _say("case", kind, "frame", k, "packed vs alone")
        px = _render(f.scene, f.width, f.height, bg)
        cells = sorted(f.cells)
        for i in rng.choice(cells, min(8, len(cells)), replace=False):
            cell = f.cells[int(i)]
            got = batch.crop(px, cell)
            want = _alone(items[int(i)], bg, cell)
            assert np.array_equal(got, want), (kind, k, int(i), cell,
                                               int(np.abs(got - want).max()))
            compared += got.size
    _say("compared", compared, "values")
