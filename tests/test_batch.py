"""Packing of many scenes or glyphs into frames (pathfinder_amd.batch)."""
import numpy as np

from pathfinder_amd import batch, synth


def _check_frames(items, frames, fw, fh):
    seen = set()
    for f in frames:
        assert f.width == fw and f.height <= fh and f.height % 16 == 0
        occ = np.zeros((f.height, f.width), np.int32)
        for i, (x, y, w, h) in f.cells.items():
            assert i not in seen
            seen.add(i)
            assert w == items[i].width and h == items[i].height
            assert x >= 0 and y >= 0 and x + w <= f.width and y + h <= f.height
            occ[y:y + h, x:x + w] += 1
        assert occ.max() <= 1  # cells never overlap
    assert seen == set(range(len(items)))


def test_glyph_pages():
    items = synth.glyphs(count=600)
    for it in items:  # the glyph stays inside its cell
        p = it.arrays.points
        assert p.min() >= 0 and (p[:, 0] <= it.width).all() and (p[:, 1] <= it.height).all()
    frames = batch.build_frames(items, 1024, 512, 10 ** 9)
    assert len(frames) > 1
    _check_frames(items, frames, 1024, 512)


def test_canvas_frames_respect_tile_budget():
    items = synth.canvases(count=24, size_hi=256, paths_hi=80)
    budget = 20000
    frames = batch.build_frames(items, 2048, 2048, budget, align=16)
    _check_frames(items, frames, 2048, 2048)
    for f in frames:
        est = sum(items[i].tile_estimate() for i in f.cells)
        assert est <= budget or len(f.cells) == 1
    # every canvas's points lie in its cell once moved
    for f in frames:
        for i, (x, y, w, h) in f.cells.items():
            p = items[i].arrays.points
            assert (p >= 0).all() and (p[:, 0] <= w).all() and (p[:, 1] <= h).all()
