"""Pins of the oracle's clip and composite stages against independent
statements of what they compute.

- Clip: fill.cs.glsl:84-86 takes min(coverage, clip texel) for an alpha tile
  with a clip. For every such tile of every batch, the oracle's mask must be
  min(FP64 exact area of the tile's own path, the clip tile's mask texel),
  to the exact-area tolerance of tests/test_exact_area.py.
- Composite: tile.cs.glsl:71-163 blends each framebuffer tile's list of path
  tiles back to front, src-over with the path colour (f16, gpu/renderer.rs:
  724-727) times its mask. Recomputed here in FP64 over every path tile that
  covers the pixel, in path order and with no occlusion culling (an opaque
  solid tile overwrites the destination exactly, so culling must not change
  the result), from the oracle's own tile words and mask. Every channel must
  agree within 1/255 and at least 99.9% exactly.
"""
import numpy as np
import pytest

import oracle
from helpers import oracle_frame
from oracle import exact
from pathfinder_amd import synth


def _tile_paths(b):
    """Path of every tile of a batch and the tile's coordinates in tiles."""
    md = np.asarray(b["propagate_metadata"]).astype(np.int64)
    offs = md[:, 4]
    t = np.arange(b["tile_count"], dtype=np.int64)
    p = np.searchsorted(offs, t, side="right") - 1
    w = np.maximum(md[p, 2] - md[p, 0], 1)
    local = t - offs[p]
    return p, md[p, 0] + local % w, md[p, 1] + local // w


def clipped_mask_errors(scene, size, bg):
    cmds, o = oracle_frame(scene, size, bg, threads=4)
    mask = o.read_mask()
    errs, checked = [], 0
    for kind, b in cmds["batches"]:
        bid = b["batch_id"]
        alpha = o.read(bid, oracle.ALPHA_TILES)
        if not len(alpha):
            continue
        ml = o.read(bid, oracle.MICROLINES)
        a0 = o.read(bid, oracle.ALPHA_RANGE)[0]
        paths, txs, tys = _tile_paths(b)
        edges = exact.decode_microlines(ml)
        for i, (t, clip) in enumerate(alpha):
            clip = int(np.int32(clip))
            if clip < 0:
                continue
            p = int(paths[t])
            sel = ml[:, 3] == p
            own = exact.exact_tile_coverage(tuple(e[sel] for e in edges), int(txs[t]),
                                            int(tys[t]))
            clip_cov = exact.mask_tile(mask, clip) / 255.0
            got = exact.mask_tile(mask, a0 + i) / 255.0
            errs.append(np.abs(got - np.minimum(own, clip_cov)).ravel())
            checked += 1
    return (np.concatenate(errs) if errs else np.zeros(0)), checked


def test_clipped_mask_is_min_of_own_area_and_clip_texel():
    scene, size, bg = synth.tiny_scenes()["clip"]
    e, checked = clipped_mask_errors(scene, size, bg)
    assert checked > 0
    assert e.max() < 12.0 / 255.0, e.max() * 255
    assert e.mean() < 3.0 / 255.0


def test_clipped_mask_small_workload():
    scene, size, bg = synth.scene_4k(paths=300, clips=20, clipped_paths=120, width=768,
                                     height=512, size_hi=300.0, seg_hi=60)
    e, checked = clipped_mask_errors(scene, size, bg)
    assert checked > 50
    assert np.percentile(e, 99.9) < 12.0 / 255.0, np.percentile(e, 99.9) * 255
    assert e.mean() < 1.0 / 255.0


def _f16_paints(rgba8):
    """ColorU::to_f32 (x * (1/255) in f32, color/src/lib.rs:70-73), then
    IEEE binary16 round-to-nearest-even, back to FP64."""
    c = np.asarray(rgba8, np.uint8).reshape(-1, 4).astype(np.float32)
    return (c * np.float32(1.0 / 255.0)).astype(np.float16).astype(np.float64)


def fp64_composite(cmds, o, width, height, background):
    fbw, fbh = (width + 15) // 16, (height + 15) // 16
    if background is None:
        dest = np.zeros((fbh * 16, fbw * 16, 4))
    else:
        dest = np.empty((fbh * 16, fbw * 16, 4))
        dest[:] = [background.r, background.g, background.b, background.a]
    paints = _f16_paints(cmds["paints"])
    mask = o.read_mask()
    for kind, b in cmds["batches"]:
        if kind != 4:
            continue
        tiles = o.read(b["batch_id"], oracle.TILES)
        paths, txs, tys = _tile_paths(b)
        order = np.lexsort((np.arange(len(tiles)), paths))  # path order, then tile
        for t in order:
            tx, ty = int(txs[t]), int(tys[t])
            if not (0 <= tx < fbw and 0 <= ty < fbh):
                continue
            w2, w3 = int(tiles[t, 2]), int(tiles[t, 3])
            alpha = (w2 << 8 & 0xffffffff) >> 8
            alpha = alpha - (1 << 24) if alpha >= 1 << 23 else alpha
            backdrop = ((w3 >> 24) & 0xff) - (256 if (w3 >> 24) & 0x80 else 0)
            if alpha >= 0:
                cov = exact.mask_tile(mask, alpha) / 255.0
            elif backdrop != 0:
                cov = np.ones((16, 16))
            else:
                continue
            color = paints[w3 & 0xffff] if (w3 & 0xffff) < len(paints) else np.zeros(4)
            a = color[3] * cov
            d = dest[ty * 16:ty * 16 + 16, tx * 16:tx * 16 + 16]
            for c in range(3):
                d[..., c] = d[..., c] * (1.0 - a) + color[c] * a
            d[..., 3] = d[..., 3] * (1.0 - a) + a
    out = np.floor(np.clip(dest, 0.0, 1.0) * 255.0 + 0.5).astype(np.int32)
    return out[:height, :width]


@pytest.mark.parametrize("name", ["square", "triangle", "curves", "overlap", "clip", "offscreen"])
def test_composite_matches_fp64_src_over_tiny(name):
    scene, size, bg = synth.tiny_scenes()[name]
    cmds, o = oracle_frame(scene, size, bg, threads=4)
    want = fp64_composite(cmds, o, size[0], size[1], bg)
    got = o.read_framebuffer().astype(np.int32)
    diff = np.abs(got - want)
    assert diff.max() <= 1, diff.max()
    assert np.mean(diff == 0) >= 0.999


def test_composite_matches_fp64_src_over_small():
    scene, size, bg = synth.small_scene(paths=64)
    cmds, o = oracle_frame(scene, size, bg, threads=4)
    want = fp64_composite(cmds, o, size[0], size[1], bg)
    got = o.read_framebuffer().astype(np.int32)
    diff = np.abs(got - want)
    print("composite vs FP64: max %d, identical %.5f" % (diff.max(), np.mean(diff == 0)))
    assert diff.max() <= 1
    assert np.mean(diff == 0) >= 0.999


def test_recomposite_detects_a_dropped_path():
    """Dropping the first draw path from the FP64 recomposite (as a wrong
    culling rule would) breaks agreement on the overlap scene."""
    scene, size, bg = synth.tiny_scenes()["overlap"]
    cmds, o = oracle_frame(scene, size, bg, threads=4)
    kind, b = [kb for kb in cmds["batches"] if kb[0] == 4][0]
    md = np.array(b["propagate_metadata"], copy=True)
    md[0, 2] = md[0, 0]  # path 0 loses its tiles in the recomposite only
    bad = dict(b, propagate_metadata=md)
    cmds_bad = dict(cmds, batches=[(k, bad if x is b else x) for k, x in cmds["batches"]])
    want = fp64_composite(cmds_bad, o, size[0], size[1], bg)
    got = o.read_framebuffer().astype(np.int32)
    assert np.abs(got - want).max() > 1
