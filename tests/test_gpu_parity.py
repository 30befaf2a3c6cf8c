"""Bit-exact parity of the HIP path with the oracle, through the C ABI.

Every batch's microlines, fills (with list links), tiles, z-buffer, first-tile
map, alpha tiles and column backdrops, the frame's alpha mask and the
framebuffer must equal the oracle's word for word.
"""
import numpy as np
import pytest

from helpers import assert_bit_exact, compare_frame, oracle_frame, product_frame
from pathfinder_amd import _lib as L
from pathfinder_amd import synth

pytestmark = pytest.mark.gpu

TINY = sorted(synth.tiny_scenes())


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", TINY)
def test_tiny_scene_bit_exact(name, renderer_factory):
    scene, size, bg = synth.tiny_scenes()[name]
    _say("case tiny", name, size)
    cmds, o = oracle_frame(scene, size, bg)
    r = renderer_factory(size[0], size[1], bg)
    product_frame(r, scene)
    report = compare_frame(r, o, cmds)
    _say(report)
    assert_bit_exact(report)


def test_small_workload_bit_exact(renderer_factory):
    scene, size, bg = synth.small_scene()
    _say("case small 1024x1024 256 paths")
    cmds, o = oracle_frame(scene, size, bg)
    r = renderer_factory(size[0], size[1], bg)
    product_frame(r, scene)
    report = compare_frame(r, o, cmds)
    _say(report)
    assert_bit_exact(report)


def test_transform_bit_exact(renderer_factory):
    scene, size, bg = synth.small_scene(paths=48)
    t = (0.8, 0.3, -0.25, 0.9, 100.5, -40.25)
    _say("case small transformed", t)
    cmds, o = oracle_frame(scene, size, bg, transform=t)
    r = renderer_factory(size[0], size[1], bg)
    product_frame(r, scene, transform=t)
    report = compare_frame(r, o, cmds)
    _say(report)
    assert_bit_exact(report)


def test_overflow_replay_bit_exact(renderer_factory):
    """Tiny starting capacities force the grow-and-render-again path."""
    scene, size, bg = synth.small_scene(paths=64)
    _say("case overflow replay")
    cmds, o = oracle_frame(scene, size, bg)
    r = renderer_factory(size[0], size[1], bg, initial_microline_capacity=64,
                         initial_fill_capacity=64)
    product_frame(r, scene)
    assert r.stats()["overflow_replays"] >= 1
    report = compare_frame(r, o, cmds)
    _say(report)
    assert_bit_exact(report)


def test_rerender_is_identical(renderer_factory):
    scene, size, bg = synth.small_scene(paths=96)
    _say("case rerender")
    r = renderer_factory(size[0], size[1], bg)
    product_frame(r, scene)
    first = r.read_pixels().copy()
    for _ in range(3):
        r.rerender()
    r.synchronize()
    assert np.array_equal(first, r.read_pixels())
    # serialized stages (PFH_PROFILE_SERIAL) give the same frame and time
    # every kernel
    first_mask = r.read_mask().copy()
    r.profile(L.PROFILE_SERIAL)
    r.profile_reset()
    r.rerender()
    prof = r.profile_read()
    r.profile(L.PROFILE_OFF)
    assert np.array_equal(first, r.read_pixels())
    assert np.array_equal(first_mask, r.read_mask())
    assert prof["fill"][0] >= 1 and prof["sort_tiles"][0] >= 1 and prof["tile"][0] == 1


@pytest.mark.slow
def test_4k_workload_bit_exact(renderer_factory):
    scene, size, bg = synth.scene_4k()
    _say("case 4k 3840x2160 20k paths 500 clips")
    cmds, o = oracle_frame(scene, size, bg, threads=16)
    r = renderer_factory(size[0], size[1], bg)
    product_frame(r, scene)
    report = compare_frame(r, o, cmds)
    _say(report)
    assert_bit_exact(report)


def _luts():
    """Area LUTs whose saturated cells differ from the shipped one: none at
    all (every 0 raised to 1, every 255 lowered to 254), and wider ones, not
    shaped as a prefix and a suffix of each row (bytes below 40 cleared and
    above 215 set, and a band of rows cleared)."""
    import oracle
    base = oracle.area_lut().reshape(256, 256, 4)
    none = np.clip(base, 1, 254).astype(np.uint8)
    wide = base.copy()
    wide[wide < 40] = 0
    wide[wide > 215] = 255
    wide[100:110] = 0
    return {"no saturated cells": none, "wide saturated cells": wide}


@pytest.mark.parametrize("which", sorted(_luts()))
def test_custom_area_lut_bit_exact(which, renderer_factory):
    """The fill samples the LUT only where its cell is not all 0 or all 255,
    from a table built from the LUT in use: the frame must equal the oracle's
    with the same LUT whatever its saturated cells are."""
    import oracle
    lut = _luts()[which]
    scene, size, bg = synth.small_scene(paths=96)
    _say("case custom LUT", which)
    from oracle import packing
    cmds = packing.build(scene, (1, 0, 0, 1, 0, 0))
    o = oracle.Oracle(lut=lut.reshape(-1), threads=8)
    o.render(cmds, size[0], size[1], bg)
    r = renderer_factory(size[0], size[1], bg, area_lut=lut.reshape(-1))
    product_frame(r, scene)
    report = compare_frame(r, o, cmds)
    _say(report)
    assert_bit_exact(report)
