"""The D3D9 level's CPU tiler in the library (pfh_scene_build_d3d9) against
the D3D9 oracle (oracle/pf_d3d9.c): fills with their alpha tiles, the draw
batch's tiles, clip tiles and z-buffer, word for word."""
import numpy as np
import pytest

from oracle import d3d9 as od3d9
from pathfinder_amd import BuildOptions, Transform2F, synth
from pathfinder_amd.scene import RendererLevel

XF = (0.8, 0.3, -0.25, 0.9, 100.5, -40.25)


def _compare(scene, transform=None):
    t = transform or (1, 0, 0, 1, 0, 0)
    want = od3d9.build(scene, t)
    cl = scene.build(BuildOptions(Transform2F(*t)), RendererLevel.D3D9)
    fills = cl.fills_d3d9()
    assert np.array_equal(fills, want["fills"])
    b = cl.batch_d3d9()
    if b is None:
        assert len(want["tiles"]) == 0
        return want
    assert np.array_equal(b["tiles"], want["tiles"])
    assert np.array_equal(b["clips"], want["clips"])
    assert b["z_rect"] == want["z_rect"]
    assert np.array_equal(b["z"], want["z"])
    return want


@pytest.mark.parametrize("name", sorted(synth.tiny_scenes()))
def test_d3d9_build_tiny(name):
    scene, _, _ = synth.tiny_scenes()[name]
    _compare(scene)
    _compare(scene, XF)


@pytest.mark.parametrize("name", sorted(synth.edge_scenes()))
def test_d3d9_build_edge(name):
    scene, _, _, xf = synth.edge_scenes()[name]
    _compare(scene, xf)


def test_d3d9_build_small():
    scene, _, _ = synth.small_scene()
    w = _compare(scene)
    assert len(w["fills"]) > 10000 and len(w["tiles"]) > 1000
    _compare(scene, XF)


def test_d3d9_build_4k_subset_with_clips():
    scene, _, _ = synth.scene_4k(paths=1500, clips=60, clipped_paths=300)
    w = _compare(scene)
    assert len(w["clips"]) > 0


def test_d3d9_build_refuses_what_it_cannot_draw():
    from pathfinder_amd import _lib as L
    scene, _, _ = synth.paint_scene(paths=8)
    with pytest.raises(L.PfhError):
        scene.build(None, RendererLevel.D3D9)
