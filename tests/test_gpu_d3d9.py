"""The D3D9 level on the HIP device (binary16 fill raster, clip combine,
tile composite) against the D3D9 oracle (oracle/pf_d3d9.c): whole frames
byte for byte, through the scene builder's CPU tiler and the C ABI."""
import numpy as np
import pytest

from oracle import d3d9 as od3d9
from pathfinder_amd import BuildOptions, ColorF, Transform2F, synth
from pathfinder_amd import _lib as L
from pathfinder_amd.scene import RendererLevel, RendererMode

pytestmark = pytest.mark.gpu

XF = (0.8, 0.3, -0.25, 0.9, 100.5, -40.25)


def _say(*a):
    print(*a, flush=True)


def _frame(renderer_factory, scene, size, bg, transform=None):
    r = renderer_factory(size[0], size[1], bg, mode=RendererMode(RendererLevel.D3D9))
    scene.build_and_render(r, BuildOptions(Transform2F(*transform)) if transform else None)
    return r.read_pixels()


def _check(got, want, name):
    bad = np.argwhere(np.any(got != want, axis=-1))
    assert bad.size == 0, (name, len(bad), bad[:4].tolist(),
                           got[tuple(bad[0])].tolist() if len(bad) else None,
                           want[tuple(bad[0])].tolist() if len(bad) else None)


@pytest.mark.parametrize("name", sorted(synth.tiny_scenes()))
def test_d3d9_tiny_bit_exact(name, renderer_factory):
    scene, size, bg = synth.tiny_scenes()[name]
    _say("case d3d9 tiny", name)
    want = od3d9.render(scene, size, bg)
    _check(_frame(renderer_factory, scene, size, bg), want, name)


@pytest.mark.parametrize("name", sorted(synth.edge_scenes()))
def test_d3d9_edge_bit_exact(name, renderer_factory):
    scene, size, bg, xf = synth.edge_scenes()[name]
    _say("case d3d9 edge", name)
    want = od3d9.render(scene, size, bg, transform=xf or (1, 0, 0, 1, 0, 0))
    _check(_frame(renderer_factory, scene, size, bg, xf), want, name)


@pytest.mark.parametrize("xf", [None, XF])
def test_d3d9_small_bit_exact(xf, renderer_factory):
    scene, size, bg = synth.small_scene()
    _say("case d3d9 small", xf)
    want = od3d9.render(scene, size, bg, transform=xf or (1, 0, 0, 1, 0, 0), threads=8)
    _check(_frame(renderer_factory, scene, size, bg, xf), want, "small")


def test_d3d9_4k_bit_exact(renderer_factory):
    """The 4K workload (20,000 paths, 500 clip paths) at the D3D9 level."""
    scene, size, bg = synth.scene_4k()
    _say("case d3d9 4k")
    want = od3d9.render(scene, size, bg, threads=16)
    _check(_frame(renderer_factory, scene, size, bg), want, "4k")


def test_d3d9_seam_in_pieces(renderer_factory):
    """The same fills handed over in several add_fills calls with flushes in
    between (the mask keeps what earlier flushes drew and grows) give the
    frame of one build_and_render."""
    import ctypes as C
    scene, size, bg = synth.small_scene(paths=64)
    want = _frame(renderer_factory, scene, size, bg)
    cl = scene.build(None, RendererLevel.D3D9)
    fills = np.ascontiguousarray(cl.fills_d3d9())
    b = cl.batch_d3d9()
    r = renderer_factory(size[0], size[1], bg, mode=RendererMode(RendererLevel.D3D9))
    lib = L.load()
    r.begin_scene()
    r.upload_texture_metadata(np.asarray([[c.r, c.g, c.b, c.a] for c in scene.paints], np.uint8))
    cuts = [0, len(fills) // 3, len(fills) // 3 + 1, 2 * len(fills) // 3, len(fills)]
    for a, z in zip(cuts, cuts[1:]):
        part = np.ascontiguousarray(fills[a:z])
        r._check(lib.pfh_d3d9_add_fills(r.handle, part.ctypes.data_as(C.c_void_p), len(part)))
        r._check(lib.pfh_d3d9_flush_fills(r.handle))
    tiles = np.ascontiguousarray(b["tiles"])
    clips = np.ascontiguousarray(b["clips"])
    z = np.ascontiguousarray(b["z"])
    db = L.DrawTileBatchD3D9(tiles.ctypes.data_as(C.c_void_p), len(tiles),
                             clips.ctypes.data_as(C.c_void_p), len(clips),
                             z.ctypes.data_as(C.c_void_p), L.RectI(*b["z_rect"]))
    r._check(lib.pfh_d3d9_draw_tiles(r.handle, C.byref(db)))
    r.end_scene()
    _check(r.read_pixels(), want, "pieces")
