"""The C++ scene front-end against the packing restatement, and the C ABI
surface of libpfhip.so (loadable and complete without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import packing
from pathfinder_amd import BuildOptions, Transform2F, synth
from pathfinder_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pfhip.h")


def _compare(scene, transform=None):
    t = transform or (1, 0, 0, 1, 0, 0)
    want = packing.build(scene, t)
    cl = scene.build(BuildOptions(Transform2F(*t)))
    d, c = cl.segments()
    for got, exp in ((d, want["draw"]), (c, want["clip"])):
        assert np.array_equal(got["points"], exp["points"])
        assert np.array_equal(got["indices"], exp["indices"])
    assert np.array_equal(cl.paints(), want["paints"])
    assert np.array_equal(cl.metadata().view(np.uint8), want["metadata"].view(np.uint8))
    assert cl.pages() == want["pages"]
    got_t = cl.texel_uploads()
    assert len(got_t) == len(want["texels"])
    for g, w in zip(got_t, want["texels"]):
        assert g[:3] == tuple(w[:3]) and np.array_equal(g[3], w[3])
    got_b = cl.batches()
    assert [k for k, _ in got_b] == [k for k, _ in want["batches"]]
    for (_, gb), (_, wb) in zip(got_b, want["batches"]):
        for key, gv in gb.items():
            wv = wb[key]
            if isinstance(gv, np.ndarray):
                assert np.array_equal(gv, np.asarray(wv).astype(gv.dtype).reshape(gv.shape)), key
            else:
                assert gv == wv, key
    return want


@pytest.mark.parametrize("name", sorted(synth.tiny_scenes()))
def test_packing_tiny(name):
    scene, _, _ = synth.tiny_scenes()[name]
    _compare(scene)


@pytest.mark.parametrize("name", sorted(synth.edge_scenes()))
def test_packing_edge_scenes(name):
    scene, _, _, xf = synth.edge_scenes()[name]
    _compare(scene, xf)


def test_packing_blend_modes():
    """Every blend mode: destructive ones tile the view box, only opaque
    SrcOver/Clear paths write z (builder.rs:83, 430-436)."""
    scene, size, _ = synth.blend_scene()
    cmds = _compare(scene)
    _compare(scene, (0.9, 0.2, -0.1, 1.1, 3.5, -7.25))
    (_, b), = [kb for kb in cmds["batches"] if kb[0] == packing.DRAW]
    vb = (0, 0, (size[0] + 15) // 16, (size[1] + 15) // 16)
    modes = scene.draw_arrays().blend_modes
    paints = cmds["paints"]
    for i, pm in enumerate(b["propagate_metadata"]):
        mode, path = int(modes[i]), int(pm[5])
        assert path == i
        assert (tuple(int(v) for v in pm[:4]) == vb) == (mode in packing.DESTRUCTIVE)
        opaque = paints[scene.draw_arrays().paints[i], 3] == 255
        assert pm[6] == int(opaque and mode in packing.OCCLUDING)


def test_packing_small_and_transformed():
    scene, _, _ = synth.small_scene()
    _compare(scene)
    _compare(scene, (0.8, 0.3, -0.25, 0.9, 100.5, -40.25))


def test_packing_4k_subset():
    scene, _, _ = synth.scene_4k(paths=1500, clips=60, clipped_paths=300)
    cmds = _compare(scene)
    kinds = [k for k, _ in cmds["batches"]]
    assert kinds == [packing.PREPARE_CLIP, packing.DRAW]


def test_packing_culls_paths_outside_view_box():
    scene, _, _ = synth.tiny_scenes()["offscreen"]
    from pathfinder_amd import PathArrays, ColorU
    pid = scene.push_paint(ColorU(1, 2, 3, 255))
    pts = np.asarray([(100, 100), (120, 100), (110, 130)], np.float32)
    scene.push_draw_paths(PathArrays(pts, np.zeros(3, np.uint8), [3], [1], [pid]))
    cmds = _compare(scene)
    assert cmds["batches"][-1][1]["path_count"] == 1  # the second path is culled


def test_segments_close_contours():
    # SegmentsD3D11::add_path: a contour of k on-curve points gives k segments
    # and its first point is appended (builder.rs:804-840)
    scene, _, _ = synth.tiny_scenes()["square"]
    d, _ = scene.build().segments()
    assert len(d["indices"]) == 4
    assert np.array_equal(d["points"][4], d["points"][0])
    assert list(d["indices"][:, 1]) == [0, 0, 0, 0]


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfh_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding declares exactly those
    assert sorted(L.SIGNATURES) == names


def test_struct_layouts_match_reference_sizes():
    # byte sizes of the #[repr(C)] structs of renderer/src/gpu_data.rs
    src = """#include "pfhip.h"
    #include <stddef.h>
    size_t sizes[] = {sizeof(pfh_tile), sizeof(pfh_fill), sizeof(pfh_microline),
      sizeof(pfh_alpha_tile), sizeof(pfh_tile_path_info), sizeof(pfh_propagate_metadata),
      sizeof(pfh_dice_metadata), sizeof(pfh_backdrop_info), sizeof(pfh_segment_indices)};
    """
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        so = os.path.join(d, "s.so")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"), c,
                        "-o", so], check=True)
        lib = ctypes.CDLL(so)
        sizes = (ctypes.c_size_t * 9).in_dll(lib, "sizes")
        assert list(sizes) == [16, 12, 16, 8, 16, 48, 16, 12, 8]


def test_library_and_torch_share_one_hip_runtime():
    # In a fresh process, load the library before torch: both must bind the
    # same libamdhip64, or the second runtime to initialise finds no GPU.
    import subprocess
    import sys
    code = (
        "import re, sys\n"
        "sys.path.insert(0, %r)\n"
        "from pathfinder_amd import _lib\n"
        "_lib.load()\n"
        "import torch\n"
        "maps = open('/proc/self/maps').read()\n"
        "print(sorted(set(re.findall(r'\\S*/libamdhip64\\S*', maps))))\n" % ROOT)
    out = subprocess.run([sys.executable, "-c", code], check=True, capture_output=True,
                         text=True).stdout.strip()
    paths = eval(out)
    assert len(paths) == 1, paths


def test_no_device_is_reported_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    opts = L.ContextOptions()
    opts.level = L.RENDERER_LEVEL_D3D11
    st = L.load().pfh_context_create(ctypes.byref(opts), ctypes.byref(h))
    assert st == 5  # PFH_ERROR_NO_DEVICE


def test_null_arguments_are_rejected():
    lib = L.load()
    assert lib.pfh_begin_scene(None, None) == 1
    assert lib.pfh_end_scene(None) == 1
    assert lib.pfh_scene_build(None, None, None) == 1
    assert lib.pfh_render_commands(None, None) == 1


@pytest.mark.parametrize("xf", [None, (0.8, 0.3, -0.25, 0.9, 100.5, -40.25)])
def test_packing_paint_scene(xf):
    """Gradients and patterns: TextureMetadataEntry bytes, texture pages,
    texel uploads (gradient ramps and images) and the colour-texture batch
    breaks equal the restatement."""
    scene, _, _ = synth.paint_scene()
    cmds = _compare(scene, xf)
    draws = [b for k, b in cmds["batches"] if k == packing.DRAW]
    assert len(draws) > 10
    assert {b["color_texture_page"] for b in draws} == set(cmds["pages"])


def test_packing_gradient_pages():
    scene, _, _ = synth.gradient_pages_scene()
    cmds = _compare(scene)
    assert cmds["pages"] == {0: (256, 256), 1: (256, 256)}
    assert [b["color_texture_page"] for _, b in cmds["batches"]] == [0, 1]


def test_in_place_point_and_paint_updates():
    """pfh_scene_set_draw_points / pfh_scene_set_paint_colors: a scene built
    once and updated in place packs as a scene built from scratch with the
    new points and one paint per path."""
    import ctypes
    from pathfinder_amd import diff
    from pathfinder_amd.scene import ColorU, CommandList, Paint, PathArrays, RectF, Scene
    d = synth.diff_paths(paths=40, size=128)
    sh = diff.Shapes(d["flags"], d["contour_ends"], d["path_contour_ends"])
    lib = L.load()
    pts0 = np.ascontiguousarray(d["points"], np.float32)
    pts1 = np.ascontiguousarray(pts0 * np.float32(0.9) + np.float32(7.3))
    cols = np.ascontiguousarray(np.tile([[10, 20, 30, 255]], (sh.path_count, 1)), np.uint8)

    def fresh(points):
        s = Scene()
        s.set_view_box(RectF(0, 0, 128, 128))
        for c in cols:  # one paint per path, equal colours kept apart
            s.paints.append(ColorU(*c))
            s.paint_objects.append(Paint(ColorU(*c)))
        s.push_draw_paths(PathArrays(points, sh.flags, sh.contour_ends, sh.path_contour_ends,
                                     np.arange(sh.path_count, dtype=np.uint32)))
        return packing.build(s)

    s = Scene()
    s.set_view_box(RectF(0, 0, 128, 128))
    s.push_draw_paths(PathArrays(pts0, sh.flags, sh.contour_ends, sh.path_contour_ends,
                                       np.arange(sh.path_count, dtype=np.uint32)))
    h = s._native()
    try:
        L.check(lib.pfh_scene_set_draw_points(h, pts1.ctypes.data_as(ctypes.c_void_p), len(pts1)))
        L.check(lib.pfh_scene_set_paint_colors(h, cols.ctypes.data_as(ctypes.c_void_p),
                                               len(cols)))
        assert lib.pfh_scene_set_draw_points(h, pts1.ctypes.data_as(ctypes.c_void_p), 3) != 0
        out = ctypes.c_void_p()
        bo = L.BuildOptions(L.Transform2F(1, 0, 0, 1, 0, 0))
        L.check(lib.pfh_scene_build(h, ctypes.byref(bo), ctypes.byref(out)))
        cl = CommandList(out)
    finally:
        lib.pfh_scene_destroy(h)
    want = fresh(pts1)
    assert np.array_equal(cl.paints(), want["paints"])
    assert np.array_equal(cl.segments()[0]["points"], want["draw"]["points"])
    for (_, gb), (_, wb) in zip(cl.batches(), want["batches"]):
        for key in ("dice_metadata", "tile_path_info", "propagate_metadata", "backdrops"):
            assert np.array_equal(gb[key], np.asarray(wb[key]).astype(gb[key].dtype)), key
