"""Colour-texture paints on the HIP path, through the C ABI: every paint case
(texture sampling, radial gradient, blur, colour matrix, SrcIn/DestIn, the
fifteen composite ops) and a scene of antialiased curves and clips that
mixes them must give the oracle's framebuffer byte for byte."""
import numpy as np
import pytest

import oracle
import paint_cases as pc
from helpers import compare_frame, assert_bit_exact, product_seam
from pathfinder_amd import ColorF
from pathfinder_amd import _lib as L

pytestmark = pytest.mark.gpu

CASES = pc.cases()


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_paint_case_bit_exact(name, renderer_factory):
    case = CASES[name]
    w, h = case["size"]
    _say("case paint", name, w, h)
    want = pc.oracle_render(case["cmds"], w, h)
    r = renderer_factory(w, h, ColorF(0, 0, 0, 0))
    product_seam(r, case["cmds"])
    got = r.read_pixels().astype(np.int32)
    bad = np.argwhere(np.any(got != want, axis=-1))
    assert bad.size == 0, (name, len(bad), bad[:4].tolist())


def test_mixed_paints_bit_exact(renderer_factory):
    cmds, size, bg = pc.mixed_case()
    _say("case paint mixed", size)
    o = oracle.Oracle(threads=8)
    o.render(cmds, size[0], size[1], bg)
    r = renderer_factory(size[0], size[1], bg)
    product_seam(r, cmds)
    report = compare_frame(r, o, cmds)
    _say(report)
    assert_bit_exact(report)


def test_seam_errors(renderer_factory):
    """Unallocated pages, texel rects outside the page and unknown filters are
    refused with an error, not rendered."""
    r = renderer_factory(32, 32, ColorF(0, 0, 0, 0))
    r.begin_scene()
    with pytest.raises(L.PfhError):
        r.upload_texel_data(5, 0, 0, np.zeros((2, 2, 4), np.uint8))
    r.allocate_texture_page(5, 8, 8)
    with pytest.raises(L.PfhError):
        r.upload_texel_data(5, 7, 0, np.zeros((2, 2, 4), np.uint8))
    bad = np.asarray([pc.entry(kind=9)], pc.MD)
    with pytest.raises(L.PfhError):
        r.upload_texture_metadata_entries(bad)


def test_backward_refuses_textured_paints(renderer_factory):
    import torch
    case = CASES["linear_clamp"]
    w, h = case["size"]
    r = renderer_factory(w, h, ColorF(0, 0, 0, 0))
    r.set_gradients(True)
    product_seam(r, case["cmds"])
    g_img = torch.zeros(h * w * 4, device="cuda")
    g_pts = torch.zeros(2 * len(case["cmds"]["draw"]["points"]) + 2, device="cuda")
    g_col = torch.zeros(4 * len(case["cmds"]["metadata"]), device="cuda")
    with pytest.raises(L.PfhError, match="solid paints"):
        r.backward(g_img.data_ptr(), g_pts.data_ptr(), len(case["cmds"]["draw"]["points"]),
                   g_col.data_ptr(), 1)


@pytest.mark.parametrize("xf", [None, (0.8, 0.3, -0.25, 0.9, 100.5, -40.25)])
def test_paint_scene_bit_exact(xf, renderer_factory):
    """A scene of gradient and pattern paints through the C++ scene builder
    and the HIP path, against the packing restatement and the oracle: every
    batch's buffers and the framebuffer."""
    from helpers import oracle_frame, product_frame
    from pathfinder_amd import synth
    scene, size, bg = synth.paint_scene()
    _say("case paint scene", size, xf)
    cmds, o = oracle_frame(scene, size, bg, transform=xf or (1, 0, 0, 1, 0, 0))
    r = renderer_factory(size[0], size[1], bg)
    product_frame(r, scene, transform=xf)
    report = compare_frame(r, o, cmds)
    _say({k: v for k, v in report.items() if v[0]})
    assert_bit_exact(report)


def test_gradient_pages_bit_exact(renderer_factory):
    from helpers import oracle_frame, product_frame
    from pathfinder_amd import synth
    scene, size, bg = synth.gradient_pages_scene()
    _say("case gradient pages", size)
    cmds, o = oracle_frame(scene, size, bg)
    r = renderer_factory(size[0], size[1], bg)
    product_frame(r, scene)
    report = compare_frame(r, o, cmds)
    assert_bit_exact(report)
