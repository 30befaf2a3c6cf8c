"""pfh_backward_noalloc against pfh_backward, through the C ABI.

Two frames in a row on one GPU, the second smaller than the first (the
static scratch array keeps the first call's contents): on each, the
allocation-free backward must give the gradients pfh_backward gives on the
same frame. Both accumulate with float atomics whose order is not fixed, so
they agree to rounding: relative error <= 1e-5 (the gradient bar is 1e-3).
A frame whose scratch does not fit is refused with PFH_ERROR_LIMIT.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _grads(dr, img, target, noalloc):
    import torch
    g = (2.0 * (img - target) / img.numel()).contiguous()
    n_packed = len(dr.shapes.packed)
    gp = torch.zeros((n_packed, 2), dtype=torch.float32, device="cuda")
    gc = torch.zeros((dr.shapes.path_count, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    fn = dr.renderer.backward_noalloc if noalloc else dr.renderer.backward
    fn(g.data_ptr(), gp.data_ptr(), n_packed, gc.data_ptr(), dr.shapes.path_count)
    return gp.cpu().numpy().astype(np.float64), gc.cpu().numpy().astype(np.float64)


def _frame(paths, size, seed):
    import torch

    from pathfinder_amd import diff, synth
    d = synth.diff_paths(seed=seed, paths=paths, size=size)
    shapes = diff.Shapes(d["flags"], d["contour_ends"], d["path_contour_ends"])
    dr = diff.DiffRenderer(shapes, size, size, (1.0, 1.0, 1.0, 1.0))
    with torch.no_grad():
        target = dr.render(torch.tensor(d["target_points"]), torch.tensor(d["colors"]))
        img = dr.render(torch.tensor(d["points"]), torch.tensor(d["colors"]))
    return dr, img, target


@pytest.mark.parametrize("paths,size,seed", [(96, 256, 11), (24, 128, 12)])
def test_noalloc_backward_matches_backward(paths, size, seed):
    _say("case noalloc backward", paths, "paths", size, "px")
    dr, img, target = _frame(paths, size, seed)
    try:
        gp0, gc0 = _grads(dr, img, target, noalloc=False)
        gp1, gc1 = _grads(dr, img, target, noalloc=True)
        gp2, gc2 = _grads(dr, img, target, noalloc=True)  # the array as the last call left it
    finally:
        dr.close()
    assert np.linalg.norm(gp0) > 0 and np.linalg.norm(gc0) > 0
    for gp, gc in ((gp1, gc1), (gp2, gc2)):
        ep = np.linalg.norm(gp - gp0) / np.linalg.norm(gp0)
        ec = np.linalg.norm(gc - gc0) / np.linalg.norm(gc0)
        _say("points rel err %.3g, colours rel err %.3g" % (ep, ec))
        assert ep <= 1e-5 and ec <= 1e-5


def test_noalloc_backward_refuses_a_frame_beyond_its_scratch():
    from pathfinder_amd import _lib as L
    _say("case noalloc limit: 1024 paths 512 px")
    dr, img, target = _frame(1024, 512, 13)
    try:
        with pytest.raises(L.PfhError) as e:
            _grads(dr, img, target, noalloc=True)
        assert e.value.status == 6  # PFH_ERROR_LIMIT
        _say(str(e.value))
    finally:
        dr.close()
