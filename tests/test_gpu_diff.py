"""The HIP adjoint (pfh_backward through pathfinder_amd.diff) against the FP64
CPU backward of oracle/diff.py.

The CPU backward is fed the GPU forward's own state: its masks (the unorm8
texels the composite read), its coverage checkpoint C and the paints as the
kernels hold them (u8, then f16), with dL/dimage from the GPU's image. The
two then evaluate the same formulas on the same inputs (fp32 against FP64, and
microline end points computed in fp32 against FP64), so the gradients must
agree closely. The FP64 formulas themselves are pinned against finite
differences in tests/test_diff_oracle.py.

At the 1024-path 512 x 512 workload, where the CPU model is too slow, a
gradient step must lower the loss (a size-independent property).
"""
import numpy as np
import pytest

from oracle import diff as odiff
from oracle import exact, packing

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _blob_arrays(rng, paths, size, k=5):
    pts, flags, ends, pends = [], [], [], []
    for p in range(paths):
        c = rng.uniform(size * 0.25, size * 0.75, 2)
        r = rng.uniform(size * 0.1, size * 0.3)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        on = [c + r * np.array([np.cos(a), np.sin(a)]) for a in ang]
        path, fl = [on[0]], [0]
        for j in range(k):
            a, b = on[j], on[(j + 1) % k]
            path += [a + (b - a) / 3 + rng.normal(0, 3, 2), a + 2 * (b - a) / 3 + rng.normal(0, 3, 2), b]
            fl += [1, 2, 0]
        path, fl = path[:-1], fl[:-1]
        pts += path
        flags += fl
        ends.append(len(pts))
        pends.append(len(ends))
    return np.asarray(pts, np.float32), np.asarray(flags, np.uint8), ends, pends


def _f16_colors(colors):
    from pathfinder_amd.diff import colors_to_u8
    c = colors_to_u8(colors).astype(np.float32) * np.float32(1.0 / 255.0)
    return c.astype(np.float16).astype(np.float64)


def _gpu_layers(dr, scene, W, H, path_count):
    """Per path (C, m) as the GPU forward left them."""
    from pathfinder_amd import _lib as L
    r = dr.renderer
    cmds = packing.build(scene)
    (kind, b), = [kb for kb in cmds["batches"] if kb[0] == 4]
    bid = b["batch_id"]
    tiles = r.read_debug(bid, L.DEBUG_TILES).view(np.uint32).reshape(-1, 4)
    a0 = int(r.read_debug(bid, L.DEBUG_ALPHA_RANGE).view(np.uint32)[0])
    cov = r.read_debug(bid, L.DEBUG_COVERAGE).view(np.float32).reshape(-1, 256)
    mask = r.read_mask()
    md = np.asarray(b["propagate_metadata"]).astype(np.int64)
    dice = np.asarray(b["dice_metadata"]).astype(np.int64)
    out = [(np.zeros((H, W)), np.zeros((H, W))) for _ in range(path_count)]
    yy, xx = np.mgrid[0:16, 0:16]
    order = (yy & 3) * 64 + (yy >> 2) * 16 + xx
    for bp in range(len(md)):
        C, m = out[int(dice[bp, 0])]
        x0, y0, x1, y1, off = (int(v) for v in md[bp, :5])
        for ty in range(y0, y1):
            for tx in range(x0, x1):
                t = off + (ty - y0) * (x1 - x0) + (tx - x0)
                w2, w3 = int(tiles[t, 2]), int(tiles[t, 3])
                alpha = ((w2 << 8) & 0xffffffff) >> 8
                alpha = alpha - (1 << 24) if alpha >= 1 << 23 else alpha
                bd = ((w3 >> 24) & 0xff) - (256 if (w3 >> 24) & 0x80 else 0)
                if alpha >= 0:
                    cv = cov[alpha - a0][order].astype(np.float64)
                    mk = exact.mask_tile(mask, alpha) / 255.0
                else:
                    cv = np.full((16, 16), float(bd))
                    mk = np.full((16, 16), 1.0 if bd != 0 else 0.0)
                ys, xs = slice(ty * 16, min(ty * 16 + 16, H)), slice(tx * 16, min(tx * 16 + 16, W))
                h, w = ys.stop - ys.start, xs.stop - xs.start
                if h > 0 and w > 0:
                    C[ys, xs] = cv[:h, :w]
                    m[ys, xs] = mk[:h, :w]
    return out


@pytest.mark.parametrize("seed,paths,size", [(11, 3, 64), (12, 8, 96)])
def test_gpu_gradient_matches_cpu_backward(seed, paths, size):
    import torch

    from pathfinder_amd import _lib as L
    from pathfinder_amd import diff
    _say("case diff gpu vs cpu", seed, paths, size)
    rng = np.random.default_rng(seed)
    pts, flags, ends, pends = _blob_arrays(rng, paths, size)
    colors = np.concatenate([rng.uniform(0, 1, (paths, 3)),
                             rng.uniform(0.3, 1.0, (paths, 1))], axis=1).astype(np.float32)
    shapes = diff.Shapes(flags, ends, pends)
    bg = (1.0, 1.0, 1.0, 1.0)
    dr = diff.DiffRenderer(shapes, size, size, bg, checkpoint=L.CHECKPOINT_COVERAGE)
    try:
        p_t = torch.tensor(pts, device="cuda", requires_grad=True)
        c_t = torch.tensor(colors, device="cuda", requires_grad=True)
        img = diff.render(dr, p_t, c_t)
        target = torch.full_like(img, 0.5)
        loss = ((img - target) ** 2).sum()
        loss.backward()
        gp = p_t.grad.cpu().numpy().astype(np.float64)
        gc = c_t.grad.cpu().numpy().astype(np.float64)
        img_np = img.detach().cpu().numpy().astype(np.float64)
        scene = dr.scene(pts, colors)
        layers = _gpu_layers(dr, scene, size, size, paths)
    finally:
        dr.close()
    osh = odiff.Shapes(flags, ends, pends)
    flat = odiff.flatten(pts.astype(np.float64), osh)
    cols16 = _f16_colors(colors)
    # the CPU composite of the GPU masks reproduces the GPU image
    cpu_img, _ = odiff.composite([m for C, m in layers], cols16, size, size, bg)
    assert np.abs(cpu_img - img_np).max() <= 1.0 / 255.0 + 1e-6
    g = 2.0 * (img_np - 0.5)
    cp, cc = odiff.backward(pts.astype(np.float64), cols16, osh, size, size, bg, g,
                            (flat, layers))
    ep = np.linalg.norm(gp - cp) / max(np.linalg.norm(cp), 1e-12)
    ec = np.linalg.norm(gc - cc) / max(np.linalg.norm(cc), 1e-12)
    _say("points: |cpu| %.4g rel err %.3g; colours: |cpu| %.4g rel err %.3g" %
         (np.linalg.norm(cp), ep, np.linalg.norm(cc), ec))
    assert np.linalg.norm(cp) > 0 and np.linalg.norm(cc) > 0
    assert ep <= 1e-3, ep
    assert ec <= 1e-4, ec


def test_gradient_step_lowers_loss_at_512():
    import torch

    from pathfinder_amd import diff, synth
    d = synth.diff_paths()
    size = d["size"]
    _say("case diff 1024 paths", size)
    shapes = diff.Shapes(d["flags"], d["contour_ends"], d["path_contour_ends"])
    dr = diff.DiffRenderer(shapes, size[0], size[1], (1.0, 1.0, 1.0, 1.0))
    try:
        with torch.no_grad():
            target = dr.render(torch.tensor(d["target_points"]), torch.tensor(d["colors"]))
        p_t = torch.tensor(d["points"], device="cuda", requires_grad=True)
        c_t = torch.tensor(d["colors"], device="cuda", requires_grad=True)
        img = diff.render(dr, p_t, c_t)
        loss = ((img - target) ** 2).sum()
        loss.backward()
        g = p_t.grad
        assert torch.isfinite(g).all() and torch.isfinite(c_t.grad).all()
        assert g.abs().sum() > 0
        with torch.no_grad():
            step = 0.5 / g.abs().max()
            img2 = dr.render(p_t - step * g, c_t)
            loss2 = ((img2 - target) ** 2).sum()
        _say("loss %.3f -> %.3f" % (loss.item(), loss2.item()))
        assert loss2.item() < loss.item()
    finally:
        dr.close()


def test_adam_fit_with_colours_lowers_loss_at_512():
    """Twelve Adam steps on points and colours together (the benchmark's fit)."""
    import torch

    from pathfinder_amd import diff, synth
    d = synth.diff_paths()
    size = d["size"]
    _say("case diff fit 1024 paths", size)
    shapes = diff.Shapes(d["flags"], d["contour_ends"], d["path_contour_ends"])
    dr = diff.DiffRenderer(shapes, size[0], size[1], (1.0, 1.0, 1.0, 1.0))
    try:
        with torch.no_grad():
            target = dr.render(torch.tensor(d["target_points"]), torch.tensor(d["colors"]))
        p_t = torch.tensor(d["points"], device="cuda", requires_grad=True)
        c_t = torch.tensor(d["colors"], device="cuda", requires_grad=True)
        losses = diff.fit(dr, p_t, c_t, target, 12)
        _say("losses", ["%.6f" % v for v in losses])
        assert all(np.isfinite(losses))
        assert losses[-1] < 0.95 * losses[0]
        assert sum(b < a for a, b in zip(losses, losses[1:])) >= 8
    finally:
        dr.close()


def _expected_states(C, clip_texels, clipped):
    """The adjoint's 2-bit states from the coverage checkpoint C (per alpha
    tile [k][lane]) and, for the clipped tiles, their clip texel bytes."""
    aC = np.abs(C)
    open_ = (aC > 0) & (aC < 1)
    if clip_texels is not None:
        cv = clip_texels.astype(np.float32) / np.float32(255.0)
        limited = (cv < np.minimum(aC, np.float32(1.0))) & clipped[:, None, None]
        open_ &= ~limited
    neg = C < 0
    bits = open_.astype(np.uint32) | (neg.astype(np.uint32) << 1)      # [tile][k][lane]
    return (bits[:, 0] | (bits[:, 1] << 2) | (bits[:, 2] << 4) | (bits[:, 3] << 6)).astype(
        np.uint8)


@pytest.mark.parametrize("name", ["clip", "small"])
def test_coverage_states_equal_the_checkpoint(name, renderer_factory):
    """The states of PFH_CHECKPOINT_STATES (written by the forward fill) and
    of PFH_CHECKPOINT_RECOMPUTE (written by the fill the backward reruns)
    equal the ones computed from the float checkpoint C and the clip texels,
    on a scene with clip paths and on 256 random paths."""
    import ctypes as C_
    import torch
    from pathfinder_amd import _lib as L
    from pathfinder_amd import synth
    from helpers import product_frame
    scene, size, bg = (synth.tiny_scenes()["clip"] if name == "clip"
                       else synth.small_scene(paths=256, size=512))
    got = {}
    for mode in (L.CHECKPOINT_COVERAGE, L.CHECKPOINT_STATES, L.CHECKPOINT_RECOMPUTE):
        _say("case states", name, mode)
        r = renderer_factory(size[0], size[1], bg)
        r.set_gradients(True)
        r.set_gradient_checkpoint(mode)
        cl = product_frame(r, scene)
        (kind, b), = [kb for kb in cl.batches() if kb[0] == 4]
        bid = b["batch_id"]
        if mode == L.CHECKPOINT_COVERAGE:
            cov = r.read_debug(bid, L.DEBUG_COVERAGE).view(np.float32).reshape(-1, 4, 64)
            alpha = r.read_debug(bid, L.DEBUG_ALPHA_TILES).view(np.uint32).reshape(-1, 2)
            mask = r.read_mask().view(np.uint8).reshape(-1, 64, 4)
            clip = alpha[:, 1].astype(np.int64)
            texels = None
            if np.any(clip != 0xFFFFFFFF):
                texels = np.zeros((len(clip), 4, 64), np.uint8)
                for i in np.nonzero(clip != 0xFFFFFFFF)[0]:
                    texels[i] = mask[clip[i]].T
            want = _expected_states(cov, texels, clip != 0xFFFFFFFF)
            assert np.any(want & 1) and (name != "clip" or texels is not None)
            continue
        if mode == L.CHECKPOINT_RECOMPUTE:
            g = torch.zeros(size[0] * size[1] * 4, device="cuda")
            n = len(cl.segments()[0]["points"])
            gp = torch.zeros(2 * n, device="cuda")
            gc = torch.zeros(4 * scene.draw_path_count(), device="cuda")
            torch.cuda.synchronize()
            r.backward(g.data_ptr(), gp.data_ptr(), n, gc.data_ptr(), scene.draw_path_count())
        got[mode] = r.read_debug(bid, L.DEBUG_COVERAGE_STATES).view(np.uint8).reshape(-1, 64)
    for mode, st in got.items():
        bad = np.argwhere(st != want)
        for t, lane in bad[:6]:
            _say("mode", mode, "tile", t, "lane", lane, "got", st[t, lane], "want", want[t, lane],
                 "C", cov[t, :, lane].tolist(), "clip", int(clip[t]),
                 "texels", None if texels is None else texels[t, :, lane].tolist())
        assert np.array_equal(st, want), (mode, int(np.count_nonzero(st != want)))


def test_checkpoint_modes_give_the_same_gradients():
    """PFH_CHECKPOINT_COVERAGE (C per pixel), _STATES (2 bits per pixel from
    the forward fill) and _RECOMPUTE (the backward reruns the fill) give the
    same image and the same point and colour gradients (the states they feed
    the adjoint are compared bit for bit above)."""
    import torch
    from pathfinder_amd import _lib as L
    from pathfinder_amd import diff, synth
    d = synth.diff_paths(paths=96, size=128)
    shapes = diff.Shapes(d["flags"], d["contour_ends"], d["path_contour_ends"])
    out = {}
    for mode in (L.CHECKPOINT_COVERAGE, L.CHECKPOINT_STATES, L.CHECKPOINT_RECOMPUTE):
        _say("case checkpoint mode", mode)
        dr = diff.DiffRenderer(shapes, 128, 128, checkpoint=mode)
        try:
            p_t = torch.tensor(d["points"], device="cuda", requires_grad=True)
            c_t = torch.tensor(d["colors"], device="cuda", requires_grad=True)
            img = diff.render(dr, p_t, c_t)
            loss = ((img - 0.5) ** 2).sum()
            loss.backward()
            out[mode] = (img.detach().cpu().numpy(), p_t.grad.cpu().numpy(),
                         c_t.grad.cpu().numpy())
        finally:
            dr.close()
    ref = out[L.CHECKPOINT_COVERAGE]
    assert np.abs(ref[1]).max() > 0
    for mode, got in out.items():
        # the image bit for bit; the gradients up to the order of the float
        # atomics that sum them (which changes from run to run in any mode)
        assert np.array_equal(got[0], ref[0]), mode
        for a, b in zip(got[1:], ref[1:]):
            assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max(), mode
