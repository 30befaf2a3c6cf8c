"""The FP64 model of the differentiable render (oracle/diff.py) pinned against
what the mathematics fixes: exact areas, and central finite differences of
the loss (epsilon 1e-3 px in FP64) for its analytic gradient.

A parameter whose finite differences at epsilon 1e-3 and 5e-4 disagree lies
next to a pixel-boundary event (a vertex or edge crossing a pixel corner,
where the loss is not differentiable); those are left out, and at most a few
percent of the parameters may be.
"""
import numpy as np
import pytest

from oracle import diff


def _blob_paths(rng, paths, size, k=4, radius=(6, 12)):
    pts, flags, ends, pends = [], [], [], []
    for p in range(paths):
        c = rng.uniform(size * 0.3, size * 0.7, 2)
        r = rng.uniform(*radius)
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        on = [c + r * np.array([np.cos(a), np.sin(a)]) for a in ang]
        path, fl = [on[0]], [0]
        for j in range(k):
            a, b = on[j], on[(j + 1) % k]
            path += [a + (b - a) / 3 + rng.normal(0, 2, 2), a + 2 * (b - a) / 3 + rng.normal(0, 2, 2), b]
            fl += [1, 2, 0]
        path, fl = path[:-1], fl[:-1]  # the contour closes back to its first point
        pts += path
        flags += fl
        ends.append(len(pts))
        pends.append(len(ends))
    return np.asarray(pts), diff.Shapes(flags, ends, pends)


def test_square_coverage_is_exact_area():
    pts = np.array([[10.3, 10.2], [30.6, 10.2], [30.6, 20.7], [10.3, 20.7]])
    sh = diff.Shapes(np.zeros(4, np.uint8), [4], [1])
    C = diff.path_coverage(diff.flatten(pts, sh)[0][0], 40, 32)
    assert abs(abs(C).sum() - 20.3 * 10.5) < 1e-9
    assert abs(abs(C[15, 20]) - 1.0) < 1e-12 and C[5, 5] == 0.0
    # the corner pixel (10, 10) is covered over [10.3, 11] x [10.2, 11]
    assert abs(abs(C[10, 10]) - 0.7 * 0.8) < 1e-12


def test_quadratic_elevation_matches_cubic():
    """A quadratic and the cubic of its elevated control points flatten alike."""
    q = np.array([[5.0, 5.0], [30.0, 2.0], [25.0, 28.0], [4.0, 25.0]])
    shq = diff.Shapes([0, 1, 0, 0], [4], [1])
    c = np.array([q[0], (q[0] + 2 * q[1]) / 3, (q[2] + 2 * q[1]) / 3, q[2], q[3]])
    shc = diff.Shapes([0, 1, 2, 0, 0], [5], [1])
    lq = diff.flatten(q, shq)[0][0]
    lc = diff.flatten(c, shc)[0][0]
    assert lq.shape == lc.shape and np.allclose(lq, lc, atol=1e-9)


def _fd_check(pts, cols, sh, W, H, bg, target):
    def L(p, c):
        img, _ = diff.render(p, c, sh, W, H, bg)
        return diff.loss(img, target)

    img, st = diff.render(pts, cols, sh, W, H, bg)
    gp, gc = diff.backward(pts, cols, sh, W, H, bg, 2 * (img - target), st)
    kept = bad = 0
    worst = 0.0
    for i in range(len(pts)):
        for d in range(2):
            fd = []
            for eps in (1e-3, 5e-4):
                pp, pm = pts.copy(), pts.copy()
                pp[i, d] += eps
                pm[i, d] -= eps
                fd.append((L(pp, cols) - L(pm, cols)) / (2 * eps))
            scale = max(1.0, abs(fd[1]))
            if abs(fd[0] - fd[1]) > 1e-4 * scale:
                bad += 1
                continue
            kept += 1
            worst = max(worst, abs(gp[i, d] - fd[0]) / scale)
    for k in range(len(cols)):
        for c in range(4):
            eps = 1e-4
            cp, cm = cols.copy(), cols.copy()
            cp[k, c] += eps
            cm[k, c] -= eps
            fd = (L(pts, cp) - L(pts, cm)) / (2 * eps)
            worst = max(worst, abs(gc[k, c] - fd) / max(1.0, abs(fd)))
    return kept, bad, worst


@pytest.mark.parametrize("seed", [3, 7])
def test_analytic_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    W = H = 40
    pts, sh = _blob_paths(rng, 2, W)
    cols = np.array([[0.9, 0.2, 0.1, 0.8], [0.1, 0.5, 0.9, 0.6]])
    target = np.empty((H, W, 4))
    target[:] = (0.5, 0.5, 0.5, 1.0)
    kept, bad, worst = _fd_check(pts, cols, sh, W, H, (1, 1, 1, 1), target)
    assert kept >= 0.9 * (kept + bad), (kept, bad)
    assert worst <= 1e-3, worst


def test_planted_sign_error_is_caught():
    """The same check with the edge normal flipped in the model's backward
    fails: the finite differences see the sign."""
    rng = np.random.default_rng(3)
    W = H = 40
    pts, sh = _blob_paths(rng, 1, W)
    cols = np.array([[0.9, 0.2, 0.1, 0.8]])
    target = np.zeros((H, W, 4))
    img, st = diff.render(pts, cols, sh, W, H, (1, 1, 1, 1))
    gp, _ = diff.backward(pts, cols, sh, W, H, (1, 1, 1, 1), 2 * (img - target), st)
    flat, cov = st
    neg = [(C, m) for C, m in cov]
    # flipping sign(C) is the same as flipping the edge normal
    gq, _ = diff.backward(pts, cols, sh, W, H, (1, 1, 1, 1), 2 * (img - target),
                          (flat, [(-C, m) for C, m in neg]))
    assert np.allclose(gq, -gp)
    eps = 1e-3
    i = int(np.argmax(np.abs(gp[:, 0])))
    pp, pm = pts.copy(), pts.copy()
    pp[i, 0] += eps
    pm[i, 0] -= eps
    Lp = diff.loss(diff.render(pp, cols, sh, W, H, (1, 1, 1, 1))[0], target)
    Lm = diff.loss(diff.render(pm, cols, sh, W, H, (1, 1, 1, 1))[0], target)
    fd = (Lp - Lm) / (2 * eps)
    assert abs(gp[i, 0] - fd) < 1e-3 * max(1, abs(fd))
    assert abs(gq[i, 0] - fd) > 0.5 * abs(fd)
