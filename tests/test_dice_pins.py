"""Pins of the oracle's dice stage against closed forms evaluated in FP64.

dice.cs.glsl:173-216 flattens every segment into n microlines: n comes from
Sederberg's error bound for the parametric chords of a cubic ("Computer-Aided
Geometric Design", section 10.6: the polyline through B(i/n) lies within
L / (8 n^2) of the curve, L bounding |B''|), or from the 16 px microline
length for lines, and microline i joins the Bezier at t = i/n and (i+1)/n,
quantised to 1/256 px after a clamp to [-32768, 32767].

These tests recompute all of that in FP64 from the float32 scene points and
hold the oracle's microlines to it:
  (a) every segment's microline count equals max(ceil(x), 1) for the FP64
      bound x (a difference of one is allowed only where x lies within the
      float32 evaluation error of x of an integer: the second differences
      cancel, so their absolute error is a few ulp of the coordinates);
  (b) every microline end lies within 1/512 px (the quantisation) plus a
      float32 evaluation allowance of the FP64 curve at t = i/n;
  (c) every microline chord stays within 0.25 px (the flattening tolerance)
      + 1/256 px of the FP64 curve between its two parameters, and lines are
      cut into pieces of at most 16 px.
Planted mistakes in the oracle (twice or half the count, a swapped mix in the
curve evaluation) must fail them.
"""
import numpy as np
import pytest

import oracle
from helpers import oracle_frame
from pathfinder_amd import synth

TOLERANCE = 0.25       # dice.cs.glsl:38 (flattening tolerance, px)
MICROLINE_LENGTH = 16  # dice.cs.glsl:39
CUBIC, QUAD = 0x40000000, 0x80000000


def _unpack(ml):
    w0, w1, w2 = (ml[:, k].astype(np.uint32) for k in range(3))
    fx = (w0 & 0xffff).astype(np.int16).astype(np.float64) + (w2 & 0xff) / 256.0
    fy = (w0 >> 16).astype(np.int16).astype(np.float64) + ((w2 >> 8) & 0xff) / 256.0
    tx = (w1 & 0xffff).astype(np.int16).astype(np.float64) + ((w2 >> 16) & 0xff) / 256.0
    ty = (w1 >> 16).astype(np.int16).astype(np.float64) + ((w2 >> 24) & 0xff) / 256.0
    return np.stack([fx, fy], 1), np.stack([tx, ty], 1)


def _segments(cmds, b):
    """FP64 control points of a batch's segments in batch order (transformed)."""
    src = cmds["clip"] if b["path_source"] == 1 else cmds["draw"]
    t = np.asarray(b["transform"], np.float32).astype(np.float64)
    pts = np.asarray(src["points"], np.float32).reshape(-1, 2).astype(np.float64)
    pts = np.stack([t[0] * pts[:, 0] + t[1] * pts[:, 1] + t[4],
                    t[2] * pts[:, 0] + t[3] * pts[:, 1] + t[5]], 1)
    dm = np.asarray(b["dice_metadata"], np.int64).reshape(-1, 4)
    nxt = np.append(dm[1:, 2], b["segment_count"])
    gsegs = np.concatenate([np.arange(d[1], d[1] + c, dtype=np.int64)
                            for d, c in zip(dm, nxt - dm[:, 2])] or [np.zeros(0, np.int64)])
    idx = np.asarray(src["indices"], np.uint32).reshape(-1, 2)[gsegs]
    first, flags = idx[:, 0].astype(np.int64), idx[:, 1]
    cubic = (flags & CUBIC) != 0
    quad = ((flags & QUAD) != 0) & ~cubic
    step = np.where(cubic, 3, np.where(quad, 2, 1))
    last = len(pts) - 1
    p0, p3 = pts[first], pts[first + step]
    k0, k1 = pts[np.minimum(first + 1, last)], pts[np.minimum(first + 2, last)]
    c0 = np.where(cubic[:, None], k0, (p0 + 2.0 * k0) / 3.0)  # quadratic elevation
    c1 = np.where(cubic[:, None], k1, (p3 + 2.0 * k0) / 3.0)
    curve = cubic | quad
    return p0, c0, c1, p3, curve


def bound_x(p0, c0, c1, p3, curve):
    """Sederberg's segment count before ceil (cubic), or length / 16 (line)."""
    dd = 6.0 * np.maximum(np.abs(c1 - 2.0 * c0 + p0), np.abs(p3 - 2.0 * c1 + c0))
    xc = np.sqrt(np.hypot(dd[:, 0], dd[:, 1]) / (8.0 * TOLERANCE))
    xl = np.hypot(*(p3 - p0).T) / MICROLINE_LENGTH
    return np.where(curve, xc, xl)


def bezier(p0, c0, c1, p3, curve, t):
    t = t[:, None]
    u = 1.0 - t
    cub = u ** 3 * p0 + 3 * u * u * t * c0 + 3 * u * t * t * c1 + t ** 3 * p3
    lin = p0 + t * (p3 - p0)
    return np.where(curve[:, None], cub, lin)


def x_error(p0, c0, c1, p3, curve, x):
    """A bound on |x32 - x| for the shader's float32 evaluation: each point
    carries <= 2 ulp (transform) + 2 ulp (quadratic elevation), a second
    difference sums four such terms (16 ulp of the largest coordinate), times
    6; x = sqrt(|dd| / 2) turns an error e of |dd| into e / (4x). A line's
    length carries a few ulp, divided by 16. sqrt and the division add 2^-22
    relative."""
    m = np.max(np.abs(np.concatenate([p0, c0, c1, p3], 1)), axis=1)
    ulp = 2.0 ** (np.floor(np.log2(np.maximum(m, 1.0))) - 23)
    dx = np.where(curve, 6.0 * 16.0 * ulp / (4.0 * np.maximum(x, 0.5)), 4.0 * ulp / 16.0)
    return dx + x * 2.0 ** -22


def eval_allowance(p0, c0, c1, p3):
    """A bound on the float32 evaluation error of a point on the curve, per
    segment: the transform and three levels of nested mixes each round at the
    magnitude of the control points (6 ulp of the largest), and t = i / n
    rounds (2^-23 relative) times |dB/dt| <= 3 x the control points' span."""
    pts = np.stack([p0, c0, c1, p3], 1)
    m = np.max(np.abs(pts), axis=(1, 2))
    ulp = 2.0 ** (np.floor(np.log2(np.maximum(m, 1.0))) - 23)
    span = np.max(np.ptp(pts, axis=1), axis=1)
    return 6.0 * ulp + 3.0 * span * 2.0 ** -23


def dice_report(cmds, o):
    """Checks (a)-(c) on every batch; returns the violation counts and worst errors."""
    rep = dict(segments=0, microlines=0, near_integer=0, count_bad=0, end_bad=0,
               chord_bad=0, worst_end=0.0, worst_chord=0.0, worst_line=0.0)
    for _, b in cmds["batches"]:
        p0, c0, c1, p3, curve = _segments(cmds, b)
        frm, to = _unpack(o.read(b["batch_id"], oracle.MICROLINES))
        M, S = len(frm), len(p0)
        rep["segments"] += S
        rep["microlines"] += M
        x = bound_x(p0, c0, c1, p3, curve)
        n = np.maximum(np.ceil(x), 1).astype(np.int64)
        near = np.abs(x - np.round(x)) <= x_error(p0, c0, c1, p3, curve, x)
        rep["near_integer"] += int(near.sum())
        end = np.clip(p3, -32768.0, 32767.0)
        allow = eval_allowance(p0, c0, c1, p3)

        def ends_ok(pos, k):
            j = pos + k - 1
            return 0 <= j < M and np.all(np.abs(to[j] - end[s]) <= 1 / 512 + allow[s])

        # (a) walk the segments: the chain of segment s must end at its end point
        # after exactly n[s] microlines (n[s] +- 1 only where x is near an integer)
        offsets = np.zeros(S + 1, np.int64)
        pos = 0
        for s in range(S):
            k = n[s]
            if not ends_ok(pos, k):
                alt = [c for c in (k - 1, k + 1) if near[s] and c >= 1 and ends_ok(pos, c)]
                if alt:
                    k = alt[0]
                    n[s] = k
                else:
                    rep["count_bad"] += 1
                    break
            offsets[s] = pos
            pos += k
        offsets[S] = pos
        if rep["count_bad"] or pos != M:
            rep["count_bad"] += int(pos != M)
            continue
        # (b) every microline end at t = (i+1)/n, every start at t = i/n
        seg = np.repeat(np.arange(S), n)
        i = np.arange(M) - offsets[seg]
        nn = n[seg].astype(np.float64)
        args = (p0[seg], c0[seg], c1[seg], p3[seg], curve[seg])
        for t, got in ((i / nn, frm), ((i + 1) / nn, to)):
            want = np.clip(bezier(*args, t), -32768.0, 32767.0)
            err = np.abs(got - want) - allow[seg][:, None]
            rep["worst_end"] = max(rep["worst_end"], float(np.max(np.abs(got - want), initial=0)))
            rep["end_bad"] += int(np.count_nonzero(np.max(err, axis=1, initial=0) > 1 / 512))
        # (c) chord within tolerance of the curve between its parameters
        inside = np.all((np.abs(frm) < 32767) & (np.abs(to) < 32767), axis=1)
        d = to - frm
        dl = np.maximum(np.hypot(d[:, 0], d[:, 1]), 1e-12)
        worst = np.zeros(M)
        for f in np.linspace(0.0, 1.0, 9)[1:-1]:
            q = bezier(*args, (i + f) / nn)
            # distance from the curve point to the chord segment
            w = np.clip(np.sum((q - frm) * d, 1) / (dl * dl), 0.0, 1.0)
            r = q - (frm + w[:, None] * d)
            worst = np.maximum(worst, np.hypot(r[:, 0], r[:, 1]))
        worst = np.where(inside, worst, 0.0)
        rep["worst_chord"] = max(rep["worst_chord"], float(worst.max(initial=0)))
        line_len = np.where(curve[seg] | ~inside, 0.0, dl)
        rep["worst_line"] = max(rep["worst_line"], float(line_len.max(initial=0)))
        rep["chord_bad"] += int(np.count_nonzero(worst > TOLERANCE + 1 / 256 + 1e-3))
        rep["chord_bad"] += int(np.count_nonzero(line_len > MICROLINE_LENGTH + 1 / 128 + 1e-3))
    return rep


def _ok(rep):
    return rep["count_bad"] == 0 and rep["end_bad"] == 0 and rep["chord_bad"] == 0


def _scenes():
    out = {"tiny_" + k: (v, None) for k, v in synth.tiny_scenes().items()}
    out["small"] = (synth.small_scene(), None)
    out["small_rotated"] = (synth.small_scene(paths=64, size=512),
                            (0.8, -0.45, 0.5, 0.9, 40.0, 120.0))
    out["4k_subset"] = (synth.scene_4k(paths=1500, clips=40, clipped_paths=150), None)
    for k, (scene, size, bg, xf) in synth.edge_scenes().items():
        out["edge_" + k] = ((scene, size, bg), xf)
    return out


SCENES = _scenes()


@pytest.mark.parametrize("name", sorted(SCENES))
def test_dice_matches_fp64_closed_forms(name):
    (scene, size, bg), xf = SCENES[name]
    cmds, o = oracle_frame(scene, size, bg, threads=4,
                           transform=xf if xf else (1, 0, 0, 1, 0, 0))
    rep = dice_report(cmds, o)
    o.close()
    assert rep["segments"] > 0 and rep["microlines"] >= rep["segments"]
    assert _ok(rep), rep
    # the pins have room to fail: curves reach their tolerance and lines their length
    if name in ("small", "4k_subset"):
        assert rep["worst_chord"] > 0.1, rep


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_planted_dice_mistakes_fail_the_pins(plant):
    (scene, size, bg), _ = SCENES["tiny_curves"]
    import oracle.packing as packing
    cmds = packing.build(scene, (1, 0, 0, 1, 0, 0))
    o = oracle.Oracle(threads=2)
    o.set_option(oracle.OPT_PLANT_DICE, plant)
    o.render(cmds, size[0], size[1], bg)
    rep = dice_report(cmds, o)
    o.close()
    assert not _ok(rep), (plant, rep)
