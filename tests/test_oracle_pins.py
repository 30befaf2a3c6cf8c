"""Pins of the CPU oracle that do not depend on the oracle's own formulas:
IEEE binary16 rounding, closed-form coverage, geometric invariants of the
diced and binned output, and brute-force restatements of propagate and sort.
"""
import numpy as np
import pytest

import oracle
from helpers import oracle_frame
from pathfinder_amd import synth


@pytest.fixture(scope="module")
def orc():
    o = oracle.Oracle(threads=4)
    yield o
    o.close()


def test_f16_matches_ieee_round_to_nearest_even(orc):
    rng = np.random.default_rng(1)
    vals = np.concatenate([np.arange(256, dtype=np.float32) * np.float32(1.0 / 255.0),
                           rng.uniform(-70000, 70000, 2000).astype(np.float32),
                           rng.uniform(-1e-4, 1e-4, 2000).astype(np.float32),
                           np.float32([0.0, -0.0, 65504.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8])])
    for v in vals:
        with np.errstate(over="ignore"):
            want = np.float32(v).astype(np.float16).view(np.uint16)
        assert orc.lib.orc_f32_to_f16(float(v)) == want, v
        back = orc.lib.orc_f16_to_f32(int(want))
        assert np.float32(back) == np.float16(np.uint16(want).view(np.float16)).astype(np.float32)


def exact_strip_coverage(fx, fy, tx, ty):
    """Signed area under segment (fx,fy)-(tx,ty) inside the pixel column
    x in [-0.5, 0.5], for the four unit rows y in [-0.5+k, 0.5+k], k = 0..3,
    counting area below the line (towards +y), sign = +1 when from.x > to.x.
    Integrated exactly in float64 by subdividing the clipped segment."""
    x0, x1 = max(min(fx, tx), -0.5), min(max(fx, tx), 0.5)
    if x1 <= x0:
        return np.zeros(4)
    sign = 1.0 if fx > tx else -1.0
    out = np.zeros(4)
    xs = np.linspace(x0, x1, 2001)
    ys = fy + (ty - fy) * (xs - fx) / (tx - fx)
    for k in range(4):
        lo, hi = -0.5 + k, 0.5 + k
        h = np.clip(hi - ys, 0.0, 1.0)  # height of the row below the line
        out[k] = sign * np.trapezoid(h, xs)
    return out


def test_coverage_against_exact_area(orc):
    """computeCoverage approximates the exact area; the bound is the stated
    per-pixel coverage error of the path (LUT quantisation, the half-texel
    sampling offset below, and the centre-line approximation of partial
    windows)."""
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(600):
        fx, tx = rng.uniform(-1.2, 1.2, 2)
        if abs(fx - tx) < 1e-3:
            continue
        fy, ty = rng.uniform(-1.0, 4.5, 2)
        got = oracle.coverage(fx, fy, tx, ty, orc)
        want = exact_strip_coverage(fx, fy, tx, ty)
        worst = max(worst, float(np.max(np.abs(got - want))))
    assert worst < 9.0 / 255.0, worst


HALF_TEXEL = 1.0 / 32.0  # fill_area.inc.glsl:26 samples texel edges: 0.5 texel = 1/32 px


def test_coverage_full_width_is_exact_area_shifted_half_texel(orc):
    """For a segment spanning the whole pixel column the LUT tabulates the exact
    area (utils/area-lut/src/main.rs:29-65), but the lookup (y + 8) * 16 lands
    on texel edges, half a texel (1/32 px) off the generator's texel centres:
    the result is the exact area of the line moved up by 1/32 px, to within the
    LUT's 1/255 quantisation."""
    rng = np.random.default_rng(11)
    for _ in range(300):
        fx = rng.uniform(0.5, 1.5) * rng.choice([-1, 1])
        tx = -np.sign(fx) * rng.uniform(0.5, 1.5)
        fy, ty = rng.uniform(-1.0, 4.5, 2)
        got = oracle.coverage(fx, fy, tx, ty, orc)
        want = exact_strip_coverage(fx, fy - HALF_TEXEL, tx, ty - HALF_TEXEL)
        assert np.max(np.abs(got - want)) < 1.5 / 255.0, (fx, fy, tx, ty, got, want)


def test_coverage_full_column_horizontal_edge(orc):
    for h in (-0.5, -0.25, 0.0, 0.375, 1.0, 2.75):
        got = oracle.coverage(0.75, h, -0.75, h, orc)
        want = np.clip(np.array([0.5, 1.5, 2.5, 3.5]) - (h - HALF_TEXEL), 0, 1)
        assert np.allclose(got, want, atol=1.0 / 255), (h, got, want)


@pytest.fixture(scope="module", params=sorted(synth.tiny_scenes()))
def tiny(request):
    scene, size, bg = synth.tiny_scenes()[request.param]
    cmds, o = oracle_frame(scene, size, bg, threads=2)
    return request.param, scene, size, cmds, o


def _unpack_microlines(ml):
    w0, w1, w2 = ml[:, 0].astype(np.uint32), ml[:, 1].astype(np.uint32), ml[:, 2].astype(np.uint32)
    fx = (w0 & 0xffff).astype(np.int16).astype(np.float64) + (w2 & 0xff) / 256.0
    fy = (w0 >> 16).astype(np.int16).astype(np.float64) + ((w2 >> 8) & 0xff) / 256.0
    tx = (w1 & 0xffff).astype(np.int16).astype(np.float64) + ((w2 >> 16) & 0xff) / 256.0
    ty = (w1 >> 16).astype(np.int16).astype(np.float64) + ((w2 >> 24) & 0xff) / 256.0
    return fx, fy, tx, ty


def test_microlines_chain_and_hit_endpoints(tiny):
    name, scene, size, cmds, o = tiny
    for kind, b in cmds["batches"]:
        src = cmds["clip"] if b["path_source"] == 1 else cmds["draw"]
        ml = o.read(b["batch_id"], oracle.MICROLINES)
        fx, fy, tx, ty = _unpack_microlines(ml)
        # walk segments in batch order: microlines of one segment chain end to start
        pos = 0
        for p in range(b["path_count"]):
            d = b["dice_metadata"][p]
            nseg = (b["dice_metadata"][p + 1][2] if p + 1 < b["path_count"]
                    else b["segment_count"]) - d[2]
            for s in range(nseg):
                first, flags = src["indices"][d[1] + s]
                step = 3 if flags & 0x40000000 else (2 if flags & 0x80000000 else 1)
                p0, p1 = src["points"][first], src["points"][first + step]
                # the segment's microlines: until the chain reaches the end point
                start = pos
                while True:
                    pos += 1
                    if abs(tx[pos - 1] - p1[0]) <= 1 / 512 and abs(ty[pos - 1] - p1[1]) <= 1 / 512:
                        break
                    assert pos < len(ml)
                    assert tx[pos - 1] == fx[pos] and ty[pos - 1] == fy[pos]
                assert abs(fx[start] - p0[0]) <= 1 / 512 and abs(fy[start] - p0[1]) <= 1 / 512
        assert pos == len(ml)


def _winding(px, py, segs):
    """Nonzero winding number of (px, py) against closed polylines (float64),
    counting crossings of a ray towards -x; edges oriented as drawn."""
    fx, fy, tx, ty = segs
    w = 0
    up = (fy <= py) & (ty > py)
    down = (ty <= py) & (fy > py)
    with np.errstate(divide="ignore", invalid="ignore"):
        xc = fx + (py - fy) * (tx - fx) / (ty - fy)
    left = xc < px
    w += np.count_nonzero(up & left)
    w -= np.count_nonzero(down & left)
    return w


def test_solid_tile_backdrop_is_winding_number(tiny):
    """A tile with no fills is uniformly covered: its propagated backdrop must be
    the winding number of its centre against the path's microlines."""
    name, scene, size, cmds, o = tiny
    for kind, b in cmds["batches"]:
        ml = o.read(b["batch_id"], oracle.MICROLINES)
        segs_all = _unpack_microlines(ml)
        path_of = ml[:, 3]
        tiles = o.read(b["batch_id"], oracle.TILES)
        for p in range(b["path_count"]):
            x0, y0, x1, y1 = b["propagate_metadata"][p][:4]
            off = b["propagate_metadata"][p][4]
            if b["propagate_metadata"][p][7] >= 0:
                continue  # clipped paths have their backdrops culled by the clip
            sel = path_of == p
            segs = tuple(s[sel] for s in segs_all)
            w = x1 - x0
            for ty in range(y0, y1):
                for tx in range(x0, x1):
                    t = tiles[off + (ty - y0) * w + (tx - x0)]
                    if np.int32(t[1]) >= 0:
                        continue
                    bd = np.int32(t[3]) >> 24
                    # the winding number of the path (even-odd is propagated as winding)
                    assert bd == _winding(tx * 16 + 8.0, ty * 16 + 8.0, segs), (name, p, tx, ty)


def test_zbuffer_lists_and_alpha_order_brute_force(tiny):
    name, scene, size, cmds, o = tiny
    fbw = (size[0] + 15) // 16
    fbh = (size[1] + 15) // 16
    for kind, b in cmds["batches"]:
        bid = b["batch_id"]
        tiles = o.read(bid, oracle.TILES)
        z = o.read(bid, oracle.ZBUFFER)[8:]
        ftm = o.read(bid, oracle.FIRST_TILE_MAP)
        alpha = o.read(bid, oracle.ALPHA_TILES)
        a0, a1 = o.read(bid, oracle.ALPHA_RANGE)
        # alpha tiles are allocated in (column, row) order and numbered from a0
        cols = []
        for p in range(b["path_count"]):
            x0, y0, x1, y1, off = b["propagate_metadata"][p][:5]
            for tx in range(x1 - x0):
                for ty in range(y1 - y0):
                    cols.append(off + ty * (x1 - x0) + tx)
        order = {t: i for i, t in enumerate(cols)}
        ranks = [order[int(t)] for t in alpha[:, 0]]
        assert ranks == sorted(ranks)
        assert a1 - a0 == len(alpha)
        for i, (t, clip) in enumerate(alpha):
            assert (np.int32(tiles[t][2] << 8) >> 8) == a0 + i
        # z-buffer and sorted lists
        want_z = np.zeros(fbw * fbh, np.int64)
        linked = [[] for _ in range(fbw * fbh)]
        for p in range(b["path_count"]):
            x0, y0, x1, y1, off, _, zw = b["propagate_metadata"][p][:7]
            for ty in range(y0, y1):
                for tx in range(x0, x1):
                    if not (0 <= tx < fbw and 0 <= ty < fbh):
                        continue
                    t = off + (ty - y0) * (x1 - x0) + (tx - x0)
                    w = tiles[t]
                    al = np.int32(w[2] << 8) >> 8
                    bd = np.int32(w[3]) >> 24
                    if bd != 0 or al >= 0:
                        linked[ty * fbw + tx].append(t)
                    if zw and bd != 0 and al < 0:
                        want_z[ty * fbw + tx] = max(want_z[ty * fbw + tx], t)
        assert np.array_equal(z, want_z)
        for g in range(fbw * fbh):
            surv = [t for t in sorted(linked[g]) if t >= want_z[g]]
            got, t = [], ftm[g]
            while t >= 0:
                got.append(int(t))
                t = np.int32(tiles[t][0])
            assert got == surv
