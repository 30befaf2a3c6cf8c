"""Paint cases: frames whose paints use the colour-texture half of
calculateColor (tile_fragment.inc.glsl:567-614) - texture sampling, the
radial gradient, blur and colour matrix filters, SrcIn/DestIn combining and
the composite ops - each with a float64 expectation that follows from the
mathematics of the feature, not from the oracle's code:

  sampling       GL_LINEAR / GL_NEAREST as the OpenGL specification defines
                 them (8.14), REPEAT or CLAMP_TO_EDGE per axis; a compute
                 shader samples at LOD 0, so the magnification filter decides
  radial         the two-point conical gradient: t solves
                 |p - (c0 + t dc)| = r0 + t dr; concentric cases in closed form
  blur           the normalised discrete Gaussian over [-support, support]
                 with clamp-to-edge (the paired bilinear taps of the shader
                 are that convolution exactly)
  colour matrix  M src + offset
  composite      the W3C Compositing and Blending formulas (separable modes;
                 colour dodge/burn without W3C's min(1, .), as the reference
                 writes them) and the HSL modes through colorsys, with the
                 reference's saturation C / V

Metadata values reach the shader as binary16 (RGBA16F texels), so the
expectations round them the same way (numpy float16).
"""
import colorsys

import numpy as np

import oracle
from oracle import packing
from pathfinder_amd import _lib as L
from pathfinder_amd.scene import ColorF, ColorU, PathArrays, RectF, Scene

MD = oracle.METADATA_DTYPE
NEAREST = 4 | 8
REPEAT_U, REPEAT_V = 1, 2
FILTER_NONE, FILTER_RADIAL, FILTER_BLUR, FILTER_COLOR_MATRIX = 0, 1, 2, 3
COMBINE_NONE, COMBINE_SRC_IN, COMBINE_DEST_IN = 0, 1, 2
# composite op -> BlendMode index (PFH_BLEND_MODE_*)
OP_BLEND = {0x1: 14, 0x2: 15, 0x3: 17, 0x4: 12, 0x5: 13, 0x6: 18, 0x7: 19, 0x8: 16,
            0x9: 20, 0xa: 21, 0xb: 22, 0xc: 23, 0xd: 24, 0xe: 25, 0xf: 26}
OP_NAMES = {0x1: "multiply", 0x2: "screen", 0x3: "overlay", 0x4: "darken", 0x5: "lighten",
            0x6: "color_dodge", 0x7: "color_burn", 0x8: "hard_light", 0x9: "soft_light",
            0xa: "difference", 0xb: "exclusion", 0xc: "hue", 0xd: "saturation",
            0xe: "color", 0xf: "luminosity"}


def h16(x):
    """float32 -> binary16 -> float64, the metadata texel encoding."""
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float64)


def to_u8(x):
    return np.floor(np.clip(x, 0.0, 1.0) * 255.0 + 0.5).astype(np.int32)


def entry(transform=(1, 0, 0, 1, 0, 0), combine=0, base=(255, 255, 255, 255), kind=0,
          params=(), blend=0):
    e = np.zeros((), MD)
    e["transform"] = transform
    e["combine"] = combine
    e["base_color"] = base
    e["filter_kind"] = kind
    p = np.zeros(20, np.float32)
    p[:len(params)] = params
    e["filter_params"] = p
    e["blend_mode"] = blend
    return e


def strip_scene(w, h, colors):
    """Vertical strips on 16-pixel tile boundaries, one path and paint each."""
    s = Scene()
    s.set_view_box(RectF(0, 0, w, h))
    sw = w // len(colors)
    for i, c in enumerate(colors):
        x0, x1 = i * sw, (i + 1) * sw
        pts = np.asarray([(x0, 0), (x1, 0), (x1, h), (x0, h)], np.float32)
        pid = s.push_paint(ColorU(*c))
        s.push_draw_paths(PathArrays(pts, np.zeros(4, np.uint8), np.asarray([4], np.uint32),
                                     [1], [pid], [L.FILL_RULE_WINDING], [-1]))
    return s


def frame(scene, entries, page=None, flags=0, page_id=0):
    """Packs the scene and attaches metadata entries and colour texture 0."""
    cmds = packing.build(scene)
    assert len(cmds["paints"]) == len(entries)
    cmds["metadata"] = np.asarray(entries, MD)
    cmds["pages"], cmds["texels"] = {}, []
    if page is not None:
        cmds["pages"][page_id] = (page.shape[1], page.shape[0])
        cmds["texels"].append((page_id, 0, 0, page))
    for kind, b in cmds["batches"]:
        if kind == 4:
            b["color_texture_page"] = page_id if page is not None else -1
            b["color_texture_sampling_flags"] = flags
    return cmds


def oracle_render(cmds, w, h, plant=0):
    o = oracle.Oracle(threads=4)
    o.set_option(oracle.OPT_PLANT_PAINT, plant)
    try:
        o.render(cmds, w, h, ColorF(0, 0, 0, 0))
        return o.read_framebuffer().astype(np.int32)
    finally:
        o.set_option(oracle.OPT_PLANT_PAINT, 0)
        o.close()


# ------------------------------------------------------------------ float64 models

def pixel_centres(w, h):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    return x + 0.5, y + 0.5


def uv_of(e, fx, fy):
    m00, m01, m10, m11, tx, ty = h16(e["transform"])
    return m00 * fx + m01 * fy + tx, m10 * fx + m11 * fy + ty


def sample(page, u, v, flags):
    """texture() at LOD 0 (OpenGL 4.6 8.14-8.15): GL_NEAREST picks texel
    floor(u W); GL_LINEAR weights the four texels around u W - 1/2."""
    H, W = page.shape[:2]
    T = page.astype(np.float64) / 255.0

    def wrap(i, n, rep):
        i = i.astype(np.int64)
        return np.mod(i, n) if rep else np.clip(i, 0, n - 1)

    ru, rv = bool(flags & REPEAT_U), bool(flags & REPEAT_V)
    if flags & 8:
        return T[wrap(np.floor(v * H), H, rv), wrap(np.floor(u * W), W, ru)]
    sx, sy = u * W - 0.5, v * H - 0.5
    x0, y0 = np.floor(sx), np.floor(sy)
    a, b = (sx - x0)[..., None], (sy - y0)[..., None]
    i0, i1 = wrap(x0, W, ru), wrap(x0 + 1, W, ru)
    j0, j1 = wrap(y0, H, rv), wrap(y0 + 1, H, rv)
    return ((1 - a) * (1 - b) * T[j0, i0] + a * (1 - b) * T[j0, i1] +
            (1 - a) * b * T[j1, i0] + a * b * T[j1, i1])


def boundary_pixels(page, u, v, flags, eps=1e-4):
    """Pixels whose nearest-texel choice sits within eps of a texel edge (a
    float32 rounding may take either side there)."""
    if not flags & 8:
        return np.zeros(u.shape, bool)
    H, W = page.shape[:2]
    fu, fv = u * W, v * H
    return (np.abs(fu - np.round(fu)) < eps) | (np.abs(fv - np.round(fv)) < eps)


def premultiplied(rgb, a):
    return np.concatenate([rgb * a[..., None], a[..., None]], axis=-1)


def conical_t(u, v, p):
    """The gradient parameter of a two-point conical gradient at (u, v): the
    roots of |p - (c0 + t dc)|^2 = (r0 + t dr)^2, the smaller if it is >= 0,
    else the larger (NaN where there is no real root)."""
    c0x, c0y, c1x, c1y, r0, r1 = [np.float64(x) for x in p[:6]]
    dcx, dcy, dr = c1x - c0x, c1y - c0y, r1 - r0
    px, py = u - c0x, v - c0y
    A = dcx * dcx + dcy * dcy - dr * dr
    B = px * dcx + py * dcy + r0 * dr
    Cc = px * px + py * py - r0 * r0
    disc = B * B - A * Cc
    with np.errstate(invalid="ignore", divide="ignore"):
        sq = np.sqrt(disc)
        ta, tb = np.minimum((B - sq) / A, (B + sq) / A), np.maximum((B - sq) / A, (B + sq) / A)
    return np.where(ta >= 0, ta, tb), disc


def gaussian_blur(page, sigma, axis, support):
    T = page.astype(np.float64) / 255.0
    k = np.arange(-support, support + 1)
    g = np.exp(-0.5 * (k / sigma) ** 2)
    g /= g.sum()
    n = T.shape[1 - axis]
    out = np.zeros_like(T)
    for kk, gg in zip(k, g):
        idx = np.clip(np.arange(n) + kk, 0, n - 1)
        out += gg * (T[:, idx] if axis == 0 else T[idx, :])
    return out


def blend_separable(op, d, s):
    """W3C Compositing and Blending Level 1, 10.1 (d = backdrop Cb, s = Cs)."""
    def hard_light(b, c):
        return np.where(c <= 0.5, b * 2 * c, b + (2 * c - 1) - b * (2 * c - 1))

    with np.errstate(divide="ignore", invalid="ignore"):
        if op == 0x1:
            return d * s
        if op == 0x2:
            return d + s - d * s
        if op == 0x3:
            return hard_light(s, d)                     # Overlay(Cb, Cs) = HardLight(Cs, Cb)
        if op == 0x4:
            return np.minimum(d, s)
        if op == 0x5:
            return np.maximum(d, s)
        if op == 0x6:
            return np.where(d == 0, 0.0, np.where(s == 1, 1.0, d / (1 - s)))
        if op == 0x7:
            return np.where(d == 1, 1.0, np.where(s == 0, 0.0, 1 - (1 - d) / s))
        if op == 0x8:
            return hard_light(d, s)
        if op == 0x9:
            D = np.where(d <= 0.25, ((16 * d - 12) * d + 4) * d, np.sqrt(d))
            return np.where(s <= 0.5, d - (1 - 2 * s) * d * (1 - d), d + (2 * s - 1) * (D - d))
        if op == 0xa:
            return np.abs(d - s)
        if op == 0xb:
            return d + s - 2 * d * s
    raise ValueError(op)


def _hsv_sat(rgb):
    return colorsys.rgb_to_hsv(*rgb)[1]


def blend_hsl(op, d, s):
    """Hue/Saturation/Color/Luminosity: HSL hue and lightness (colorsys), the
    reference's saturation C / V (HSV's), recombined with HSL's inverse."""
    out = np.zeros_like(d)
    for idx in np.ndindex(d.shape[:-1]):
        dh, dl, _ = colorsys.rgb_to_hls(*d[idx])
        sh, sl, _ = colorsys.rgb_to_hls(*s[idx])
        ds, ss = _hsv_sat(d[idx]), _hsv_sat(s[idx])
        h, sat, l = {0xc: (sh, ds, dl), 0xd: (dh, ss, dl), 0xe: (sh, ss, dl),
                     0xf: (dh, ds, sl)}[op]
        out[idx] = colorsys.hls_to_rgb(h, l, sat)
    return out


def composite(op, dest, src_rgb, src_a):
    """composite(): a src colour over the destination texel with op; the
    output alpha is 1 (tile_fragment.inc.glsl:508-534)."""
    d, da = dest[..., :3], dest[..., 3:4]
    s = np.broadcast_to(src_rgb, d.shape)
    B = blend_hsl(op, d, s) if op >= 0xc else blend_separable(op, d, s)
    sa = src_a[..., None]
    rgb = sa * (1 - da) * s + sa * da * B + (1 - sa) * d
    return np.concatenate([rgb, np.ones(d.shape[:-1] + (1,))], axis=-1)


# ------------------------------------------------------------------ the cases

def _page(rng, w, h, opaque=False):
    p = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    if opaque:
        p[..., 3] = 255
    return p


def cases():
    """name -> dict(cmds, size, expect (h, w, 4) float64 premultiplied or None,
    valid (h, w) mask of pixels the expectation covers, tol, plants)."""
    rng = np.random.default_rng(20261202)
    out = {}

    # the colour transform, transposed: u = fy / 32, v = fx / 64 on a page of
    # 32 columns and 64 rows, so the frame is the page transposed
    w, h = 64, 32
    page = _page(rng, 32, 64, opaque=True)
    e = entry((0.0, 1 / 32, 1 / 64, 0.0, 0.0, 0.0), COMBINE_SRC_IN)
    fx, fy = pixel_centres(w, h)
    u, v = uv_of(e, fx, fy)
    c = sample(page, u, v, NEAREST)
    assert np.array_equal(c, page.transpose(1, 0, 2) / 255.0)
    out["nearest_transposed"] = dict(
        cmds=frame(strip_scene(w, h, [(255, 255, 255, 255)]), [e], page, NEAREST),
        size=(w, h), expect=premultiplied(c[..., :3], c[..., 3]),
        valid=~boundary_pixels(page, u, v, NEAREST), tol=0, plants=(1,))

    for name, flags, t in (
            ("linear_clamp", 0, (0.025, 0.004, -0.006, 0.04, -0.3, -0.2)),
            ("linear_repeat", REPEAT_U | REPEAT_V, (0.055, 0.011, -0.013, 0.09, -1.2, -0.9)),
            ("nearest_repeat_u", REPEAT_U | NEAREST, (0.047, 0.013, 0.009, 0.031, -1.1, -0.2))):
        w, h = 64, 32
        page = _page(rng, 16, 12)
        e = entry(t, COMBINE_SRC_IN)
        fx, fy = pixel_centres(w, h)
        u, v = uv_of(e, fx, fy)
        c = sample(page, u, v, flags)
        out[name] = dict(cmds=frame(strip_scene(w, h, [(255, 255, 255, 255)]), [e], page, flags),
                         size=(w, h), expect=premultiplied(c[..., :3], c[..., 3]),
                         valid=~boundary_pixels(page, u, v, flags), tol=1, plants=(2,) if
                         not flags & 8 else ())

    # radial gradients over a ramp page: texel i holds (i, 255 - i, 0, 255), so
    # GL_LINEAR at column coordinate s gives red clamp(s, 0, 255) / 255
    ramp = np.zeros((4, 256, 4), np.uint8)
    ramp[..., 0] = np.arange(256)
    ramp[..., 1] = 255 - np.arange(256)
    ramp[..., 3] = 255
    for name, p, plants in (
            ("radial_concentric", (0.5, 0.5, 0.5, 0.5, 0.0, 0.45, 0.0, 0.375), ()),
            ("radial_annulus", (0.4, 0.55, 0.4, 0.55, 0.1, 0.4, 0.02, 0.625), ()),
            ("radial_two_point", (0.25, 0.5, 0.75, 0.5, 0.05, 0.15, 0.1, 0.125), (6,))):
        w = h = 64
        e = entry((1 / 64, 0, 0, 1 / 64, 0, 0), COMBINE_SRC_IN, kind=FILTER_RADIAL, params=p)
        fx, fy = pixel_centres(w, h)
        u, v = uv_of(e, fx, fy)
        # the shader's operands: from, to - from, radii, uv origin, all binary16
        hp = h16([p[0], p[1], np.float32(p[2]) - np.float32(p[0]),
                  np.float32(p[3]) - np.float32(p[1]), p[4], p[5], p[6], p[7]])
        q = (hp[0], hp[1], hp[0] + hp[2], hp[1] + hp[3], hp[4], hp[5])
        t, disc = conical_t(u, v, q)
        s = np.clip(256.0 * (hp[6] + t) - 0.5, 0.0, 255.0) / 255.0
        c = np.stack([s, 1.0 - s, np.zeros_like(s), np.ones_like(s)], -1)
        valid = np.isfinite(t) & (np.abs(disc) > 1e-4)
        if name == "radial_two_point":
            # where both roots are >= 0 the choice matters: keep those pixels
            # away from the root crossing zero
            tb = np.where(np.isfinite(t), t, 0)
            valid &= np.abs(tb) > 1e-3
        out[name] = dict(cmds=frame(strip_scene(w, h, [(255, 255, 255, 255)]), [e], ramp, 0),
                         size=(w, h), expect=c, valid=valid, tol=1, plants=plants,
                         min_valid=0.15 if name == "radial_two_point" else 0.9)

    # blur along x and along y, sigma 2.5 (support 8)
    for name, ydir in (("blur_x", 0.0), ("blur_y", 1.0)):
        w, h = 64, 16
        page = _page(rng, 64, 16)
        sigma = 2.5
        e = entry((1 / 64, 0, 0, 1 / 16, 0, 0), COMBINE_SRC_IN, kind=FILTER_BLUR,
                  params=(ydir, sigma))
        c = gaussian_blur(page, sigma, int(ydir), int(np.ceil(1.5 * sigma) * 2))
        out[name] = dict(cmds=frame(strip_scene(w, h, [(255, 255, 255, 255)]), [e], page, 0),
                         size=(w, h), expect=premultiplied(c[..., :3], c[..., 3]),
                         valid=np.ones((h, w), bool), tol=1, plants=(4,))

    # colour matrix: params[4 j + k] = M[k, j] (column j), params[16 + k] = offset
    w = h = 32
    page = _page(rng, 16, 16)
    M = rng.uniform(-0.8, 1.2, (4, 4)).astype(np.float32)
    off = rng.uniform(-0.3, 0.4, 4).astype(np.float32)
    params = np.concatenate([M.T.ravel(), off])
    e = entry((1 / 32, 0, 0, 1 / 32, 0, 0), COMBINE_SRC_IN, kind=FILTER_COLOR_MATRIX,
              params=params)
    fx, fy = pixel_centres(w, h)
    u, v = uv_of(e, fx, fy)
    src = sample(page, u, v, NEAREST)
    c = src @ h16(M).T + h16(off)
    out["color_matrix"] = dict(
        cmds=frame(strip_scene(w, h, [(255, 255, 255, 255)]), [e], page, NEAREST),
        size=(w, h), expect=premultiplied(np.clip(c[..., :3], -1e9, 1e9), c[..., 3]),
        valid=~boundary_pixels(page, u, v, NEAREST), tol=1, plants=(7,))

    # combining: DestIn keeps the paint's colour with the texel's alpha; SrcIn
    # takes the texel's colour with the paint's alpha
    w = h = 32
    page = _page(rng, 32, 32)
    for name, combine, base, plants in (
            ("dest_in", COMBINE_DEST_IN, (40, 200, 120, 200), (8,)),
            ("src_in_alpha", COMBINE_SRC_IN, (10, 20, 30, 128), ())):
        e = entry((1 / 32, 0, 0, 1 / 32, 0, 0), combine, base=base)
        fx, fy = pixel_centres(w, h)
        u, v = uv_of(e, fx, fy)
        tx = sample(page, u, v, NEAREST)
        bc = h16(np.float32(base) * np.float32(1.0 / 255.0))
        if combine == COMBINE_DEST_IN:
            rgb, a = np.broadcast_to(bc[:3], tx[..., :3].shape), tx[..., 3] * bc[3]
        else:
            rgb, a = tx[..., :3], tx[..., 3] * bc[3]
        out[name] = dict(cmds=frame(strip_scene(w, h, [base]), [e], page, NEAREST),
                         size=(w, h), expect=premultiplied(rgb, a),
                         valid=~boundary_pixels(page, u, v, NEAREST), tol=1, plants=plants)

    # composite ops over a destination page sampled at fragCoord / framebuffer
    # size (texel centres exactly): four strips, four source colours
    w, h = 64, 32
    dest = _page(rng, 64, 32)
    dest[::7, ::3, :3] = 0
    dest[3::7, 1::3, :3] = 255
    bases = [(255, 0, 128, 255), (30, 200, 90, 160), (128, 128, 128, 255), (250, 240, 10, 60)]
    for op in range(1, 16):
        ents = [entry(base=b, blend=OP_BLEND[op]) for b in bases]
        T = dest.astype(np.float64) / 255.0
        exp = np.zeros((h, w, 4))
        for i, b in enumerate(bases):
            bc = h16(np.float32(b) * np.float32(1.0 / 255.0))
            sl = slice(16 * i, 16 * i + 16)
            exp[:, sl] = composite(op, T[:, sl], bc[:3], np.full((h, 16), bc[3]))
        plants = {0x3: (3,), 0xc: (5,)}.get(op, ())
        out["composite_" + OP_NAMES[op]] = dict(
            cmds=frame(strip_scene(w, h, bases), ents, dest, 0), size=(w, h), expect=exp,
            valid=np.ones((h, w), bool), tol=1, plants=plants)
    return out


def mixed_case():
    """Antialiased curves and clips with every kind of paint over one colour
    texture page: only for comparing the HIP path with the oracle."""
    from pathfinder_amd import synth
    rng = np.random.default_rng(7)
    scene, size, bg = synth.small_scene(paths=64, size=256)
    n = len(packing.build(scene)["paints"])
    kinds = [
        entry(),
        entry((0.011, 0.002, -0.003, 0.013, 0.1, -0.2), COMBINE_SRC_IN, (255, 255, 255, 200)),
        entry((0.004, 0, 0, 0.004, 0, 0), COMBINE_SRC_IN, kind=FILTER_RADIAL,
              params=(0.5, 0.5, 0.6, 0.45, 0.05, 0.6, 0.0, 0.5)),
        entry((0.0078125, 0, 0, 0.0078125, 0, 0), COMBINE_SRC_IN, kind=FILTER_BLUR,
              params=(0.0, 3.0)),
        entry((0.0078125, 0, 0, 0.0078125, 0, 0), COMBINE_SRC_IN, kind=FILTER_BLUR,
              params=(1.0, 1.5)),
        entry((0.009, 0.001, 0, 0.009, 0, 0), COMBINE_SRC_IN, kind=FILTER_COLOR_MATRIX,
              params=rng.uniform(-0.5, 1.0, 20)),
        entry((0.01, 0, 0, 0.01, 0.3, 0.1), COMBINE_DEST_IN, (200, 40, 90, 230)),
    ] + [entry(base=(int(a), int(b), int(c), int(d)), blend=OP_BLEND[op])
         for op, (a, b, c, d) in zip(range(1, 16), rng.integers(0, 256, (15, 4)))]
    ents = [kinds[i % len(kinds)] for i in range(n)]
    page = _page(rng, 128, 96)
    return frame(scene, ents, page, REPEAT_U), size, bg
