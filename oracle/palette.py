"""Restatement of the reference's palette build for colour-texture paints.

TEST INFRASTRUCTURE ONLY: the checker for pathfinder_amd/csrc/pfh_palette.cpp;
the product never imports it.

Follows renderer/src/paint.rs (assign_paint_locations :453-583,
calculate_texture_transforms :585-639, create_texture_metadata :641-659,
GradientTileBuilder :813-880, PaintMetadata::filter :747-769) and
content/src/gradient.rs (sample :188-213), in binary32 with the reference's
operation order (geometry/src/transform2d.rs for the transforms), written in
numpy scalars. Paints are read by attribute: .color (r, g, b, a), .overlay
(None, a gradient with .geometry/.from_/.to/.radii/.transform/.stops/.wrap,
or a pattern with .image.pixels/.transform/.filter/.flags()) and
.composite_op.
"""
import numpy as np

from . import METADATA_DTYPE

F = np.float32
TILE = 256                     # GRADIENT_TILE_LENGTH (paint.rs:35)
LINEAR, RADIAL = 0, 1
WRAP_REPEAT = 1
REPEAT_X, REPEAT_Y, NO_SMOOTHING = 1, 2, 4
FILTER_RADIAL = 1
COMBINE_SRC_IN = 1
IDENTITY = (F(1), F(0), F(0), F(1), F(0), F(0))


def xmul(a, b):
    """Transform2F * Transform2F on (m00, m01, m10, m11, tx, ty)."""
    a00, a01, a10, a11, atx, aty = [F(v) for v in a]
    b00, b01, b10, b11, btx, bty = [F(v) for v in b]
    return (a00 * b00 + a01 * b10, a00 * b01 + a01 * b11,
            a10 * b00 + a11 * b10, a10 * b01 + a11 * b11,
            (a00 * btx + a01 * bty) + atx, (a10 * btx + a11 * bty) + aty)


def xinverse(a):
    """Transform2F::inverse: adjugate times 1 / det, then -(M^-1 v)."""
    m00, m01, m10, m11, tx, ty = [F(v) for v in a]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        r = F(1) / (m00 * m11 - m01 * m10)
        i00, i01, i10, i11 = r * m11, r * -m01, r * -m10, r * m00
        return (i00, i01, i10, i11, -(i00 * tx + i01 * ty), -(i10 * tx + i11 * ty))


def gradient_sample(stops, t):
    """Gradient::sample for stops [(offset, (r, g, b, a))] sorted by offset."""
    if not stops:
        return (0, 0, 0, 0)
    t = np.fmin(F(1), np.fmax(F(0), F(t)))
    n = 0
    while n < len(stops) and (F(stops[n][0]) < t or F(stops[n][0]) == 0):
        n += 1
    upper = min(n, len(stops) - 1)
    lower = upper - 1 if upper > 0 else upper
    (ao, ac), (bo, bc) = stops[lower], stops[upper]
    denom = F(bo) - F(ao)
    if denom == 0:
        return tuple(int(v) for v in ac)
    ratio = np.fmin((t - F(ao)) / denom, F(1))
    inv = F(1) / F(255)
    out = []
    for c in range(4):
        fa, fb = F(ac[c]) * inv, F(bc[c]) * inv
        out.append(int(np.rint((fa + (fb - fa) * ratio) * F(255))) & 0xFF)
    return tuple(out)


def _is_gradient(o):
    return hasattr(o, "stops")


def is_opaque(paint):
    """Paint::is_opaque (paint.rs:163-176)."""
    if tuple(paint.color)[3] != 255:
        return False
    o = paint.overlay
    if o is None:
        return True
    if _is_gradient(o):
        return all(tuple(s.color)[3] == 255 for s in o.stops)
    return bool(np.all(o.image.pixels[..., 3] == 255))


def build(paints, build_transform):
    """-> dict(metadata, opaque, textures, pages, texels) for a fresh texture
    manager; textures[i] is (page, sampling flags, composite op) or None."""
    render = xinverse(build_transform)
    metadata = np.zeros(len(paints), METADATA_DTYPE)
    opaque, textures, pages, images = [], [], {}, []
    tiles = []                 # [page, texels (256, 256, 4), next row]
    image_pages = {}

    def new_page(w, h):
        pid = len(pages)
        pages[pid] = (w, h)
        return pid

    for i, p in enumerate(paints):
        e = metadata[i]
        e["transform"] = IDENTITY
        e["base_color"] = tuple(p.color)
        tex = None
        o = p.overlay
        if o is not None and _is_gradient(o):
            if not tiles or tiles[-1][2] == TILE:
                tex_rgba = np.zeros((TILE, TILE, 4), np.uint8)
                tex_rgba[..., 3] = 255          # ColorU::black()
                tiles.append([new_page(TILE, TILE), tex_rgba, 0])
            tile = tiles[-1]
            row = tile[2]
            tile[2] += 1
            stops = [(F(s.offset), tuple(s.color)) for s in o.stops]
            for x in range(TILE):
                tile[1][row, x] = gradient_sample(stops, (F(x) + F(0.5)) / F(TILE))
            scale = F(1) / F(TILE)
            fx, fy = F(o.from_[0]), F(o.from_[1])
            tx_, ty_ = F(o.to[0]), F(o.to[1])
            if o.geometry == LINEAR:
                v0 = (F(row) + F(1) * F(0.5)) * scale
                dx, dy = tx_ - fx, ty_ - fy
                with np.errstate(divide="ignore", invalid="ignore"):
                    l2 = dx * dx + dy * dy
                    m0x, m0y = dx / l2, dy / l2
                    xf = (m0x, m0y, F(0), F(0), m0x * -fx + m0y * -fy, v0)
            else:
                xf = xinverse(o.transform)
                e["filter_kind"] = FILTER_RADIAL
                fp = np.zeros(20, np.float32)
                fp[:8] = [fx, fy, tx_, ty_, F(o.radii[0]), F(o.radii[1]),
                          F(0) * scale + F(0), F(row) * scale + scale * F(0.5)]
                e["filter_params"] = fp
            e["transform"] = xmul(xf, render)
            e["combine"] = COMBINE_SRC_IN
            tex = (tile[0], 1 if o.wrap == WRAP_REPEAT else 0, p.composite_op)
        elif o is not None:
            flags = o.flags()
            bx = 0 if flags & REPEAT_X else 1
            by = 0 if flags & REPEAT_Y else 1
            px = o.image.pixels
            h, w = px.shape[:2]
            key = (w, h, bx, by, px.tobytes())
            if key not in image_pages:
                image_pages[key] = new_page(w + 2 * bx, h + 2 * by)
            page = image_pages[key]
            images.append((page, bx, by, px))
            pw, ph = pages[page]
            sx, sy = F(1) / F(pw), F(1) / F(ph)
            base = xmul((F(1), F(0), F(0), F(1), F(0) * sx, F(0) * sy),
                        (sx, F(0), F(0), sy, F(0), F(0)))
            e["transform"] = xmul(xmul(base, xinverse(o.transform)), render)
            e["combine"] = COMBINE_SRC_IN
            kind, params = o.filter
            e["filter_kind"] = kind
            fp = np.zeros(20, np.float32)
            fp[:len(params)] = params
            e["filter_params"] = fp
            sampling = ((1 if flags & REPEAT_X else 0) | (2 if flags & REPEAT_Y else 0) |
                        (0 if not flags & NO_SMOOTHING else 4 | 8))
            tex = (page, sampling, p.composite_op)
        opaque.append(is_opaque(p))
        textures.append(tex)
    texels = [(t[0], 0, 0, t[1]) for t in tiles] + images
    return {"metadata": metadata, "opaque": opaque, "textures": textures, "pages": pages,
            "texels": texels}
