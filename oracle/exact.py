"""FP64 exact-area coverage of the diced outlines.

TEST INFRASTRUCTURE ONLY.

For a pixel P and a closed polygon (the batch's microlines of one path, decoded
exactly from their 8.8 fixed point), the integral of the winding number over
P equals the sum over edges of sign(e) * (area of the part of P below e and
inside e's x-range), with sign +1 for edges running towards -x: the same
quantity fill_area.inc.glsl approximates with the area LUT, here integrated
in closed form. The D3D11 level stores clamp(|integral|, 0, 1) per pixel
(fill.cs.glsl:78), so that is what exact_tile_coverage returns.
"""
import numpy as np


def _clamped_integral(g0, g1, length):
    """Integral over s in [0, length] of clamp(g0 + (g1 - g0) s / length, 0, 1)."""
    def F(u):
        return np.where(u <= 0, 0.0, np.where(u >= 1, u - 0.5, 0.5 * u * u))

    dg = g1 - g0
    flat = np.abs(dg) < 1e-15
    with np.errstate(divide="ignore", invalid="ignore"):
        slope = length * (F(g1) - F(g0)) / dg
    return np.where(flat, length * np.clip(g0, 0.0, 1.0), slope)


def decode_microlines(ml):
    w0, w1, w2 = (ml[:, k].astype(np.uint32) for k in range(3))
    fx = (w0 & 0xffff).astype(np.int16).astype(np.float64) + (w2 & 0xff) / 256.0
    fy = (w0 >> 16).astype(np.int16).astype(np.float64) + ((w2 >> 8) & 0xff) / 256.0
    tx = (w1 & 0xffff).astype(np.int16).astype(np.float64) + ((w2 >> 16) & 0xff) / 256.0
    ty = (w1 >> 16).astype(np.int16).astype(np.float64) + ((w2 >> 24) & 0xff) / 256.0
    return fx, fy, tx, ty


def exact_tile_coverage(edges, tile_x, tile_y):
    """clamp(|integral of winding|, 0, 1) for the 16x16 pixels of a tile,
    indexed [row, column]."""
    fx, fy, tx, ty = edges
    out = np.zeros((16, 16))
    x_lo, x_hi = np.minimum(fx, tx), np.maximum(fx, tx)
    sign = np.where(fx > tx, 1.0, -1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        m = (ty - fy) / (tx - fx)
    rows_bottom = tile_y * 16.0 + np.arange(1, 17, dtype=np.float64)  # lower edge of each row
    for c in range(16):
        px0 = tile_x * 16.0 + c
        xa = np.maximum(x_lo, px0)
        xb = np.minimum(x_hi, px0 + 1.0)
        sel = (xb > xa) & (tx != fx)
        if not np.any(sel):
            continue
        a, b = xa[sel], xb[sel]
        ya = fy[sel] + m[sel] * (a - fx[sel])
        yb = fy[sel] + m[sel] * (b - fx[sel])
        L = b - a
        # rows x edges: part of the pixel row below the edge
        g0 = rows_bottom[:, None] - ya[None, :]
        g1 = rows_bottom[:, None] - yb[None, :]
        area = _clamped_integral(g0, g1, L[None, :])
        out[:, c] = np.sum(sign[sel][None, :] * area, axis=1)
    return np.clip(np.abs(out), 0.0, 1.0)


def mask_tile(mask_words, alpha_id):
    """The 16x16 coverage values (0..255) of one alpha tile of the mask."""
    w = mask_words[alpha_id * 64:(alpha_id + 1) * 64].reshape(4, 16)
    out = np.zeros((16, 16), np.uint8)
    for ly in range(4):
        for k in range(4):
            out[4 * ly + k] = (w[ly] >> (8 * k)) & 0xff
    return out
