"""FP64 CPU model of the differentiable render and its analytic gradient.

TEST INFRASTRUCTURE ONLY (the checker of the HIP adjoint; nothing in the
product imports it).

The model is the continuous version of what the D3D11-level kernels compute:
- dice (dice.cs.glsl:132-217): each segment becomes n microlines, with n
  from the shader's float32 formula (replicated here in float32, so it is the
  GPU's n) and the end points the Bernstein polynomial at t = i / n, here in
  FP64 without the 8.8 rounding;
- per path and pixel, C = the integral of the winding number over the pixel:
  the sum over edges of sign * (area of the pixel column strip below the edge
  within the edge's x-range), sign +1 for edges running towards -x (the
  quantity fill_area.inc.glsl tabulates in the area LUT, oracle/exact.py);
- mask m = clamp(|C|, 0, 1) (fill.cs.glsl:77-78);
- src-over of the paths in order over a clear colour (tile.cs.glsl:71-163):
  D = D (1 - a) + (rgb a, a) with a = alpha m.

The gradient of L = sum (image - target)^2 follows the chain:
- composite: with T_k the product of (1 - a_j) over the paths above k,
  dL/da_k = T_k (g . ((rgb_k, 1) - D_below)), dL/drgb_k = g T_k a_k,
  dL/dalpha_k = dL/da_k m_k, dL/dm_k = dL/da_k alpha_k;
- mask: dL/dC = dL/dm sign(C) where 0 < |C| < 1;
- edges: moving a directed edge e = p1 - p0 sweeps area of winding +-1, so
  dC/dp0 = (-e_y, e_x) int (1 - t) dt and dC/dp1 = (-e_y, e_x) int t dt over
  the parameters where the edge lies inside the pixel;
- control points: Bernstein weights, the quadratic-to-cubic elevation
  (dice.cs.glsl:183-185) and the transform's transpose.
"""
import numpy as np

from . import exact

CUBIC = 0x40000000
QUAD = 0x80000000
CONTROL_0 = 0x01
CONTROL_1 = 0x02


class Shapes:
    """Flags and contour structure of filled paths (coordinates are leaves)."""

    def __init__(self, flags, contour_ends, path_contour_ends):
        self.flags = np.asarray(flags, np.uint8)
        self.contour_ends = np.asarray(contour_ends, np.int64)
        self.path_contour_ends = np.asarray(path_contour_ends, np.int64)

    @property
    def path_count(self):
        return len(self.path_contour_ends)

    def segments(self):
        """Per path: list of (kind, source point indices) in packing order
        (SegmentsD3D11::add_path, builder.rs:804-840): every on-curve point
        starts a segment, each contour closes back to its first point."""
        out = []
        c0 = 0
        start = 0
        for p in range(self.path_count):
            c1 = int(self.path_contour_ends[p])
            segs = []
            for c in range(c0, c1):
                end = int(self.contour_ends[c])
                n = end - start
                if n > 0:
                    idx = list(range(start, end)) + [start]
                    fl = [int(self.flags[i]) for i in range(start, end)] + [0]
                    for i in range(n):
                        if fl[i] & (CONTROL_0 | CONTROL_1):
                            continue
                        if fl[i + 1] & CONTROL_0 and i + 2 <= n and fl[i + 2] & CONTROL_1:
                            segs.append(("cubic", idx[i:i + 4]))
                        elif fl[i + 1] & CONTROL_0:
                            segs.append(("quad", idx[i:i + 3]))
                        else:
                            segs.append(("line", idx[i:i + 2]))
                start = max(start, end)
            out.append(segs)
            c0 = c1
        return out


def _segment_count(kind, q):
    """n of dice.cs.glsl:189-197 in float32 (q: transformed float32 points)."""
    f = np.float32
    if kind == "line":
        dx, dy = f(q[1][0] - q[0][0]), f(q[1][1] - q[0][1])
        count = f(np.sqrt(f(f(dx * dx) + f(dy * dy)))) / f(16.0)
    else:
        if kind == "quad":
            third = f(1.0) / f(3.0)
            k2 = (f(q[1][0] * f(2.0)), f(q[1][1] * f(2.0)))
            c0 = (f(f(q[0][0] + k2[0]) * third), f(f(q[0][1] + k2[1]) * third))
            c1 = (f(f(q[2][0] + k2[0]) * third), f(f(q[2][1] + k2[1]) * third))
            p0, p3 = q[0], q[2]
        else:
            p0, c0, c1, p3 = q
        ax = abs(f(f(c1[0] - f(f(2.0) * c0[0])) + p0[0]))
        ay = abs(f(f(c1[1] - f(f(2.0) * c0[1])) + p0[1]))
        bx = abs(f(f(p3[0] - f(f(2.0) * c1[0])) + c0[0]))
        by = abs(f(f(p3[1] - f(f(2.0) * c1[1])) + c0[1]))
        mx = f(f(6.0) * (bx if ax < bx else ax))
        my = f(f(6.0) * (by if ay < by else ay))
        length = f(np.sqrt(f(f(mx * mx) + f(my * my))))
        count = f(np.sqrt(f(length / f(2.0))))
    n = int(np.ceil(count))
    return max(n, 1)


def _apply(xform, p):
    m00, m01, m10, m11, tx, ty = xform
    return np.array([m00 * p[0] + m01 * p[1] + tx, m10 * p[0] + m11 * p[1] + ty])


def _bernstein(kind, t):
    if kind == "line":
        return np.array([1.0 - t, t])
    u = 1.0 - t
    return np.array([u * u * u, 3.0 * t * u * u, 3.0 * t * t * u, t * t * t])


def flatten(points, shapes, xform=(1.0, 0.0, 0.0, 1.0, 0.0, 0.0)):
    """Microlines of every path in FP64: per path an array (M, 4) of
    (x0, y0, x1, y1) and, per microline, (control point weights at both ends
    as a (2, K) array over the K source points it depends on, those indices)."""
    pts = np.asarray(points, np.float64)
    x32 = np.asarray(xform, np.float32)
    out = []
    for segs in shapes.segments():
        lines, deps = [], []
        for kind, idx in segs:
            q32 = [tuple(np.float32(v) for v in (
                np.float32(x32[0] * np.float32(pts[i][0])) + np.float32(x32[1] * np.float32(pts[i][1])) + x32[4],
                np.float32(x32[2] * np.float32(pts[i][0])) + np.float32(x32[3] * np.float32(pts[i][1])) + x32[5]))
                for i in idx]
            n = _segment_count(kind, q32)
            q = [_apply(xform, pts[i]) for i in idx]
            if kind == "quad":
                # elevation: c0 = (p0 + 2k) / 3, c1 = (p3 + 2k) / 3
                ctrl = [q[0], (q[0] + 2 * q[1]) / 3, (q[2] + 2 * q[1]) / 3, q[2]]
                elev = np.array([[1, 0, 0], [1 / 3, 2 / 3, 0], [0, 2 / 3, 1 / 3], [0, 0, 1]])
                bk = "cubic"
            else:
                ctrl = q
                elev = np.eye(len(q))
                bk = kind
            ctrl = np.asarray(ctrl)
            for i in range(n):
                ta, tb = i / n, (i + 1) / n
                wa, wb = _bernstein(bk, ta), _bernstein(bk, tb)
                a, b = wa @ ctrl, wb @ ctrl
                lines.append((a[0], a[1], b[0], b[1]))
                deps.append((np.stack([wa @ elev, wb @ elev]), list(idx)))
        out.append((np.asarray(lines, np.float64).reshape(-1, 4), deps))
    return out


def path_coverage(lines, width, height):
    """C[y, x]: the integral of the winding number over every pixel."""
    C = np.zeros((height, width))
    full = np.zeros((height + 1, width))  # rows wholly below an edge, summed down the column
    for x0, y0, x1, y1 in lines:
        if x0 == x1:
            continue
        sign = 1.0 if x0 > x1 else -1.0
        lo, hi = min(x0, x1), max(x0, x1)
        m = (y1 - y0) / (x1 - x0)
        for c in range(max(int(np.floor(lo)), 0), min(int(np.ceil(hi)), width)):
            a, b = max(lo, c), min(hi, c + 1.0)
            if b <= a:
                continue
            ya, yb = y0 + m * (a - x0), y0 + m * (b - x0)
            L = b - a
            r0 = int(np.floor(min(ya, yb)))
            r1 = int(np.floor(max(ya, yb)))
            rows = np.arange(max(r0, 0), min(r1, height - 1) + 1)
            if len(rows):
                bottom = rows + 1.0
                C[rows, c] += sign * exact._clamped_integral(bottom - ya, bottom - yb, L)
            below = max(r1 + 1, 0)
            if below < height:
                full[below, c] += sign * L
    return C + np.cumsum(full[:height], axis=0)


def composite(masks, colors, width, height, background):
    """src-over of the paths' masks m (per path (H, W)) over a clear colour;
    returns the image and, per path, (a, D below it)."""
    cols = np.asarray(colors, np.float64)
    D = np.empty((height, width, 4))
    D[:] = np.asarray(background, np.float64)
    layers = []
    for k, m in enumerate(masks):
        a = cols[k, 3] * m
        below = D.copy()
        D = D * (1.0 - a)[..., None]
        D[..., 0] += cols[k, 0] * a
        D[..., 1] += cols[k, 1] * a
        D[..., 2] += cols[k, 2] * a
        D[..., 3] += a
        layers.append((a, below))
    return D, layers


def render(points, colors, shapes, width, height, background, xform=(1, 0, 0, 1, 0, 0)):
    """Image (H, W, 4) of the FP64 model, plus (microlines, per path (C, m))."""
    flat = flatten(points, shapes, xform)
    cov = []
    for lines, deps in flat:
        C = path_coverage(lines, width, height)
        cov.append((C, np.clip(np.abs(C), 0.0, 1.0)))
    img, _ = composite([m for C, m in cov], colors, width, height, background)
    return img, (flat, cov)


def _edge_pixels(x0, y0, x1, y1, width, height):
    """(x, y, ta, tb) for every pixel the edge crosses with ta < tb."""
    out = []
    dx, dy = x1 - x0, y1 - y0
    for y in range(max(int(np.floor(min(y0, y1))), 0), min(int(np.floor(max(y0, y1))), height - 1) + 1):
        for x in range(max(int(np.floor(min(x0, x1))), 0), min(int(np.floor(max(x0, x1))), width - 1) + 1):
            ta, tb = 0.0, 1.0
            ok = True
            for p, q in ((-dx, x0 - x), (dx, x + 1 - x0), (-dy, y0 - y), (dy, y + 1 - y0)):
                if p == 0:
                    if q < 0:
                        ok = False
                else:
                    r = q / p
                    if p < 0:
                        ta = max(ta, r)
                    else:
                        tb = min(tb, r)
            if ok and tb > ta:
                out.append((x, y, ta, tb))
    return out


def backward(points, colors, shapes, width, height, background, grad_image, state,
             xform=(1, 0, 0, 1, 0, 0)):
    """dL/dpoints (N, 2) and dL/dcolors (P, 4) for dL/dimage = grad_image.
    state = (microlines per path, per path (C, m)) as render returns it; C and
    m may come from another forward (the GPU's checkpoint and masks)."""
    flat, cov = state
    cols = np.asarray(colors, np.float64)
    g = np.asarray(grad_image, np.float64)
    n = len(cov)
    _, layers = composite([m for C, m in cov], colors, width, height, background)
    T = [None] * n
    acc = np.ones((height, width))
    for k in range(n - 1, -1, -1):
        T[k] = acc.copy()
        acc = acc * (1.0 - layers[k][0])
    gpts = np.zeros((len(points), 2))
    gcol = np.zeros((n, 4))
    m00, m01, m10, m11 = xform[:4]
    for k in range(n):
        C, m = cov[k]
        a, below = layers[k]
        rgb1 = np.array([cols[k, 0], cols[k, 1], cols[k, 2], 1.0])
        dal = T[k] * np.sum(g * (rgb1[None, None, :] - below), axis=2)
        for c in range(3):
            gcol[k, c] = np.sum(g[..., c] * T[k] * a)
        gcol[k, 3] = np.sum(dal * m)
        aC = np.abs(C)
        dC = np.where((aC > 0) & (aC < 1), dal * cols[k, 3] * np.sign(C), 0.0)
        lines, deps = flat[k]
        for (x0, y0, x1, y1), (w, idx) in zip(lines, deps):
            w0 = w1 = 0.0
            for x, y, ta, tb in _edge_pixels(x0, y0, x1, y1, width, height):
                gc = dC[y, x]
                sq = 0.5 * (tb * tb - ta * ta)
                w0 += gc * ((tb - ta) - sq)
                w1 += gc * sq
            if w0 == 0.0 and w1 == 0.0:
                continue
            nrm = np.array([-(y1 - y0), x1 - x0])
            gq = np.outer(w[0], nrm * w0) + np.outer(w[1], nrm * w1)  # (K, 2) per dependency
            for j, i in enumerate(idx):
                gx, gy = gq[j]
                gpts[i, 0] += m00 * gx + m10 * gy
                gpts[i, 1] += m01 * gx + m11 * gy
    return gpts, gcol


def loss(image, target):
    return float(np.sum((np.asarray(image) - np.asarray(target)) ** 2))
