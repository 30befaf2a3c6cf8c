"""Many scenes or glyphs of uneven sizes rendered together.

Each item (a canvas, or a glyph at its size) gets a cell in a large frame:
cells are placed on shelves (rows of cells, tallest first), and the item's
points are moved by its cell's origin. A frame is then one ordinary scene:
one pass of the pipeline's fixed launch sequence renders every item in it,
and the per-item table of cells (scene id -> frame, x, y, w, h) is how the
results are found. Items never interfere: an item's geometry stays inside
its cell (the generators guarantee it; a caller must), so a path's coverage
is zero outside its own cell, and a tile can only be solid for a path that
covers it entirely, i.e. a tile inside that path's cell.

Frames are limited by a tile budget (the sum of path tile-rect areas,
estimated from each item's paths' bounds) so the 24-bit alpha tile ids of
the D3D11 formats (gpu_data.rs AlphaTileD3D11) cannot run out, and by the
16-bit paint ids (paint.rs:81): at most 65536 colours per frame.
"""
import numpy as np

from .scene import ColorU, PathArrays, RectF, Scene


class Item:
    """One canvas or glyph: its paths in local coordinates (origin at the
    cell's top-left), a paint (ColorU) per path and its size in pixels."""

    def __init__(self, arrays, colors, width, height):
        self.arrays = arrays
        self.colors = list(colors)
        self.width, self.height = int(width), int(height)
        if len(self.colors) != len(arrays):
            raise ValueError("%d colours for %d paths" % (len(self.colors), len(arrays)))

    def tile_estimate(self):
        """Sum of the paths' tile-rect areas (bounds rounded out to tiles,
        within the cell)."""
        a = self.arrays
        total = 0
        pce = a.path_contour_ends.astype(np.int64)
        ce = a.contour_ends.astype(np.int64)
        c0 = 0
        for c1 in pce:
            p0 = int(ce[c0 - 1]) if c0 > 0 else 0
            p1 = int(ce[c1 - 1]) if c1 > 0 else p0
            if p1 > p0:
                lo = np.clip(a.points[p0:p1].min(axis=0), 0, [self.width, self.height])
                hi = np.clip(a.points[p0:p1].max(axis=0), 0, [self.width, self.height])
                tw = int(np.ceil(hi[0] / 16.0)) - int(np.floor(lo[0] / 16.0))
                th = int(np.ceil(hi[1] / 16.0)) - int(np.floor(lo[1] / 16.0))
                total += max(tw, 1) * max(th, 1)
            c0 = int(c1)
        return total


class Frame:
    """A packed frame: its scene, size and the cells of its items."""

    def __init__(self, scene, width, height, cells):
        self.scene = scene
        self.width, self.height = width, height
        self.cells = cells  # item index -> (x, y, w, h)


def build_frames(items, frame_width, frame_height, tile_budget, align=1, max_paints=65536):
    """Places items on shelves (rows of cells, tallest item first) in frames
    of frame_width x at most frame_height pixels, opening a new frame when
    the next shelf would not fit or the frame's estimated tiles would exceed
    tile_budget, and builds every frame's Scene. align rounds cell sizes up
    (16 puts every cell on whole tiles)."""
    est = [it.tile_estimate() for it in items]
    size = [(-(-it.width // align) * align, -(-it.height // align) * align) for it in items]
    for w, h in size:
        if w > frame_width or h > frame_height:
            raise ValueError("a %dx%d cell does not fit a %dx%d frame" %
                             (w, h, frame_width, frame_height))
    order = sorted(range(len(items)), key=lambda i: (-size[i][1], -size[i][0], i))
    groups = []  # per frame: [(item, x, y)], used height
    cur, x, y, shelf_h, budget, colors = [], 0, 0, 0, 0, set()
    for i in order:
        w, h = size[i]
        if x + w > frame_width:  # next shelf
            y += shelf_h
            x = shelf_h = 0
        new = set(tuple(c) for c in items[i].colors)
        if cur and (y + h > frame_height or budget + est[i] > tile_budget or
                    len(colors | new) > max_paints):
            groups.append((cur, y + shelf_h if x else y))
            cur, x, y, shelf_h, budget, colors = [], 0, 0, 0, 0, set()
        cur.append((i, x, y))
        x += w
        shelf_h = max(shelf_h, h)
        budget += est[i]
        colors |= new
    if cur:
        groups.append((cur, y + shelf_h))
    out = []
    for cur, used in groups:
        height = -(-max(used, 1) // 16) * 16
        scene = Scene()
        scene.set_view_box(RectF(0, 0, frame_width, height))
        parts, cells = [], {}
        for i, x, y in cur:
            it = items[i]
            paints = np.asarray([scene.push_paint(c if isinstance(c, ColorU) else ColorU(*c))
                                 for c in it.colors], np.uint32)
            a = it.arrays
            moved = np.asarray(a.points, np.float32) + np.asarray([x, y], np.float32)
            parts.append(PathArrays(moved, a.flags, a.contour_ends, a.path_contour_ends, paints,
                                    a.fill_rules))
            cells[i] = (x, y, it.width, it.height)
        scene.push_draw_paths(PathArrays.concat(parts))
        out.append(Frame(scene, frame_width, height, cells))
    return out


def crop(pixels, cell):
    x, y, w, h = cell
    return pixels[y:y + h, x:x + w]
