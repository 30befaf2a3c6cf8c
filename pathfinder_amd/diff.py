"""Inverse rendering through the D3D11-level path.

`render(dr, points, colors)` is a torch.autograd.Function: its forward renders
filled paths with the HIP pipeline (the same kernels and launch shapes as any
frame, plus the fill's coverage checkpoint) and returns the image as floats;
its backward runs the library's hand-written adjoint (pfh_backward) and
returns dL/dpoints and dL/dcolors.

The gradient is that of the continuous model the kernels evaluate: per path
and pixel the signed integral C of the winding number over the pixel, mask
clamp(|C|, 0, 1) and the src-over composite; the 8.8 fixed point, unorm8
masks and pixels and the paints' u8/f16 rounding pass gradients through
unchanged. This has no counterpart in the reference, which renders only.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .scene import (ColorF, ColorU, CommandList, DestFramebuffer, PathArrays, RectF, Renderer,
                    RendererOptions, Scene)


class Shapes:
    """The fixed structure of a set of filled paths: per-point flags
    (PointFlags), contour ends and per-path contour ends, as PathArrays holds
    them. Coordinates and colours are the leaves."""

    def __init__(self, flags, contour_ends, path_contour_ends, fill_rules=None):
        self.flags = np.ascontiguousarray(flags, np.uint8)
        self.contour_ends = np.ascontiguousarray(contour_ends, np.uint32)
        self.path_contour_ends = np.ascontiguousarray(path_contour_ends, np.uint32)
        self.fill_rules = None if fill_rules is None else np.ascontiguousarray(fill_rules, np.int32)
        self.point_count = len(self.flags)
        self.path_count = len(self.path_contour_ends)
        self.packed = self._packed_index()

    def _packed_index(self):
        """The source point of every point of the packed draw segments: each
        contour's points, then its first point again (SegmentsD3D11::add_path,
        builder.rs:804-840)."""
        out = []
        start = 0
        for end in self.contour_ends.astype(np.int64):
            if end > start:
                out.extend(range(start, end))
                out.append(start)
            start = max(start, int(end))
        return np.asarray(out, np.int64)


def colors_to_u8(colors):
    """ColorU of float RGBA in [0, 1] (round to nearest)."""
    c = np.clip(np.asarray(colors, np.float64), 0.0, 1.0)
    return np.floor(c * 255.0 + 0.5).astype(np.uint8)


class DiffRenderer:
    """A renderer with gradients enabled for one Shapes and frame size."""

    def __init__(self, shapes, width, height, background=(1.0, 1.0, 1.0, 1.0), device=None,
                 checkpoint=None):
        self.shapes = shapes
        self.width, self.height = int(width), int(height)
        self.background = tuple(float(v) for v in background)
        self.device = torch.device(device or "cuda")
        self.renderer = Renderer(RendererOptions(DestFramebuffer.full_window(width, height),
                                                 ColorF(*self.background)))
        self.renderer.set_gradients(True)
        if checkpoint is not None:
            self.renderer.set_gradient_checkpoint(checkpoint)
        self.packed = torch.as_tensor(shapes.packed, device=self.device)
        self.frame = 0
        self._scene = None  # native scene of the structure, updated in place

    def close(self):
        if self._scene:
            L.load().pfh_scene_destroy(self._scene)
            self._scene = None
        self.renderer.close()

    def scene(self, points, colors):
        """The Scene of these leaves (host arrays)."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
        if len(pts) != self.shapes.point_count:
            raise ValueError("%d points for shapes of %d" % (len(pts), self.shapes.point_count))
        col8 = colors_to_u8(colors).reshape(-1, 4)
        if len(col8) != self.shapes.path_count:
            raise ValueError("%d colours for %d paths" % (len(col8), self.shapes.path_count))
        s = Scene()
        s.set_view_box(RectF(0, 0, self.width, self.height))
        paints = np.asarray([s.push_paint(ColorU(*[int(v) for v in c])) for c in col8], np.uint32)
        s.push_draw_paths(PathArrays(pts, self.shapes.flags, self.shapes.contour_ends,
                                     self.shapes.path_contour_ends, paints,
                                     self.shapes.fill_rules))
        return s

    def commands(self, points, colors):
        """The command list of these leaves (host arrays): the structure is
        pushed into a native scene once, and later frames replace its points
        and paint colours in place (paint i is path i's colour) before the
        C++ build."""
        lib = L.load()
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 2)
        col8 = np.ascontiguousarray(colors_to_u8(colors).reshape(-1, 4))
        if len(pts) != self.shapes.point_count or len(col8) != self.shapes.path_count:
            raise ValueError("%d points, %d colours for shapes of %d points, %d paths" %
                             (len(pts), len(col8), self.shapes.point_count,
                              self.shapes.path_count))
        first = self._scene is None
        if first:
            s = Scene()
            s.set_view_box(RectF(0, 0, self.width, self.height))
            s.push_draw_paths(PathArrays(pts, self.shapes.flags, self.shapes.contour_ends,
                                         self.shapes.path_contour_ends,
                                         np.arange(self.shapes.path_count, dtype=np.uint32),
                                         self.shapes.fill_rules))
            self._scene = s._native()
        L.check(lib.pfh_scene_set_draw_points(self._scene, pts.ctypes.data_as(C.c_void_p),
                                              len(pts)))
        L.check(lib.pfh_scene_set_paint_colors(self._scene, col8.ctypes.data_as(C.c_void_p),
                                               len(col8)))
        bo = L.BuildOptions(L.Transform2F(1, 0, 0, 1, 0, 0))
        out = C.c_void_p()
        L.check(lib.pfh_scene_build(self._scene, C.byref(bo), C.byref(out)))
        cmds = CommandList(out)
        if first:
            draw, _ = cmds.segments()
            if len(draw["points"]) != len(self.shapes.packed):
                raise RuntimeError("packed points do not follow the contour layout")
        return cmds

    def render(self, points, colors):
        """Renders and returns the image (H, W, 4) float32 on the device."""
        cmds = self.commands(points.detach().to("cpu", torch.float32).numpy(),
                             colors.detach().to("cpu", torch.float32).numpy())
        r = self.renderer
        r.begin_scene()
        r.render_commands(cmds)
        r.end_scene()
        img = torch.empty((self.height, self.width, 4), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize(self.device)
        r.framebuffer_to_float(img.data_ptr())
        self.frame += 1
        return img

    def gradients(self, grad_image):
        """dL/dpoints (N, 2) and dL/dcolors (P, 4) of the last render."""
        g = grad_image.detach().to(self.device, torch.float32).contiguous()
        gp = torch.zeros((len(self.shapes.packed), 2), dtype=torch.float32, device=self.device)
        gc = torch.zeros((self.shapes.path_count, 4), dtype=torch.float32, device=self.device)
        torch.cuda.synchronize(self.device)  # g and the zeros are ready for the library's stream
        self.renderer.backward(g.data_ptr(), gp.data_ptr(), len(self.shapes.packed),
                               gc.data_ptr(), self.shapes.path_count)
        out = torch.zeros((self.shapes.point_count, 2), dtype=torch.float32, device=self.device)
        out.index_add_(0, self.packed, gp)
        return out, gc


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, colors, dr):
        img = dr.render(points, colors)
        ctx.dr = dr
        ctx.frame = dr.frame
        return img

    @staticmethod
    def backward(ctx, grad_image):
        dr = ctx.dr
        if dr.frame != ctx.frame:
            raise RuntimeError("the renderer has drawn another frame since this forward")
        gp, gc = dr.gradients(grad_image)
        return gp.to(grad_image.dtype), gc.to(grad_image.dtype), None


def render(dr, points, colors):
    """Differentiable image (H, W, 4) of the paths: points (N, 2) and colors
    (P, 4) RGBA in [0, 1] are the leaves."""
    return _Render.apply(points, colors, dr)


def fit(dr, points, colors, target, steps, lr_points=0.2, lr_colors=0.005, callback=None):
    """Inverse rendering: Adam on the points (lr in pixels) and colours
    towards `target` (H, W, 4) under the mean squared error. points and
    colors are leaf tensors on the device (requires_grad); returns the loss
    before each step."""
    opt = torch.optim.Adam([{"params": [points], "lr": lr_points},
                            {"params": [colors], "lr": lr_colors}])
    losses = []
    for _ in range(steps):
        opt.zero_grad()
        img = render(dr, points, colors)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            colors.clamp_(0.0, 1.0)
        losses.append(loss.item())
        if callback:
            callback(len(losses))
    return losses
