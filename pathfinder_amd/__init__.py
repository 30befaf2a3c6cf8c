"""pathfinder_amd: Pathfinder's D3D11-level compute rasterization path
(bound -> dice -> bin -> propagate -> sort -> fill -> tile) as hand-written
HIP kernels for MI355X (gfx950) behind the C ABI of include/pfhip.h.
"""
from .scene import (BlendMode, BuildOptions, ClipPath, ColorF, ColorU, CommandList, Contour,
                    DestFramebuffer, DrawPath, FillRule, Outline, Paint, PathArrays, RectF,
                    Renderer, RendererLevel, RendererMode, RendererOptions, Scene, Transform2F)

__all__ = [
    "BlendMode", "BuildOptions", "ClipPath", "ColorF", "ColorU", "CommandList", "Contour",
    "DestFramebuffer", "DrawPath", "FillRule", "Outline", "Paint", "PathArrays", "RectF",
    "Renderer", "RendererLevel", "RendererMode", "RendererOptions", "Scene", "Transform2F",
]
