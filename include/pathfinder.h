/*
 * pathfinder.h - the reference's C API (c/src/lib.rs) over this library.
 *
 * The backend-neutral entry points keep the reference's names, argument
 * meaning and ownership rules: canvas (c/src/lib.rs:225-440, without text),
 * paths (:443-524), fill styles (:526-536), resource loaders (:538-563),
 * render transforms and build options (:738-783), scenes and scene proxies
 * (:785-800) and SVG scenes (:802-836). The GL device, framebuffer and
 * renderer calls (:565-625) become PFHIP* calls for a HIP device; there is
 * no window, so PFHIPRendererReadPixels returns the destination framebuffer.
 *
 * Like the reference: handles are created by *Create and released by
 * *Destroy; a call documented to consume a handle takes ownership of it;
 * nothing returns an error code, and only the SVG loaders and resource
 * loaders return NULL on failure. Unlike the reference, which panics, a
 * failing renderer call records its message (PFHIPRendererGetLastError).
 */
#ifndef PATHFINDER_H
#define PATHFINDER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* c/src/lib.rs:58-94 */
#define PF_LINE_CAP_BUTT 0
#define PF_LINE_CAP_SQUARE 1
#define PF_LINE_CAP_ROUND 2
#define PF_LINE_JOIN_MITER 0
#define PF_LINE_JOIN_BEVEL 1
#define PF_LINE_JOIN_ROUND 2
#define PF_ARC_DIRECTION_CW 0
#define PF_ARC_DIRECTION_CCW 1
#define PF_RENDERER_OPTIONS_FLAGS_HAS_BACKGROUND_COLOR 0x1
#define PF_RENDERER_OPTIONS_FLAGS_SHOW_DEBUG_UI 0x2
#define PF_RENDERER_LEVEL_D3D9 0x1
#define PF_RENDERER_LEVEL_D3D11 0x2

typedef uint8_t PFLineCap;
typedef uint8_t PFLineJoin;
typedef uint8_t PFArcDirection;
typedef uint8_t PFRendererOptionsFlags;
typedef uint8_t PFRendererLevel;

/* c/src/lib.rs:116-170 (row-major matrices) */
typedef struct PFColorF { float r, g, b, a; } PFColorF;
typedef struct PFColorU { uint8_t r, g, b, a; } PFColorU;
typedef struct PFVector2F { float x, y; } PFVector2F;
typedef struct PFVector2I { int32_t x, y; } PFVector2I;
typedef struct PFRectF { PFVector2F origin, lower_right; } PFRectF;
typedef struct PFRectI { PFVector2I origin, lower_right; } PFRectI;
typedef struct PFMatrix2x2F { float m00, m01, m10, m11; } PFMatrix2x2F;
typedef struct PFTransform2F { PFMatrix2x2F matrix; PFVector2F vector; } PFTransform2F;

typedef struct PFCanvas *PFCanvasRef;
typedef struct PFPath *PFPathRef;
typedef struct PFFillStyle *PFFillStyleRef;
typedef struct PFResourceLoader *PFResourceLoaderRef;
typedef struct PFRenderTransform *PFRenderTransformRef;
typedef struct PFBuildOptions *PFBuildOptionsRef;
typedef struct pfh_scene *PFSceneRef;
typedef struct PFSceneProxy *PFSceneProxyRef;
typedef struct PFSVGScene *PFSVGSceneRef;
typedef struct PFHIPDevice *PFHIPDeviceRef;
typedef struct PFHIPDestFramebuffer *PFHIPDestFramebufferRef;
typedef struct PFHIPRenderer *PFHIPRendererRef;
typedef void *PFDestFramebufferRef;

/* c/src/lib.rs:203-214 */
typedef struct PFRendererMode { PFRendererLevel level; } PFRendererMode;
typedef struct PFRendererOptions {
    PFDestFramebufferRef dest;  /* a PFHIPDestFramebufferRef */
    PFColorF background_color;
    PFRendererOptionsFlags flags;
} PFRendererOptions;

/* ------------------------------------------------------------ canvas
 * Canvas::new(size).get_context_2d (c/src/lib.rs:225-232). There is no font
 * context (text is not on this path), so PFCanvasCreate takes the size only. */
PFCanvasRef PFCanvasCreate(const PFVector2F *size);
void PFCanvasDestroy(PFCanvasRef canvas);
/* Consumes the canvas (c/src/lib.rs:269-274). */
PFSceneRef PFCanvasCreateScene(PFCanvasRef canvas);
void PFCanvasFillRect(PFCanvasRef canvas, const PFRectF *rect);
void PFCanvasStrokeRect(PFCanvasRef canvas, const PFRectF *rect);
void PFCanvasSetLineWidth(PFCanvasRef canvas, float new_line_width);
void PFCanvasSetLineCap(PFCanvasRef canvas, PFLineCap new_line_cap);
void PFCanvasSetLineJoin(PFCanvasRef canvas, PFLineJoin new_line_join);
void PFCanvasSetMiterLimit(PFCanvasRef canvas, float new_miter_limit);
void PFCanvasSetTransform(PFCanvasRef canvas, const PFTransform2F *transform);
void PFCanvasResetTransform(PFCanvasRef canvas);
void PFCanvasSave(PFCanvasRef canvas);
void PFCanvasRestore(PFCanvasRef canvas);
/* Do not consume the fill style (c/src/lib.rs:413-426). */
void PFCanvasSetFillStyle(PFCanvasRef canvas, PFFillStyleRef fill_style);
void PFCanvasSetStrokeStyle(PFCanvasRef canvas, PFFillStyleRef stroke_style);
/* Consume the path (c/src/lib.rs:428-441). */
void PFCanvasFillPath(PFCanvasRef canvas, PFPathRef path);
void PFCanvasStrokePath(PFCanvasRef canvas, PFPathRef path);

/* ------------------------------------------------------------ paths (Path2D) */
PFPathRef PFPathCreate(void);
void PFPathDestroy(PFPathRef path);
PFPathRef PFPathClone(PFPathRef path);
void PFPathMoveTo(PFPathRef path, const PFVector2F *to);
void PFPathLineTo(PFPathRef path, const PFVector2F *to);
void PFPathQuadraticCurveTo(PFPathRef path, const PFVector2F *ctrl, const PFVector2F *to);
void PFPathBezierCurveTo(PFPathRef path, const PFVector2F *ctrl0, const PFVector2F *ctrl1,
                         const PFVector2F *to);
void PFPathArc(PFPathRef path, const PFVector2F *center, float radius, float start_angle,
               float end_angle, PFArcDirection direction);
void PFPathArcTo(PFPathRef path, const PFVector2F *ctrl, const PFVector2F *to, float radius);
void PFPathRect(PFPathRef path, const PFRectF *rect);
void PFPathEllipse(PFPathRef path, const PFVector2F *center, const PFVector2F *axes,
                   float rotation, float start_angle, float end_angle);
void PFPathClosePath(PFPathRef path);

/* ------------------------------------------------------------ fill styles */
PFFillStyleRef PFFillStyleCreateColor(const PFColorU *color);
void PFFillStyleDestroy(PFFillStyleRef fill_style);

/* ------------------------------------------------------------ resources
 * The renderer reads one resource: the area LUT (textures/area-lut.png). The
 * embedded loader holds a copy; the filesystem loader decodes the PNG from
 * <path>/textures/area-lut.png (NULL if it cannot). */
PFResourceLoaderRef PFEmbeddedResourceLoaderCreate(void);
PFResourceLoaderRef PFFilesystemResourceLoaderFromPath(const char *path);
void PFResourceLoaderDestroy(PFResourceLoaderRef loader);

/* ------------------------------------------------------------ HIP device and renderer
 * PFGLDeviceCreate / PFGLDestFramebufferCreateFullWindow / PFGLRendererCreate
 * (c/src/lib.rs:573-625) for a HIP device. */
PFHIPDeviceRef PFHIPDeviceCreate(int32_t device_ordinal);
void PFHIPDeviceDestroy(PFHIPDeviceRef device);
PFHIPDestFramebufferRef PFHIPDestFramebufferCreateFullWindow(const PFVector2I *window_size);
void PFHIPDestFramebufferDestroy(PFHIPDestFramebufferRef dest_framebuffer);
/* Takes ownership of device and options->dest, not of resources
 * (c/src/lib.rs:601-615). Either level: PF_RENDERER_LEVEL_D3D11 bins on the
 * GPU; PF_RENDERER_LEVEL_D3D9 tiles on the CPU when the scene is built and
 * rasterizes masks on the GPU (solid paints and SrcOver only). NULL if the
 * device cannot be used. */
PFHIPRendererRef PFHIPRendererCreate(PFHIPDeviceRef device, PFResourceLoaderRef resources,
                                     const PFRendererMode *mode,
                                     const PFRendererOptions *options);
void PFHIPRendererDestroy(PFHIPRendererRef renderer);
PFHIPDeviceRef PFHIPRendererGetDevice(PFHIPRendererRef renderer);
/* The destination framebuffer, RGBA8, rows of `stride` bytes (0: width * 4),
 * top row first. Returns 0 on success. */
int32_t PFHIPRendererReadPixels(PFHIPRendererRef renderer, uint8_t *rgba, size_t stride);
/* The message of the last failure ("" if none). */
const char *PFHIPRendererGetLastError(PFHIPRendererRef renderer);

/* ------------------------------------------------------------ build options */
PFRenderTransformRef PFRenderTransformCreate2D(const PFTransform2F *transform);
void PFRenderTransformDestroy(PFRenderTransformRef transform);
PFBuildOptionsRef PFBuildOptionsCreate(void);
void PFBuildOptionsDestroy(PFBuildOptionsRef options);
/* Consumes the transform (c/src/lib.rs:765-770). */
void PFBuildOptionsSetTransform(PFBuildOptionsRef options, PFRenderTransformRef transform);
/* Recorded; a nonzero dilation or subpixel AA makes the build fail (neither
 * is on the D3D11-level path this library renders). */
void PFBuildOptionsSetDilation(PFBuildOptionsRef options, const PFVector2F *dilation);
void PFBuildOptionsSetSubpixelAAEnabled(PFBuildOptionsRef options, int subpixel_aa_enabled);

/* ------------------------------------------------------------ scenes */
void PFSceneDestroy(PFSceneRef scene);
/* Consumes the scene (c/src/lib.rs:790-796); the build runs at render time
 * (the D3D9 level's tiler on the calling thread and worker threads). The
 * level must be the renderer's. */
PFSceneProxyRef PFSceneProxyCreateFromSceneAndRayonExecutor(PFSceneRef scene,
                                                            PFRendererLevel level);
void PFSceneProxyDestroy(PFSceneProxyRef scene_proxy);
/* Does not take ownership of renderer or build_options (c/src/lib.rs:674-684). */
void PFSceneProxyBuildAndRenderHIP(PFSceneProxyRef scene_proxy, PFHIPRendererRef renderer,
                                   PFBuildOptionsRef build_options);

/* ------------------------------------------------------------ SVG
 * NULL on failure (c/src/lib.rs:804-829). */
PFSVGSceneRef PFSVGSceneCreateWithMemory(const char *bytes, size_t byte_len);
PFSVGSceneRef PFSVGSceneCreateWithPath(const char *path);
/* Destroys the SVG and returns its scene. */
PFSceneRef PFSVGSceneIntoScene(PFSVGSceneRef svg);

#ifdef __cplusplus
}
#endif

#endif /* PATHFINDER_H */
