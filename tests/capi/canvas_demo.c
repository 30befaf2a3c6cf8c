/* A C client of include/pathfinder.h: builds a canvas scene (a filled
 * rectangle, a stroked arc path, a stroked rectangle) and prints the number
 * of draw paths and the first path's points. With a device ordinal argument
 * it also renders the scene at 160 x 120 and writes the RGBA8 pixels to the
 * file named by the second argument. */
#include <stdio.h>
#include <stdlib.h>

#include "pathfinder.h"
#include "pfhip.h"

int main(int argc, char **argv) {
    PFVector2F size = {160.0f, 120.0f};
    PFCanvasRef canvas = PFCanvasCreate(&size);
    PFColorU red = {200, 30, 30, 255}, blue = {20, 60, 220, 200};
    PFFillStyleRef fill = PFFillStyleCreateColor(&red);
    PFFillStyleRef stroke = PFFillStyleCreateColor(&blue);
    PFCanvasSetFillStyle(canvas, fill);
    PFRectF rect = {{10.0f, 10.0f}, {70.0f, 50.0f}};
    PFCanvasFillRect(canvas, &rect);
    PFCanvasSetStrokeStyle(canvas, stroke);
    PFCanvasSetLineWidth(canvas, 6.0f);
    PFCanvasSetLineCap(canvas, PF_LINE_CAP_ROUND);
    PFPathRef path = PFPathCreate();
    PFVector2F p0 = {90.0f, 90.0f}, c = {120.0f, 60.0f};
    PFPathMoveTo(path, &p0);
    PFPathArc(path, &c, 30.0f, 0.0f, 3.14159265f, PF_ARC_DIRECTION_CW);
    PFCanvasStrokePath(canvas, path);
    PFCanvasSave(canvas);
    PFTransform2F t = {{1.0f, 0.0f, 0.0f, 1.0f}, {80.0f, 0.0f}};
    PFCanvasSetTransform(canvas, &t);
    PFCanvasSetLineJoin(canvas, PF_LINE_JOIN_BEVEL);
    PFCanvasStrokeRect(canvas, &rect);
    PFCanvasRestore(canvas);
    PFFillStyleDestroy(fill);
    PFFillStyleDestroy(stroke);
    PFSceneRef scene = PFCanvasCreateScene(canvas);

    pfh_path_arrays a;
    if (pfh_scene_path_arrays(scene, PFH_PATH_SOURCE_DRAW, &a) != PFH_OK) return 1;
    printf("paths %u\n", a.path_count);
    for (uint32_t i = 0; i < a.contour_ends[0]; i++)
        printf("%g %g\n", a.points[i].x, a.points[i].y);

    if (argc > 2) {
        PFHIPDeviceRef device = PFHIPDeviceCreate(atoi(argv[1]));
        PFResourceLoaderRef loader = PFEmbeddedResourceLoaderCreate();
        PFVector2I win = {160, 120};
        PFRendererMode mode = {PF_RENDERER_LEVEL_D3D11};
        PFRendererOptions options = {PFHIPDestFramebufferCreateFullWindow(&win),
                                     {1.0f, 1.0f, 1.0f, 1.0f},
                                     PF_RENDERER_OPTIONS_FLAGS_HAS_BACKGROUND_COLOR};
        PFHIPRendererRef renderer = PFHIPRendererCreate(device, loader, &mode, &options);
        PFResourceLoaderDestroy(loader);
        if (!renderer) {
            fprintf(stderr, "PFHIPRendererCreate failed\n");
            return 2;
        }
        PFSceneProxyRef proxy = PFSceneProxyCreateFromSceneAndRayonExecutor(scene,
                                                                            PF_RENDERER_LEVEL_D3D11);
        PFBuildOptionsRef build = PFBuildOptionsCreate();
        PFSceneProxyBuildAndRenderHIP(proxy, renderer, build);
        static uint8_t pixels[160 * 120 * 4];
        int rc = PFHIPRendererReadPixels(renderer, pixels, 0);
        if (rc != 0 || PFHIPRendererGetLastError(renderer)[0]) {
            fprintf(stderr, "render failed: %s\n", PFHIPRendererGetLastError(renderer));
            return 3;
        }
        FILE *f = fopen(argv[2], "wb");
        if (!f || fwrite(pixels, 1, sizeof(pixels), f) != sizeof(pixels)) return 4;
        fclose(f);
        PFBuildOptionsDestroy(build);
        PFSceneProxyDestroy(proxy);
        PFHIPRendererDestroy(renderer);
        return 0;
    }
    PFSceneDestroy(scene);
    return 0;
}
