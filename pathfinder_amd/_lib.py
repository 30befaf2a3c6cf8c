"""ctypes binding of libpfhip.so (include/pfhip.h).

The library is the product: this module only loads it and declares the C
signatures. It raises if the library has not been built; there is no
fallback implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfhip.so")

# pfh_status
OK = 0
STATUS_NAMES = {
    0: "PFH_OK",
    1: "PFH_ERROR_INVALID_ARGUMENT",
    2: "PFH_ERROR_HIP",
    3: "PFH_ERROR_OUT_OF_MEMORY",
    4: "PFH_ERROR_STATE",
    5: "PFH_ERROR_NO_DEVICE",
    6: "PFH_ERROR_LIMIT",
}

RENDERER_LEVEL_D3D9 = 1
RENDERER_LEVEL_D3D11 = 2
PATH_SOURCE_DRAW = 0
PATH_SOURCE_CLIP = 1
FILL_RULE_WINDING = 0
FILL_RULE_EVEN_ODD = 1
POINT_CONTROL_0 = 0x01
POINT_CONTROL_1 = 0x02
CURVE_IS_QUADRATIC = 0x80000000
CURVE_IS_CUBIC = 0x40000000

COMMAND_START = 0
COMMAND_UPLOAD_TEXTURE_METADATA = 1
COMMAND_UPLOAD_SCENE_D3D11 = 2
COMMAND_PREPARE_CLIP_TILES_D3D11 = 3
COMMAND_DRAW_TILES_D3D11 = 4
COMMAND_FINISH = 5
COMMAND_ALLOCATE_TEXTURE_PAGE = 6
COMMAND_UPLOAD_TEXEL_DATA = 7
COMMAND_ADD_FILLS_D3D9 = 8
COMMAND_FLUSH_FILLS_D3D9 = 9
COMMAND_DRAW_TILES_D3D9 = 10

DEBUG_MICROLINES = 0
DEBUG_FILLS = 1
DEBUG_TILES = 2
DEBUG_Z_BUFFER = 3
DEBUG_FIRST_TILE_MAP = 4
DEBUG_ALPHA_TILES = 5
DEBUG_BACKDROPS = 6
DEBUG_ALPHA_RANGE = 7
DEBUG_COVERAGE = 8
DEBUG_COVERAGE_STATES = 9

SAMPLING_REPEAT_U, SAMPLING_REPEAT_V, SAMPLING_NEAREST_MIN, SAMPLING_NEAREST_MAG = 1, 2, 4, 8
COMBINE_NONE, COMBINE_SRC_IN, COMBINE_DEST_IN = 0, 1, 2
FILTER_NONE, FILTER_RADIAL_GRADIENT, FILTER_BLUR, FILTER_COLOR_MATRIX = 0, 1, 2, 3
GRADIENT_LINEAR, GRADIENT_RADIAL = 0, 1
GRADIENT_WRAP_CLAMP, GRADIENT_WRAP_REPEAT = 0, 1
PATTERN_REPEAT_X, PATTERN_REPEAT_Y, PATTERN_NO_SMOOTHING = 1, 2, 4
PAINT_COMPOSITE_SRC_IN, PAINT_COMPOSITE_DEST_IN = 0, 1
INVALID_PAINT = 0xFFFFFFFF
CHECKPOINT_COVERAGE, CHECKPOINT_STATES, CHECKPOINT_RECOMPUTE = 0, 1, 2

NOALLOC_SCRATCH_BYTES = 64 << 20  # PFH_NOALLOC_SCRATCH_BYTES


class PathArraysC(C.Structure):
    _fields_ = [("points", C.c_void_p), ("flags", C.c_void_p), ("point_count", C.c_uint32),
                ("contour_ends", C.c_void_p), ("contour_count", C.c_uint32),
                ("path_contour_ends", C.c_void_p), ("path_count", C.c_uint32),
                ("paints", C.c_void_p), ("fill_rules", C.c_void_p), ("clip_paths", C.c_void_p),
                ("blend_modes", C.c_void_p)]


class SvgInfo(C.Structure):
    _fields_ = [("view_box", C.c_float * 4), ("fills", C.c_uint32), ("strokes", C.c_uint32),
                ("skipped_elements", C.c_uint32), ("unsupported_paints", C.c_uint32)]


class StrokeStyle(C.Structure):
    _fields_ = [("line_width", C.c_float), ("line_cap", C.c_int32), ("line_join", C.c_int32),
                ("miter_limit", C.c_float)]


class Vector2F(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class RectF(C.Structure):
    _fields_ = [("origin", Vector2F), ("lower_right", Vector2F)]


class ColorU(C.Structure):
    _fields_ = [("r", C.c_uint8), ("g", C.c_uint8), ("b", C.c_uint8), ("a", C.c_uint8)]


class ColorF(C.Structure):
    _fields_ = [("r", C.c_float), ("g", C.c_float), ("b", C.c_float), ("a", C.c_float)]


class Transform2F(C.Structure):
    _fields_ = [("m00", C.c_float), ("m01", C.c_float), ("m10", C.c_float), ("m11", C.c_float),
                ("tx", C.c_float), ("ty", C.c_float)]


class Segments(C.Structure):
    _fields_ = [("points", C.c_void_p), ("point_count", C.c_uint32),
                ("indices", C.c_void_p), ("index_count", C.c_uint32)]


class TileBatch(C.Structure):
    _fields_ = [("batch_id", C.c_uint32), ("path_source", C.c_uint32),
                ("path_count", C.c_uint32), ("tile_count", C.c_uint32),
                ("segment_count", C.c_uint32), ("transform", Transform2F),
                ("dice_metadata", C.c_void_p), ("tile_path_info", C.c_void_p),
                ("propagate_metadata", C.c_void_p), ("backdrops", C.c_void_p),
                ("backdrop_count", C.c_uint32), ("clip_batch_id", C.c_int32),
                ("color_texture_page", C.c_int32),
                ("color_texture_sampling_flags", C.c_uint32)]


class Filter(C.Structure):
    _fields_ = [("kind", C.c_int32), ("params", C.c_float * 20)]


class TextureMetadataEntry(C.Structure):
    _fields_ = [("color_0_transform", Transform2F), ("color_0_combine_mode", C.c_int32),
                ("base_color", ColorU), ("filter", Filter), ("blend_mode", C.c_int32)]


class FillD3D9(C.Structure):
    _fields_ = [("from_x", C.c_uint16), ("from_y", C.c_uint16), ("to_x", C.c_uint16),
                ("to_y", C.c_uint16), ("alpha_tile_index", C.c_uint32)]


class TileObjectD3D9(C.Structure):
    _fields_ = [("tile_x", C.c_int16), ("tile_y", C.c_int16), ("alpha_tile_id", C.c_uint32),
                ("path_id", C.c_uint32), ("color", C.c_uint16), ("ctrl", C.c_uint8),
                ("backdrop", C.c_int8)]


class ClipD3D9(C.Structure):
    _fields_ = [("dest_tile_id", C.c_uint32), ("dest_backdrop", C.c_int32),
                ("src_tile_id", C.c_uint32), ("src_backdrop", C.c_int32)]


class ColorStop(C.Structure):
    _fields_ = [("offset", C.c_float), ("color", ColorU)]


class Gradient(C.Structure):
    _fields_ = [("geometry", C.c_int32), ("from_", Vector2F), ("to", Vector2F),
                ("radius0", C.c_float), ("radius1", C.c_float), ("transform", Transform2F),
                ("stops", C.POINTER(ColorStop)), ("stop_count", C.c_uint32), ("wrap", C.c_int32)]


class Pattern(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("pixels", C.c_void_p),
                ("transform", Transform2F), ("flags", C.c_uint32), ("filter", Filter)]


class RectI(C.Structure):
    _fields_ = [("min_x", C.c_int32), ("min_y", C.c_int32), ("max_x", C.c_int32),
                ("max_y", C.c_int32)]


class DrawTileBatchD3D9(C.Structure):
    _fields_ = [("tiles", C.c_void_p), ("tile_count", C.c_uint32), ("clips", C.c_void_p),
                ("clip_count", C.c_uint32), ("z_buffer", C.c_void_p), ("z_buffer_rect", RectI)]


class ContextOptions(C.Structure):
    _fields_ = [("device", C.c_int32), ("level", C.c_int32), ("area_lut_rgba", C.c_void_p),
                ("initial_microline_capacity", C.c_uint32),
                ("initial_fill_capacity", C.c_uint32), ("enable_timing", C.c_uint32)]


class FrameOptions(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("has_background", C.c_int32),
                ("background_color", ColorF)]


class RenderStats(C.Structure):
    _fields_ = [("path_count", C.c_uint64), ("fill_count", C.c_uint64),
                ("alpha_tile_count", C.c_uint64), ("total_tile_count", C.c_uint64),
                ("drawcall_count", C.c_uint64), ("microline_count", C.c_uint64),
                ("segment_count", C.c_uint64), ("column_count", C.c_uint64),
                ("linked_tile_count", C.c_uint64), ("composited_tile_count", C.c_uint64),
                ("composited_alpha_tile_count", C.c_uint64),
                ("clipped_alpha_tile_count", C.c_uint64),
                ("gpu_bytes_allocated", C.c_uint64), ("overflow_replays", C.c_uint32),
                ("pad", C.c_uint32)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char_p), ("launches", C.c_uint64), ("total_ms", C.c_double)]


PROFILE_OFF = 0
PROFILE_ALL = 1
PROFILE_ONE = 2
PROFILE_SERIAL = 3


class BuildOptions(C.Structure):
    _fields_ = [("transform", Transform2F)]


class Outline(C.Structure):
    _fields_ = [("points", C.c_void_p), ("flags", C.c_void_p), ("point_count", C.c_uint32),
                ("contour_ends", C.c_void_p), ("contour_count", C.c_uint32)]


class LaunchShape(C.Structure):
    _fields_ = [("grid_cap", C.c_uint32), ("fill_grid", C.c_uint32)]


P = C.c_void_p
U32 = C.c_uint32
I32 = C.c_int32
SZ = C.c_size_t

# name -> (restype, argtypes); every entry point of include/pfhip.h
SIGNATURES = {
    "pfh_context_create": (I32, [C.POINTER(ContextOptions), C.POINTER(P)]),
    "pfh_context_destroy": (None, [P]),
    "pfh_last_error": (C.c_char_p, [P]),
    "pfh_begin_scene": (I32, [P, C.POINTER(FrameOptions)]),
    "pfh_upload_texture_metadata": (I32, [P, P, U32]),
    "pfh_upload_scene": (I32, [P, C.POINTER(Segments), C.POINTER(Segments)]),
    "pfh_prepare_tiles": (I32, [P, C.POINTER(TileBatch)]),
    "pfh_prepare_and_draw_tiles": (I32, [P, C.POINTER(TileBatch)]),
    "pfh_end_scene": (I32, [P]),
    "pfh_rerender_scene": (I32, [P]),
    "pfh_synchronize": (I32, [P]),
    "pfh_read_framebuffer": (I32, [P, P, SZ]),
    "pfh_read_debug": (I32, [P, U32, I32, P, SZ, C.POINTER(SZ)]),
    "pfh_read_mask": (I32, [P, P, SZ, C.POINTER(SZ)]),
    "pfh_get_stats": (I32, [P, C.POINTER(RenderStats)]),
    "pfh_profile_kernel_count": (U32, []),
    "pfh_profile_kernel_name": (C.c_char_p, [U32]),
    "pfh_profile_set": (I32, [P, I32, I32]),
    "pfh_profile_reset": (I32, [P]),
    "pfh_profile_read": (I32, [P, C.POINTER(KernelTime), U32, C.POINTER(U32)]),
    "pfh_set_gradients": (I32, [P, I32]),
    "pfh_set_graphs": (I32, [P, I32]),
    "pfh_set_launch_shape": (I32, [P, C.POINTER(LaunchShape)]),
    "pfh_get_launch_shape": (I32, [P, C.POINTER(LaunchShape)]),
    "pfh_autotune": (I32, [P, U32, C.POINTER(LaunchShape)]),
    "pfh_framebuffer_to_float": (I32, [P, P]),
    "pfh_backward": (I32, [P, P, P, U32, P, U32]),
    "pfh_backward_noalloc": (I32, [P, P, P, U32, P, U32]),
    "pfh_upload_texture_metadata_entries": (I32, [P, C.POINTER(TextureMetadataEntry), U32]),
    "pfh_allocate_texture_page": (I32, [P, U32, I32, I32]),
    "pfh_upload_texel_data": (I32, [P, U32, C.POINTER(RectI), P]),
    "pfh_scene_create": (P, []),
    "pfh_scene_destroy": (None, [P]),
    "pfh_scene_set_view_box": (None, [P, C.POINTER(RectF)]),
    "pfh_scene_push_paint": (U32, [P, C.POINTER(ColorU)]),
    "pfh_scene_set_draw_points": (I32, [P, P, U32]),
    "pfh_set_gradient_checkpoint": (I32, [P, I32]),
    "pfh_get_graphs": (I32, [P, C.POINTER(I32)]),
    "pfh_scene_set_paint_colors": (I32, [P, P, U32]),
    "pfh_scene_push_gradient_paint": (U32, [P, C.POINTER(ColorU), C.POINTER(Gradient), I32]),
    "pfh_scene_push_pattern_paint": (U32, [P, C.POINTER(ColorU), C.POINTER(Pattern), I32]),
    "pfh_scene_push_clip_path": (U32, [P, C.POINTER(Outline), I32, I32]),
    "pfh_scene_push_draw_path": (U32, [P, C.POINTER(Outline), U32, I32, I32, I32]),
    "pfh_scene_push_draw_paths": (U32, [P, P, P, U32, P, U32, P, U32, P, P, P, P]),
    "pfh_scene_push_clip_paths": (U32, [P, P, P, U32, P, U32, P, U32, P, P]),
    "pfh_scene_draw_path_count": (U32, [P]),
    "pfh_scene_path_arrays": (I32, [P, U32, C.POINTER(PathArraysC)]),
    "pfh_scene_paints": (I32, [P, C.POINTER(P), C.POINTER(U32)]),
    "pfh_scene_get_view_box": (I32, [P, C.POINTER(RectF)]),
    "pfh_svg_scene_create": (P, [C.c_char_p, SZ, C.POINTER(SvgInfo)]),
    "pfh_stroke_outline": (I32, [P, P, C.POINTER(StrokeStyle), P, C.POINTER(P)]),
    "pfh_outline_buffer_view": (C.POINTER(Outline), [P]),
    "pfh_outline_buffer_destroy": (None, [P]),
    "pfh_scene_build": (I32, [P, C.POINTER(BuildOptions), C.POINTER(P)]),
    "pfh_command_list_destroy": (None, [P]),
    "pfh_command_list_len": (U32, [P]),
    "pfh_command_kind_at": (I32, [P, U32]),
    "pfh_command_batch": (I32, [P, U32, C.POINTER(TileBatch)]),
    "pfh_command_segments": (I32, [P, U32, C.POINTER(Segments), C.POINTER(Segments)]),
    "pfh_command_paints": (I32, [P, U32, C.POINTER(P), C.POINTER(U32)]),
    "pfh_command_texture_metadata": (I32, [P, U32, C.POINTER(P), C.POINTER(U32)]),
    "pfh_command_texture_page": (I32, [P, U32, C.POINTER(U32), C.POINTER(I32),
                                       C.POINTER(I32)]),
    "pfh_command_texel_data": (I32, [P, U32, C.POINTER(U32), C.POINTER(RectI), C.POINTER(P)]),
    "pfh_command_fills_d3d9": (I32, [P, U32, C.POINTER(P), C.POINTER(U32)]),
    "pfh_command_batch_d3d9": (I32, [P, U32, C.POINTER(DrawTileBatchD3D9)]),
    "pfh_scene_build_d3d9": (I32, [P, C.POINTER(BuildOptions), C.POINTER(P)]),
    "pfh_d3d9_add_fills": (I32, [P, P, U32]),
    "pfh_d3d9_flush_fills": (I32, [P]),
    "pfh_d3d9_draw_tiles": (I32, [P, C.POINTER(DrawTileBatchD3D9)]),
    "pfh_render_commands": (I32, [P, P]),
    "pfh_scene_build_and_render": (I32, [P, P, C.POINTER(FrameOptions),
                                         C.POINTER(BuildOptions)]),
    "pfh_scene_build_and_render_d3d9": (I32, [P, P, C.POINTER(FrameOptions),
                                         C.POINTER(BuildOptions)]),
}

_lib = None
_hip = None


def _share_torch_hip_runtime():
    """Loads the HIP runtime that torch ships, if torch is installed.

    libpfhip.so needs libamdhip64.so.7. torch ships its own copy with that
    soname, and its libraries ask for it as `libamdhip64.so` from their own
    directory. Loaded first under RTLD_GLOBAL, torch's copy satisfies both
    requests, so the process holds one HIP runtime whichever of torch and
    this library loads first. Two runtimes in one process do not share
    devices: the second to initialise finds no GPU, and torch's device
    pointers and streams would be foreign to this library.
    """
    global _hip
    import importlib.util
    spec = importlib.util.find_spec("torch")
    for d in (spec.submodule_search_locations or []) if spec else []:
        path = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(path):
            _hip = C.CDLL(path, mode=C.RTLD_GLOBAL)
            return


def hip_runtime():
    """The process's HIP runtime (the one libpfhip.so is bound to)."""
    load()
    return _hip if _hip is not None else C.CDLL("libamdhip64.so.7")


def load():
    """Loads libpfhip.so once; raises if it is missing (build it first)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("pathfinder_amd: %s is not built; run "
                           "`python -m pathfinder_amd.build`" % LIB_PATH)
    _share_torch_hip_runtime()
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PfhError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))
        self.status = status


def check(status, ctx=None):
    if status != OK:
        msg = ""
        if ctx:
            raw = load().pfh_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise PfhError(status, msg)
