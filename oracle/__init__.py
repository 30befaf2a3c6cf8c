"""CPU oracle for the D3D11-level compute rasterization path.

TEST INFRASTRUCTURE ONLY. The oracle is the checker: tests/, the smoke test
of __graft_entry__ and the cpu_baseline leg of bench.py load it; the product
(pathfinder_amd) never does.

  pf_oracle.c   the seven compute shaders restated with sequential-invocation
                semantics (liboracle.so, built by `make -C oracle`)
  packing.py    the scene packing of renderer/src/builder.rs
  exact.py      FP64 exact-area coverage of the diced outlines
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LUT_PATH = os.path.join(HERE, "..", "pathfinder_amd", "data", "area_lut.rgba")

MICROLINES, FILLS, TILES, ZBUFFER, FIRST_TILE_MAP, ALPHA_TILES, BACKDROPS, ALPHA_RANGE = range(8)
OPT_CLIP_STRIDE_LITERAL = 1
OPT_PLANT_DICE = 2  # planted dice mistakes, only for tests that show a pin catches them
OPT_PLANT_PAINT = 3  # planted paint mistakes, likewise
_ELEM = {MICROLINES: (np.uint32, 4), FILLS: (np.uint32, 3), TILES: (np.uint32, 4),
         ZBUFFER: (np.int32, 1), FIRST_TILE_MAP: (np.int32, 1), ALPHA_TILES: (np.uint32, 2),
         BACKDROPS: (np.int32, 3), ALPHA_RANGE: (np.uint32, 1)}

_lib = None


def build():
    subprocess.run(["make", "-B", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P, U32, I32 = C.c_void_p, C.c_uint32, C.c_int32
    sig = {
        "orc_create": (P, [P]),
        "orc_destroy": (None, [P]),
        "orc_begin_frame": (None, [P, I32, I32, I32, P]),
        "orc_upload_paints": (None, [P, P, U32]),
        "orc_upload_scene": (None, [P, P, U32, P, U32, P, U32, P, U32]),
        "orc_prepare_batch": (I32, [P, U32, U32, U32, U32, U32, P, P, P, P, P, U32, I32]),
        "orc_draw_batch": (I32, [P, U32]),
        "orc_end_frame": (None, [P]),
        "orc_set_option": (None, [P, I32, I32]),
        "orc_count": (C.c_int64, [P, U32, I32]),
        "orc_read": (C.c_int64, [P, U32, I32, P, C.c_size_t]),
        "orc_alpha_tile_count": (U32, [P]),
        "orc_read_mask": (C.c_int64, [P, P, C.c_size_t]),
        "orc_read_framebuffer": (C.c_int64, [P, P, C.c_size_t]),
        "orc_set_threads": (None, [I32]),
        "orc_f32_to_f16": (C.c_uint16, [C.c_float]),
        "orc_f16_to_f32": (C.c_float, [C.c_uint16]),
        "orc_compute_coverage": (None, [P, C.c_float, C.c_float, C.c_float, C.c_float, P]),
        "orc_upload_metadata": (I32, [P, P, U32]),
        "orc_allocate_page": (None, [P, U32, I32, I32]),
        "orc_upload_texels": (I32, [P, U32, I32, I32, I32, I32, P]),
        "orc_set_batch_color_texture": (I32, [P, U32, I32, U32]),
        "orc_stage_reset": (None, []),
        "orc_stage_times": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    _lib = lib
    return lib


# orc_metadata_entry: colour transform (m00 m01 m10 m11 tx ty), combine mode,
# base colour, filter kind and 20 parameters, blend mode
METADATA_DTYPE = np.dtype([("transform", np.float32, 6), ("combine", np.int32),
                           ("base_color", np.uint8, 4), ("filter_kind", np.int32),
                           ("filter_params", np.float32, 20), ("blend_mode", np.int32)])


def area_lut():
    return np.fromfile(LUT_PATH, dtype=np.uint8)


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


class Oracle:
    """Mirrors the renderer seam: begin_frame, upload_*, prepare/draw batches."""

    def __init__(self, lut=None, threads=None):
        self.lib = load()
        self._lut = np.ascontiguousarray(area_lut() if lut is None else lut, np.uint8)
        self.h = self.lib.orc_create(_p(self._lut))
        if threads is not None:
            self.lib.orc_set_threads(int(threads))

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_option(self, opt, value):
        self.lib.orc_set_option(self.h, opt, value)

    def begin_frame(self, width, height, background=None):
        bg = None if background is None else np.asarray(
            [background.r, background.g, background.b, background.a], np.float32)
        self.lib.orc_begin_frame(self.h, width, height, 0 if background is None else 1, _p(bg))
        self.width, self.height = width, height

    def upload_paints(self, rgba):
        a = np.ascontiguousarray(rgba, np.uint8).reshape(-1, 4)
        self.lib.orc_upload_paints(self.h, _p(a), len(a))

    def upload_metadata(self, entries):
        a = np.ascontiguousarray(entries, METADATA_DTYPE)
        if self.lib.orc_upload_metadata(self.h, _p(a), len(a)) != 0:
            raise ValueError("oracle: unknown filter kind")

    def allocate_page(self, page, width, height):
        self.lib.orc_allocate_page(self.h, page, width, height)

    def upload_texels(self, page, x, y, rgba):
        a = np.ascontiguousarray(rgba, np.uint8)
        if self.lib.orc_upload_texels(self.h, page, x, y, a.shape[1], a.shape[0], _p(a)) != 0:
            raise ValueError("oracle: texel rect outside page %d" % page)

    def upload_scene(self, draw, clip):
        arrs = []
        for s in (draw, clip):
            arrs.append(np.ascontiguousarray(s["points"], np.float32))
            arrs.append(np.ascontiguousarray(s["indices"], np.uint32))
        self.lib.orc_upload_scene(self.h, _p(arrs[0]), len(arrs[0]), _p(arrs[1]), len(arrs[1]),
                                  _p(arrs[2]), len(arrs[2]), _p(arrs[3]), len(arrs[3]))

    def prepare_batch(self, b):
        arrs = [np.ascontiguousarray(b[k]) for k in
                ("transform", "dice_metadata", "tile_path_info", "propagate_metadata",
                 "backdrops")]
        rc = self.lib.orc_prepare_batch(self.h, b["batch_id"], b["path_source"], b["path_count"],
                                        b["tile_count"], b["segment_count"], _p(arrs[0]),
                                        _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), _p(arrs[4]),
                                        len(arrs[4]), b["clip_batch_id"])
        if rc != 0:
            raise RuntimeError("oracle prepare_batch failed: %d" % rc)
        if self.lib.orc_set_batch_color_texture(self.h, b["batch_id"],
                                                int(b.get("color_texture_page", -1)),
                                                int(b.get("color_texture_sampling_flags", 0))):
            raise RuntimeError("oracle set_batch_color_texture failed")

    def draw_batch(self, batch_id):
        if self.lib.orc_draw_batch(self.h, batch_id) != 0:
            raise RuntimeError("oracle draw_batch failed")

    def end_frame(self):
        self.lib.orc_end_frame(self.h)

    def read(self, batch_id, what):
        dt, per = _ELEM[what]
        n = self.lib.orc_count(self.h, batch_id, what)
        if n < 0:
            raise KeyError(batch_id)
        out = np.zeros(max(n * per, 1), dt)
        got = self.lib.orc_read(self.h, batch_id, what, _p(out), out.nbytes)
        if got < 0:
            raise RuntimeError("oracle read failed")
        return out[:n * per].reshape(-1, per) if per > 1 else out[:n]

    def read_mask(self):
        n = self.lib.orc_alpha_tile_count(self.h) * 64
        out = np.zeros(max(n, 1), np.uint32)
        self.lib.orc_read_mask(self.h, _p(out), out.size)
        return out[:n]

    def read_framebuffer(self):
        out = np.zeros(self.width * self.height, np.uint32)
        self.lib.orc_read_framebuffer(self.h, _p(out), out.size)
        return out.view(np.uint8).reshape(self.height, self.width, 4)

    def render(self, cmds, width, height, background):
        """Runs a packed command stream (packing.build or CommandList dicts),
        dropping the previous frame's batches."""
        self.end_frame()
        self.begin_frame(width, height, background)
        for page, (w, h) in cmds.get("pages", {}).items():
            self.allocate_page(page, w, h)
        for page, x, y, rgba in cmds.get("texels", ()):
            self.upload_texels(page, x, y, rgba)
        if "metadata" in cmds:
            self.upload_metadata(cmds["metadata"])
        else:
            self.upload_paints(cmds["paints"])
        self.upload_scene(cmds["draw"], cmds["clip"])
        for kind, b in cmds["batches"]:
            self.prepare_batch(b)
            if kind == 4:
                self.draw_batch(b["batch_id"])
        return self


STAGES = ("dice", "bound", "bin", "propagate", "fill", "sort", "tile",
          "d3d9_tiler", "d3d9_raster", "d3d9_composite")


def stage_reset():
    load().orc_stage_reset()


def stage_seconds():
    """Wall-clock seconds per stage since stage_reset() (D3D11 stages, then the
    D3D9 level's phases)."""
    out = np.zeros(len(STAGES), np.float64)
    load().orc_stage_times(_p(out))
    return dict(zip(STAGES, out.tolist()))


def coverage(fx, fy, tx, ty, oracle=None):
    o = oracle or Oracle()
    out = np.zeros(4, np.float32)
    o.lib.orc_compute_coverage(o.h, fx, fy, tx, ty, _p(out))
    return out
