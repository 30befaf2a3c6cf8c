#!/usr/bin/env python3
"""Frame benchmark of the D3D11-level compute path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload 4k|small] [--no-cpu-baseline]

A step is one frame of the 3840x2160 workload (20,000 paths, 500 clip paths,
about 5M microlines): every batch runs bound, dice, bin, propagate, sort and
fill, and the draw batch composites into the RGBA8 framebuffer. The scene and
batch metadata are resident in HBM before the timed region (the frame is
re-rendered from device memory); packing on the host is not in the timed
region. Prints one JSON line.

value = microlines per second of frame time (Msegments/s); ms_per_step = frame
time, the timed frames running with no events on the streams. roofline: the
kernel with the largest time per frame when every kernel runs alone (a
warm-up frame with the stages serialized), its algorithmic bytes per frame
(formulas below) over its time per frame measured with HIP events on the
launch stream in as many frames again after the timed region (where it
overlaps other kernels as in the timed frames).

diff: the differentiable render (1024 closed cubic paths, 512 x 512, points
and colours as leaves, L2 loss against the image of a copy moved by
N(0, 4 px)), measured in a child process started before this one touches the
GPU: forward = one frame with the coverage checkpoint (re-rendered from
device memory), backward = pfh_backward (composite and coverage adjoint),
ratio = (forward + backward) / forward, all device time with the scene
resident. step_ms is one whole optimisation step through torch.autograd
(the points copied to the host, the C++ repack of a scene kept alive with its
points and colours replaced in place, uploads, forward, loss, backward,
Adam).

batch_glyphs / batch_canvases: many items of uneven sizes packed into frames
(pathfinder_amd.batch): 50,000 glyphs at 8-72 px on 4096 x 4096 atlas pages,
and 512 canvases of 64^2-1024^2 with 10-2000 paths each on 8192-wide frames
(a tile budget per frame keeps the 24-bit alpha tile ids). One frame renders
all of its items with the pipeline's fixed launch sequence; each frame is
autotuned (launch shape, then eager launches or graph replay, whichever is
faster; `tuned`, with the eager default's times beside). aggregate
Msegments/s = all microlines / the frames' summed device time (re-rendered,
resident). per_scene_ms_in_batch: an item's latency inside the batch, the
device time of the frame that renders it (p50 / p99 over items);
per_item_alone_ms: 32 sampled items rendered alone, for comparison.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "4K frame ms & Msegments/s vs llvmpipe ref; fwd+bwd/fwd ratio at ≤1 LSB"
HBM_PEAK_GBS = 8000.0
# HBM bytes per kernel per frame from two rocprofv3 counter runs of this
# benchmark (FETCH_SIZE, WRITE_SIZE), scaled by a device copy of known size
TRAFFIC_TABLE = {"4k": os.path.join(ROOT, "profiles", "traffic_4k.md")}


def algorithmic_bytes(st, fb_tiles, pixels, points, load_action):
    """Compulsory bytes per frame of each kernel, every buffer touched once.

    Stage totals follow the per-stage model of the reference's buffers
    (microline 16 B, fill 12 B, tile 16 B, alpha tile 8 B, backdrop 12 B,
    mask 256 B per alpha tile, RGBA8 pixels); count passes are charged for
    what they read and the counts they write."""
    S, P, M, F = st["segment_count"], st["path_count"], st["microline_count"], st["fill_count"]
    T, C, A = st["total_tile_count"], st["column_count"], st["alpha_tile_count"]
    Lp, L, AL = st["linked_tile_count"], st["composited_tile_count"], \
        st["composited_alpha_tile_count"]
    Ac = st["clipped_alpha_tile_count"]
    Np, Fb, N = points, fb_tiles, pixels
    return {
        "dice_count": 8 * S + 8 * Np + 16 * P + 4 * S,
        "dice_emit": 8 * S + 8 * Np + 16 * P + 4 * S + 16 * M,
        "bound": 16 * P + 16 * T,
        "bin_count": 16 * M + 48 * P + 4 * M,
        "bin_emit": 16 * M + 48 * P + 4 * M + 12 * F + 4 * F,
        # fill tiles and words read, sorted words and links written, and per
        # tile with fills its head, run and flag (one run per alpha tile)
        "fill_link": (4 + 8) * F + (8 + 4) * F + (4 + 8 + 1) * A,
        "propagate_count": 12 * C + 16 * T + 4 * C,
        "propagate_emit": 12 * C + 16 * T + 12 * T + 8 * A + 4 * C,
        "sort_tiles": 8 * Fb + 4 * Lp + 4 * L,
        "fill": 24 * A + 12 * F + 256 * A + 256 * Ac,
        "tile": 4 * Fb + 16 * L + 256 * AL + 4 * N * (2 if load_action else 1),
    }


def copy_calibration(nbytes):
    """Three device-to-device copies of nbytes with the HIP runtime's own copy
    kernel: a copy of known size for scaling memory counters."""
    import ctypes
    from pathfinder_amd import _lib
    hip = _lib.hip_runtime()
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print("bench: copy calibration: 3 device copies of %d bytes" % nbytes, file=sys.stderr,
          flush=True)


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def median_run(fn, runs=5):
    """One warm-up call, then the median wall-clock time of `runs` calls and
    the per-stage split of that run (oracle.stage_seconds)."""
    import oracle
    fn()
    results = []
    for _ in range(runs):
        oracle.stage_reset()
        t0 = time.perf_counter()
        fn()
        results.append((time.perf_counter() - t0, oracle.stage_seconds()))
    results.sort(key=lambda r: r[0])
    return results[runs // 2]


def cpu_baseline(scene, size, bg, microlines, sample_scene, level):
    """The CPU restatement of one level timed on the host: the whole frame at
    the host's cores (at most 16, the box's CPU share) and a tenth-size sample
    of the same generator at one thread; 1 warm-up + median of 5 each."""
    import oracle
    from oracle import d3d9, packing
    nproc, model = cpu_info()
    threads = min(16, nproc)

    def runner(sc, sz, th):
        if level == "d3d9":
            return lambda: d3d9.render(sc, sz, bg, threads=th)
        cmds = packing.build(sc)
        o = oracle.Oracle(threads=th)
        return lambda: o.render(cmds, sz[0], sz[1], bg)

    def stages_ms(st):
        keys = [k for k in oracle.STAGES if k.startswith("d3d9_") == (level == "d3d9")]
        return {k: round(st[k] * 1e3, 1) for k in keys}

    dt, st = median_run(runner(scene, size, threads))
    s_scene, s_size, s_ml = sample_scene
    dt1, st1 = median_run(runner(s_scene, s_size, 1))
    what = ("the D3D11 oracle (the seven shaders restated in C; dice, bin and propagate "
            "sequential, the other stages threaded)" if level == "d3d11" else
            "the D3D9-level CPU restatement (tiler.rs flattening and tiling on a thread per "
            "path, binary16 mask raster, per-tile RGBA8 composite), rated on the D3D11 "
            "microline count")
    return {"value": round(microlines / dt / 1e6, 4), "unit": "Msegments/s", "cores": threads,
            "kind": "port", "nproc": nproc, "cpu_model": model, "runs": 5, "warmup": 1,
            "frame_ms": round(dt * 1e3, 1), "stages_ms": stages_ms(st),
            "one_thread": {"value": round(s_ml / dt1 / 1e6, 4), "frame_ms": round(dt1 * 1e3, 1),
                           "stages_ms": stages_ms(st1), "microlines": s_ml},
            "sample": "whole frames of the same workload through %s, median of 5 after a "
                      "warm-up at %d threads; one_thread: 2,000 paths and 50 clip paths of the "
                      "same generator (a tenth of the frame), median of 5 after a warm-up"
                      % (what, threads)}


def diff_leg(steps, warmup):
    """The differentiable render's timings (runs in its own process: torch's
    HIP runtime is loaded before the library)."""
    import torch

    from pathfinder_amd import diff, synth
    d = synth.diff_paths()
    w, h = d["size"]
    shapes = diff.Shapes(d["flags"], d["contour_ends"], d["path_contour_ends"])
    dr = diff.DiffRenderer(shapes, w, h, (1.0, 1.0, 1.0, 1.0))
    with torch.no_grad():
        target = dr.render(torch.tensor(d["target_points"]), torch.tensor(d["colors"]))
    p = torch.tensor(d["points"], device="cuda", requires_grad=True)
    c = torch.tensor(d["colors"], device="cuda", requires_grad=True)
    diff.fit(dr, p, c, target, max(warmup, 1))  # warm-up steps of the same fit
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = diff.fit(dr, p, c, target, steps)
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) / steps * 1e3
    # device times of the last frame: forward (re-rendered) and backward
    img = diff.render(dr, p, c)
    g = (2.0 * (img - target) / img.numel()).detach().contiguous()
    gp = torch.zeros((len(shapes.packed), 2), dtype=torch.float32, device="cuda")
    gc = torch.zeros((shapes.path_count, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    r = dr.renderer
    for _ in range(warmup):
        r.rerender()
    r.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.rerender()
    r.synchronize()
    fwd_ms = (time.perf_counter() - t0) / steps * 1e3
    for _ in range(warmup):
        r.backward(g.data_ptr(), gp.data_ptr(), len(shapes.packed), gc.data_ptr(),
                   shapes.path_count)
    t0 = time.perf_counter()
    for _ in range(steps):
        r.backward(g.data_ptr(), gp.data_ptr(), len(shapes.packed), gc.data_ptr(),
                   shapes.path_count)
    bwd_ms = (time.perf_counter() - t0) / steps * 1e3
    st = r.stats()
    # what the forward keeps for the backward, and its cost: the float
    # checkpoint of C, 2-bit states written by the forward fill (the
    # library's default, timed above), or nothing with the fill run again in
    # the backward
    from pathfinder_amd import _lib as L
    modes = {}
    for name, mode, per_tile in (("coverage", L.CHECKPOINT_COVERAGE, 1024),
                                 ("states", L.CHECKPOINT_STATES, 64),
                                 ("recompute", L.CHECKPOINT_RECOMPUTE, 64)):
        r.set_gradient_checkpoint(mode)
        diff.render(dr, p, c)
        torch.cuda.synchronize()
        for _ in range(warmup):
            r.rerender()
        r.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.rerender()
        r.synchronize()
        f_ms = (time.perf_counter() - t0) / steps * 1e3
        for _ in range(warmup):
            r.backward(g.data_ptr(), gp.data_ptr(), len(shapes.packed), gc.data_ptr(),
                       shapes.path_count)
        t0 = time.perf_counter()
        for _ in range(steps):
            r.backward(g.data_ptr(), gp.data_ptr(), len(shapes.packed), gc.data_ptr(),
                       shapes.path_count)
        b_ms = (time.perf_counter() - t0) / steps * 1e3
        modes[name] = {"fwd_ms": round(f_ms, 4), "bwd_ms": round(b_ms, 4),
                       "fwd_bwd_over_fwd": round((f_ms + b_ms) / f_ms, 3),
                       "kept_bytes": st["alpha_tile_count"] * per_tile if mode != 2 else 0,
                       "backward_scratch_bytes": st["alpha_tile_count"] * 64 if mode == 2 else 0}
    dr.close()
    return {"workload": "diff-1024-cubic-paths-512", "width": w, "height": h,
            "paths": shapes.path_count, "points": shapes.point_count,
            "microlines": st["microline_count"], "alpha_tiles": st["alpha_tile_count"],
            "fwd_ms": round(fwd_ms, 4), "bwd_ms": round(bwd_ms, 4),
            "fwd_bwd_ratio": round((fwd_ms + bwd_ms) / fwd_ms, 3),
            "step_ms": round(step_ms, 3),
            "step_over_fwd_bwd": round(step_ms / (fwd_ms + bwd_ms), 3),
            "optimizer": "adam, lr 0.2 px (points) and 0.005 (colours), mean squared error",
            "loss_first_last": [round(losses[0], 6), round(losses[-1], 6)],
            "loss_curve": [round(v, 6) for v in losses],
            "loss_fell": bool(losses[-1] < losses[0]),
            "checkpoint_modes": modes}


def batch_leg(kind, steps, warmup, sample=32):
    import numpy as np

    from pathfinder_amd import (ColorF, DestFramebuffer, RectF, Renderer, RendererOptions,
                                Scene, batch, synth)
    from pathfinder_amd.scene import ColorU, PathArrays
    if kind == "glyphs":
        items = synth.glyphs()
        fw = fh = 4096
        budget, align, bg = 6_000_000, 1, ColorF(0, 0, 0, 0)
    else:
        items = synth.canvases()
        fw, fh = 8192, 8192
        budget, align, bg = 14_000_000, 16, ColorF(1, 1, 1, 1)
    t0 = time.perf_counter()
    frames = batch.build_frames(items, fw, fh, budget, align=align)
    pack_s = time.perf_counter() - t0

    def timed(scene, w, h, tune=False):
        """Device time of re-rendering the scene; with tune, the launch
        shape and graph replay come from the autotuner, and the eager
        default is timed beside it."""
        r = Renderer(RendererOptions(DestFramebuffer.full_window(w, h), bg))
        try:
            cmds = scene.build()
            r.begin_scene()
            r.render_commands(cmds)
            r.end_scene()

            def clock():
                for _ in range(max(warmup, 1)):
                    r.rerender()
                r.synchronize()
                t = time.perf_counter()
                for _ in range(steps):
                    r.rerender()
                r.synchronize()
                return (time.perf_counter() - t) / steps * 1e3

            eager = clock()
            choice = None
            ms = eager
            if tune:
                choice = r.autotune(frames=5)
                ms = clock()
            return ms, eager, choice, r.stats()["microline_count"]
        finally:
            r.close()

    frame_ms, eager_ms, choices, ml = [], [], [], 0
    for f in frames:
        ms, eager, choice, m = timed(f.scene, f.width, f.height, tune=True)
        frame_ms.append(round(ms, 3))
        eager_ms.append(round(eager, 3))
        choices.append(choice)
        ml += m
    rng = np.random.default_rng(0)
    per = []
    for i in rng.choice(len(items), min(sample, len(items)), replace=False):
        it = items[int(i)]
        s = Scene()
        s.set_view_box(RectF(0, 0, it.width, it.height))
        a = it.arrays
        paints = np.asarray([s.push_paint(ColorU(*c)) for c in it.colors], np.uint32)
        s.push_draw_paths(PathArrays(a.points, a.flags, a.contour_ends, a.path_contour_ends,
                                     paints, a.fill_rules))
        per.append(timed(s, it.width, it.height)[0])
    total = sum(frame_ms)
    # an item's latency inside the batch: the device time of the frame that
    # renders it (every item of a frame completes with its last kernel)
    lat = np.repeat(frame_ms, [len(f.cells) for f in frames])
    return {"workload": "%s-%d" % (kind, len(items)), "items": len(items),
            "frames": len(frames), "frame_size": [fw, fh], "microlines": ml,
            "frames_ms": round(total, 3), "aggregate_msegments_s": round(ml / total / 1e3, 1),
            "frame_ms": frame_ms, "frame_ms_eager_default": eager_ms,
            "aggregate_msegments_s_eager_default": round(ml / sum(eager_ms) / 1e3, 1),
            "tuned": choices, "pack_s": round(pack_s, 2),
            "per_scene_ms_in_batch_p50": round(float(np.percentile(lat, 50)), 3),
            "per_scene_ms_in_batch_p99": round(float(np.percentile(lat, 99)), 3),
            "per_item_alone_ms_p50": round(float(np.percentile(per, 50)), 3),
            "per_item_alone_ms_p99": round(float(np.percentile(per, 99)), 3),
            "sampled_alone": len(per)}


def run_diff_leg(steps, warmup):
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--diff-leg", "--steps", str(steps),
           "--warmup", str(warmup)]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        print("bench: diff leg failed (%d): %s" % (out.returncode, out.stderr[-2000:]),
              file=sys.stderr)
        return None
    return json.loads(out.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="4k", choices=["4k", "small"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-diff", action="store_true",
                    help="skip the differentiable-render leg")
    ap.add_argument("--diff-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the glyph-atlas and multi-canvas legs")
    ap.add_argument("--shape", default=None, metavar="GRID_CAP,FILL_GRID",
                    help="render with this launch shape and eager launches instead of "
                         "autotuning (counter runs: no graph capture under the profiler)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the frames, run a plain 256 MiB device-to-device copy so a "
                         "counter run can scale FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.diff_leg:
        print(json.dumps(diff_leg(args.steps, args.warmup)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(each GPU would render its own band of framebuffer tiles).", file=sys.stderr)
        sys.exit(2)

    # the differentiable leg first, in a child process, before this one uses the GPU
    diff_result = None if args.no_diff else run_diff_leg(args.steps, args.warmup)

    from pathfinder_amd import DestFramebuffer, Renderer, RendererOptions, synth
    from pathfinder_amd import _lib as L

    if args.workload == "4k":
        scene, size, bg = synth.scene_4k()
        workload = "4k-20k-paths"
    else:
        scene, size, bg = synth.small_scene()
        workload = "1024-256-cubic-paths"
    print("bench: workload %s %dx%d" % (workload, size[0], size[1]), file=sys.stderr, flush=True)

    r = Renderer(RendererOptions(DestFramebuffer.full_window(*size), bg))
    t0 = time.perf_counter()
    commands = scene.build()
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    r.begin_scene()
    r.render_commands(commands)
    r.end_scene()
    t_first = time.perf_counter() - t0
    # the first frame may have grown buffers; this one is the steady state with uploads
    t0 = time.perf_counter()
    r.begin_scene()
    r.render_commands(commands)
    r.end_scene()
    t_upload_frame = time.perf_counter() - t0
    st = r.stats()

    # launch shapes for this workload (pfh_autotune; results do not depend on them)
    if args.shape:
        grid_cap, fill_grid = (int(v) for v in args.shape.split(","))
        r.set_launch_shape(grid_cap, fill_grid)
        shape = {"grid_cap": grid_cap, "fill_grid": fill_grid, "graphs": False}
        print("bench: launch shape %s" % shape, file=sys.stderr, flush=True)
    else:
        shape = r.autotune(frames=5)
        print("bench: autotuned launch shape %s" % shape, file=sys.stderr, flush=True)

    # warm-up with every kernel timed alone (stages serialized on one stream:
    # overlapping kernels would share the time) to find the dominant one
    r.profile(L.PROFILE_SERIAL)
    r.profile_reset()
    for _ in range(max(args.warmup, 1)):
        r.rerender()
    prof = r.profile_read()
    per_frame = {k: v[1] / max(args.warmup, 1) for k, v in prof.items() if v[0]}
    dominant = max(per_frame, key=per_frame.get)

    # timed region: the K frames alone on the streams (no events)
    r.profile(L.PROFILE_OFF)
    r.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.rerender()
    r.synchronize()
    dt = time.perf_counter() - t0
    # the dominant kernel's duration in the frame: the same K frames again with
    # HIP events around that kernel on its stream, after the timed region (the
    # events' marker packets would add to the frame time)
    r.profile(L.PROFILE_ONE, dominant)
    r.profile_reset()
    for _ in range(args.steps):
        r.rerender()
    r.synchronize()
    prof_t = r.profile_read()[dominant]
    r.profile(L.PROFILE_OFF)
    launches = prof_t[0] / args.steps

    ms_per_step = dt / args.steps * 1e3
    M = st["microline_count"]
    fb_tiles = ((size[0] + 15) // 16) * ((size[1] + 15) // 16)
    draw_pts, clip_pts = commands.segments()
    points = len(draw_pts["points"]) + len(clip_pts["points"])
    alg = algorithmic_bytes(st, fb_tiles, size[0] * size[1], points, bg is None)
    k_ms = prof_t[1] / args.steps
    bytes_k = alg.get(dominant)
    achieved = bytes_k / (k_ms * 1e-3) / 1e9 if bytes_k else None
    from pathfinder_amd import profiling
    measured = profiling.read_traffic(TRAFFIC_TABLE.get(args.workload, ""))
    traffic = measured.get("k_" + dominant)

    print("bench: first frame %.1f ms (with buffer growth), frame with uploads %.2f ms, "
          "build %.1f ms" % (t_first * 1e3, t_upload_frame * 1e3, t_build * 1e3),
          file=sys.stderr)
    print("bench: per-kernel ms/frame, each kernel alone (warm-up): " +
          ", ".join("%s %.3f" % (k, v) for k, v in sorted(per_frame.items(),
                                                        key=lambda kv: -kv[1])),
          file=sys.stderr)
    print("bench: stats %s" % json.dumps(st), file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": round(M / (ms_per_step * 1e-3) / 1e6, 3),
        "unit": "Msegments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, pathfinder_amd/synth.py)",
        "config": {"workload": workload, "width": size[0], "height": size[1],
                   "paths": st["path_count"], "segments": st["segment_count"],
                   "microlines": M, "fills": st["fill_count"],
                   "tiles": st["total_tile_count"], "alpha_tiles": st["alpha_tile_count"],
                   "frame_ms_with_uploads": round(t_upload_frame * 1e3, 3),
                   "launch_shape": shape},
        "roofline": {"bound": "hbm", "kernel": dominant,
                     "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic[3] if traffic else None,
                     "scope": "all launches of the kernel in one frame (%g per frame); "
                              "achieved = algorithmic bytes / their summed duration"
                              % launches,
                     "kernel_ms_per_frame": round(k_ms, 4),
                     "kernel_ms_per_frame_alone": round(per_frame[dominant], 4),
                     "kernel_us_per_launch": round(k_ms * 1e3 / max(launches, 1), 2),
                     "algorithmic_bytes_per_frame": bytes_k,
                     "traffic_source": (os.path.relpath(TRAFFIC_TABLE[args.workload], ROOT)
                                        if traffic else None)},
    }
    if diff_result is not None:
        line["diff"] = diff_result
    if not args.no_batch:
        for kind in ("glyphs", "canvases"):
            print("bench: batch leg %s" % kind, file=sys.stderr, flush=True)
            line["batch_" + kind] = batch_leg(kind, min(args.steps, 10), args.warmup)
    if not args.no_cpu_baseline:
        print("bench: cpu baselines", file=sys.stderr, flush=True)
        s_scene, s_size, _ = synth.scene_4k(paths=2000, clips=50, clipped_paths=200)
        from oracle import packing
        import oracle as _oracle
        o = _oracle.Oracle(threads=min(16, os.cpu_count() or 1))
        s_cmds = packing.build(s_scene)
        o.render(s_cmds, s_size[0], s_size[1], bg)
        s_ml = sum(len(o.read(b["batch_id"], _oracle.MICROLINES)) for _, b in s_cmds["batches"])
        o.close()
        line["cpu_baseline"] = cpu_baseline(scene, size, bg, M, (s_scene, s_size, s_ml), "d3d11")
        line["cpu_baseline_d3d9"] = cpu_baseline(scene, size, bg, M, (s_scene, s_size, s_ml),
                                                 "d3d9")
    r.close()
    if args.copy_calibration:
        copy_calibration(256 << 20)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
