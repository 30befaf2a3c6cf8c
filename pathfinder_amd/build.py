"""Builds pathfinder_amd/libpfhip.so for gfx950 with hipcc (no cmake).

    python -m pathfinder_amd.build        # or: from pathfinder_amd.build import build

All HIP code is compiled with -ffp-contract=off (the kernels are held bit for
bit to binary32 formulas evaluated without fused multiply-add) and with packed
FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpfhip.so")
LUT = os.path.join(HERE, "data", "area_lut.rgba")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = COMMON + [
    "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    if verbose and r.stdout.strip():
        print(r.stdout)


def build(verbose=False):
    """Compiles every source and links libpfhip.so. Always recompiles."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []

    def obj(name):
        p = os.path.join(BUILD, name + ".o")
        objs.append(p)
        return p

    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "pfh_kernels.hip"),
          "-o", obj("pfh_kernels")], verbose)
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "pfh_renderer.cpp"),
          "-o", obj("pfh_renderer")], verbose)
    cxx = shutil.which("g++") or "g++"
    _run([cxx, *COMMON, "-ffp-contract=off", "-c", os.path.join(CSRC, "pfh_scene.cpp"),
          "-o", obj("pfh_scene")], verbose)
    _run([cxx, *COMMON, "-ffp-contract=off", "-c", os.path.join(CSRC, "pfh_palette.cpp"),
          "-o", obj("pfh_palette")], verbose)
    _run([cxx, *COMMON, "-ffp-contract=off", "-pthread", "-c", os.path.join(CSRC, "pfh_d3d9.cpp"),
          "-o", obj("pfh_d3d9")], verbose)
    _run([cxx, *COMMON, "-c", os.path.join(CSRC, "pfh_svg.cpp"), "-o", obj("pfh_svg")],
         verbose)
    _run([cxx, *COMMON, "-c", os.path.join(CSRC, "pfh_capi.cpp"), "-o", obj("pfh_capi")],
         verbose)
    _run([cxx, "-fPIC", "-c", os.path.join(CSRC, "pfh_lut.cpp"),
          '-DPFH_AREA_LUT_PATH="%s"' % LUT, "-o", obj("pfh_lut")], verbose)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-lz", "-Wl,--no-undefined"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
