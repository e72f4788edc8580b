"""Builds libforkergl.so (HIP kernels for gfx950 + host C++) in-tree.

Every HIP source is compiled with -ffp-contract=off (the reference's float
expressions are separately rounded operations) and with packed FP32 code
generation disabled.  Objects go to forkerrenderer_amd/_build, the shared
library to forkerrenderer_amd/_lib.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libforkergl.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["raster.hip", "sort.hip", "shade.hip", "rng.hip", "nao.hip", "batch.hip", "context.cpp", "host_io.cpp", "batch.cpp"]
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-command-line-argument",
]


# host-only sources (no device code), built by the host compiler
HOST_SOURCES = ["mt_jump.cpp"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src in HOST_SOURCES:
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-I", CSRC, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    sources = [s for s in SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        results = list(ex.map(_compile, sources))
    failed = [r for r in results if r[2] != 0]
    for src, _, rc, out in results:
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if verbose or rc != 0:
            if lines:
                print(f"--- {src} ---\n" + "\n".join(lines), file=sys.stderr)
    if failed:
        raise RuntimeError("HIP compile failed: " + ", ".join(r[0] for r in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [r[1] for r in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    # command-line renderer (the reference's ForkerRenderer <scene>)
    exe = os.path.join(LIBDIR, "ForkerRendererAMD")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "main.cpp"),
           "-L", LIBDIR, "-lforkergl", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("CLI link failed")
    os.replace(exe + ".tmp", exe)
    # the frame driver written against the C++ facade (include/forkergl_amd.hpp)
    exe = os.path.join(LIBDIR, "facade_render")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "facade_render.cpp"),
           "-L", LIBDIR, "-lforkergl", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("facade program link failed")
    os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
