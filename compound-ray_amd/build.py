"""Build recipe for libEyeRenderer3.so (gfx950).

Every translation unit is compiled by hipcc for gfx950 only, with FMA
contraction disabled and correctly rounded FP32 divide/sqrt, so that the
device arithmetic matches IEEE single precision operation by operation.
Objects are compiled in parallel and linked into one shared library in
``compound-ray_amd/lib``.
"""
from __future__ import annotations

import concurrent.futures
import os
import pathlib
import subprocess
import sys

PKG_DIR = pathlib.Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_DIR = PKG_DIR / "lib"
OBJ_DIR = PKG_DIR / "_build"
LIB_NAME = "libEyeRenderer3.so"

SOURCES = [
    "json.cpp",
    "jpeg.cpp",
    "png.cpp",
    "scene_loader.cpp",
    "scan.hip",
    "lbvh.hip",
    "render.hip",
    "projection.hip",
    "batch.cpp",
    "mlp.hip",
    "adaptive.hip",
    "knn.cpp",
    "noalloc.hip",
    "eye_renderer.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-ffp-contract=off",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
]
DEVICE_FLAGS = [
    "--offload-arch=gfx950",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    # keep the compiler from forming packed-FP32 instructions on its own
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: pathlib.Path, obj: pathlib.Path, verbose: bool) -> None:
    cmd = [HIPCC, *COMMON_FLAGS, *DEVICE_FLAGS, "-I", str(CSRC), "-I", str(PKG_DIR.parent / "include")]
    # host-only translation units also go through hipcc so they share the HIP headers
    cmd += ["-x", "hip"]
    cmd += ["-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)


def build(verbose: bool = False, jobs: int | None = None) -> pathlib.Path:
    """Compiles every source and links lib/libEyeRenderer3.so; returns its path."""
    OBJ_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    objs = []
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for name in SOURCES:
            src = CSRC / name
            obj = OBJ_DIR / (src.stem + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, src, obj, verbose))
        for f in futs:
            f.result()
    out = LIB_DIR / LIB_NAME
    tmp = LIB_DIR / (LIB_NAME + ".tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
