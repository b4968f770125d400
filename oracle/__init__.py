"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``liboracle.so`` (built from ``oracle.cpp`` by the Makefile
here), the scalar C++ restatement of the reference's rendering path.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only to check or time against it:
nothing in the product library calls it.

Functions and the reference code they restate are listed in oracle.cpp.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess
from ctypes import c_double, c_float, c_int, c_int64, c_uint32, c_void_p

import numpy as np

ORACLE_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"
MLP_IN = 16  # inputs of the learned sampler's network

_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR), "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.oracle_philox.argtypes = [c_void_p, c_uint32, c_uint32, c_void_p]
        L.oracle_math.argtypes = [c_int, c_void_p, c_void_p, ctypes.c_size_t, c_void_p]
        L.oracle_scene_create.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int]
        L.oracle_scene_create.restype = c_void_p
        L.oracle_scene_free.argtypes = [c_void_p]
        L.oracle_scene_set_textures.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_int64]
        L.oracle_texture_sample.argtypes = [c_void_p, c_uint32, c_float, c_float, c_void_p]
        L.oracle_scene_info.argtypes = [c_void_p, c_void_p]
        L.oracle_scene_nodes.argtypes = [c_void_p, c_void_p]
        L.oracle_scene_order.argtypes = [c_void_p, c_void_p, c_void_p]
        L.oracle_generate_rays.argtypes = [c_void_p, c_int64, c_void_p, c_int, c_void_p, c_int64, c_uint32,
                                           c_uint32, c_uint32, c_void_p]
        L.oracle_trace.argtypes = [c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p, c_int]
        L.oracle_render.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_int, c_int, c_void_p, c_int64,
                                    c_uint32, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_int]
        L.oracle_make_color.argtypes = [c_void_p, c_int64, c_void_p]
        L.oracle_spherical_ids.argtypes = [c_void_p, c_int64, c_int, c_int, c_int, c_void_p]
        L.oracle_camera_rays.argtypes = [c_int, c_void_p, c_void_p, c_float, c_int, c_int, c_void_p]
        L.oracle_radiance.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_int]
        L.oracle_pose.argtypes = [c_void_p, c_void_p, c_void_p]
        _lib = L
    return _lib




def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def default_threads() -> int:
    return max(1, len(os.sched_getaffinity(0)))


def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox(_p(c), int(key[0]), int(key[1]), _p(out))
    return out


MATH_KINDS = {"sin": 0, "cos": 1, "log": 2, "exp2": 3, "pow": 4, "asin": 5, "atan2": 6}


def math_fn(kind: str, x, y=None) -> np.ndarray:
    xa = np.ascontiguousarray(x, dtype=np.float32)
    ya = np.ascontiguousarray(y if y is not None else np.zeros_like(xa), dtype=np.float32)
    out = np.zeros_like(xa)
    lib().oracle_math(MATH_KINDS[kind], _p(xa), _p(ya), xa.size, _p(out))
    return out


def pose12(pose6) -> np.ndarray:
    p = np.ascontiguousarray(np.asarray(pose6, dtype=np.float32)[:3])
    r = np.ascontiguousarray(np.asarray(pose6, dtype=np.float32)[3:6])
    out = np.zeros(12, dtype=np.float32)
    lib().oracle_pose(_p(p), _p(r), _p(out))
    return out


def make_color(rgb) -> np.ndarray:
    a = np.ascontiguousarray(rgb, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((a.shape[0], 4), dtype=np.uint8)
    lib().oracle_make_color(_p(a), a.shape[0], _p(out))
    return out


class OracleScene:
    """A scene plus its CPU-built LBVH (same rules as the GPU builder)."""

    def __init__(self, tri_pos, tri_col, tri_mode, sky_mode: int = 1, leaf_size: int = 4, textures=None):
        """textures: optional (descriptors (n, 6) int32, texels uint32 RGBA8)."""
        self.pos = np.ascontiguousarray(tri_pos, dtype=np.float32).reshape(-1, 9)
        self.col = np.ascontiguousarray(tri_col, dtype=np.float32).reshape(-1, 9)
        self.mode = np.ascontiguousarray(tri_mode, dtype=np.uint32).reshape(-1)
        self.n = self.pos.shape[0]
        self.h = lib().oracle_scene_create(_p(self.pos), _p(self.col), _p(self.mode), self.n, sky_mode, leaf_size)
        if textures is not None and len(textures[0]):
            self.tex_desc = np.ascontiguousarray(textures[0], dtype=np.int32).reshape(-1, 6)
            self.texels = np.ascontiguousarray(textures[1], dtype=np.uint32).reshape(-1)
            lib().oracle_scene_set_textures(self.h, _p(self.tex_desc), self.tex_desc.shape[0], _p(self.texels),
                                            self.texels.size)

    @classmethod
    def from_scene(cls, scene, leaf_size: int = 4):
        return cls(scene.tri_pos(), scene.tri_col(), scene.tri_mode(), scene.sky_mode(), leaf_size,
                   textures=scene.texture_arrays())

    def radiance(self, rays, threads: int | None = None) -> np.ndarray:
        """(n, 3) shaded colour (or sky) of each ray's closest hit."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        out = np.zeros((r.shape[0], 3), np.float32)
        lib().oracle_radiance(self.h, _p(r), r.shape[0], _p(out), threads or default_threads())
        return out

    def texture_sample(self, tex: int, u: float, v: float) -> np.ndarray:
        """One lookup of texture `tex` at (u, v): linear RGB before linearize."""
        out = np.zeros(3, np.float32)
        lib().oracle_texture_sample(self.h, tex, u, v, _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_scene_free(self.h)
            self.h = None

    def info(self) -> dict:
        info = np.zeros(12, dtype=np.int64)
        lib().oracle_scene_info(self.h, _p(info))
        return {"tris": int(info[0]), "nodes": int(info[1]), "stack_bound": int(info[2]), "levels": int(info[3]),
                "leaf_size": int(info[4]), "node_bytes": int(info[11]),
                "centroid_lo": info[5:8].astype(np.uint32).view(np.float32),
                "centroid_hi": info[8:11].astype(np.uint32).view(np.float32)}

    def nodes(self) -> np.ndarray:
        info = self.info()
        out = np.zeros(info["nodes"] * info["node_bytes"], dtype=np.uint8)
        lib().oracle_scene_nodes(self.h, _p(out))
        return out

    def order(self):
        ids = np.zeros(self.n, dtype=np.uint32)
        codes = np.zeros(self.n, dtype=np.uint32)
        lib().oracle_scene_order(self.h, _p(ids), _p(codes))
        return ids, codes

    def trace(self, rays, bvh: bool = True, threads: int | None = None, visits: bool = False):
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        ids = np.zeros(r.shape[0], dtype=np.int32)
        tbb = np.zeros((r.shape[0], 3), dtype=np.float32)
        vis = np.zeros((r.shape[0], 2), dtype=np.uint64) if visits else None
        lib().oracle_trace(self.h, _p(r), r.shape[0], 1 if bvh else 0, _p(ids), _p(tbb), _p(vis),
                           threads or default_threads())
        return (ids, tbb, vis) if visits else (ids, tbb)

    def render(self, omm, poses12, spp: int = 32, counts=None, seed: int = 42, eye: int = 0, frame: int = 0,
               want_samples: bool = False, want_f64: bool = False, threads: int | None = None,
               sample_base: int = 0) -> dict:
        """counts: per ommatidium (shared by the poses) or, with shape
        (poses, ommatidia), per unit."""
        o = np.ascontiguousarray(omm, dtype=np.float32).reshape(-1, 8)
        P = np.ascontiguousarray(poses12, dtype=np.float32).reshape(-1, 12)
        c = None if counts is None else np.ascontiguousarray(counts, dtype=np.uint32)
        n_omm, n_pose = o.shape[0], P.shape[0]
        per_unit = c is not None and c.ndim == 2
        if per_unit:
            assert c.shape == (n_pose, n_omm)
            total = int(c.sum())
        else:
            total = n_pose * (int(c.sum()) if c is not None else n_omm * spp)
        samples = np.zeros((total, 4), dtype=np.float32) if want_samples else None
        gpu = np.zeros((n_pose, n_omm, 3), dtype=np.float32)
        seq = np.zeros((n_pose, n_omm, 3), dtype=np.float32)
        f64 = np.zeros((n_pose, n_omm, 3), dtype=np.float64) if want_f64 else None
        visits = np.zeros(2, dtype=np.uint64)
        lib().oracle_render(self.h, _p(o), n_omm, _p(c), 1 if per_unit else 0, spp, _p(P), n_pose, seed, eye,
                            frame, sample_base, _p(samples), _p(gpu), _p(seq), _p(f64), _p(visits),
                            threads or default_threads())
        return {"samples": samples, "colour_gpu_order": gpu, "colour_seq": seq, "colour_f64": f64,
                "nodes_visited": int(visits[0]), "tris_tested": int(visits[1])}


def generate_rays(omm, poses12, spp: int = 32, counts=None, seed: int = 42, eye: int = 0, frame: int = 0):
    o = np.ascontiguousarray(omm, dtype=np.float32).reshape(-1, 8)
    P = np.ascontiguousarray(poses12, dtype=np.float32).reshape(-1, 12)
    c = None if counts is None else np.ascontiguousarray(counts, dtype=np.uint32)
    per_pose = int(c.sum()) if c is not None else o.shape[0] * spp
    out = np.zeros((P.shape[0] * per_pose, 8), dtype=np.float32)
    lib().oracle_generate_rays(_p(o), o.shape[0], _p(c), spp, _p(P), P.shape[0], seed, eye, frame, _p(out))
    return out


def mlp_loss_grad(params, X, T=None, mode: int = 0):
    """FP64 MLP (mode 0 exact, mode 1 with the GPU's BF16 rounding points).
    Returns (y, grad or None, loss)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MLP_IN)
    n = x.shape[0]
    t = None if T is None else np.ascontiguousarray(T, dtype=np.float32)
    y = np.zeros(n, dtype=np.float64)
    grad = np.zeros(p.size, dtype=np.float64) if t is not None else None
    loss = np.zeros(1, dtype=np.float64)
    L = lib()
    L.oracle_mlp_loss_grad.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p]
    L.oracle_mlp_loss_grad(_p(p), _p(x), _p(t), n, mode, _p(y), _p(grad), _p(loss))
    return y, grad, float(loss[0])


def spherical_ids(omm, width: int, height: int, variant: int) -> np.ndarray:
    """(height, width) ommatidium index per pixel of a spherical projection
    (0 positionwise, 1 orientationwise, 2 split orientationwise)."""
    o = np.ascontiguousarray(omm, dtype=np.float32).reshape(-1, 8)
    out = np.zeros((height, width), np.uint32)
    lib().oracle_spherical_ids(_p(o), o.shape[0], width, height, variant, _p(out))
    return out


def camera_rays(kind: int, pose12, scale, width: int, height: int, start_radius: float = 0.0) -> np.ndarray:
    """(h * w, 8) primary rays of a regular camera (0 pinhole, 1 orthographic,
    2 panoramic)."""
    P = np.ascontiguousarray(pose12, dtype=np.float32).reshape(12)
    sc = np.ascontiguousarray(scale, dtype=np.float32).reshape(3)
    out = np.zeros((height * width, 8), np.float32)
    lib().oracle_camera_rays(kind, _p(P), _p(sc), start_radius, width, height, _p(out))
    return out
