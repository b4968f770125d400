"""Seeded synthetic scenes and eyes of the benchmark configurations.

Every scene is emitted as glTF 2.0 (JSON + one external .bin buffer) with
per-vertex VEC4 float colours and a ``simple_sky`` background, and every eye
as a .eye file referenced by a compound-eye camera, so that the library is
always driven through its real loading boundary (loadGlTFscene).  The same
arrays are returned to the caller for the CPU oracle.

Configurations (all distances in metres, angles in radians):
  C1  100x100-vertex value-noise heightfield (19,602 triangles) over 10 m,
      amplitude 0.5 m; one 1,000-ommatidium Fibonacci-sphere eye of radius
      1 mm, acceptance 2 deg, 32 samples.
  C2  1024x1024-vertex fBm heightfield over 40 m (2,093,058 triangles) plus
      2,000 16-segment cylinders (48 triangles each); two 5,500-ommatidium
      ellipsoidal-patch eyes spanning 160 x 120 deg, acceptance U[1.5, 3] deg,
      128 samples.
  C3  4,096 poses of a 5,000-ommatidium eye over the C1 scene, per-ommatidium
      samples log-uniform in [8, 256].
"""
from __future__ import annotations

import json
import math
import os
import pathlib
from dataclasses import dataclass, field

import numpy as np

DEG = math.pi / 180.0


GL_REPEAT, GL_CLAMP_TO_EDGE, GL_MIRRORED_REPEAT = 10497, 33071, 33648
GL_NEAREST, GL_LINEAR = 9728, 9729
_ADDRESS = {GL_REPEAT: 0, GL_CLAMP_TO_EDGE: 1, GL_MIRRORED_REPEAT: 2}


@dataclass
class Texture:
    """RGBA8 image (row 0 = top) with a glTF sampler."""
    rgba: np.ndarray                 # (H, W, 4) uint8
    wrap_s: int = GL_REPEAT
    wrap_t: int = GL_REPEAT
    min_filter: int | None = GL_LINEAR   # None: texture written without a sampler
    embed: bool = True               # PNG in the .bin buffer, else a file next to the .gltf

    def desc(self) -> tuple:
        """(wrap s, wrap t, filter) as the renderer encodes them."""
        if self.min_filter is None:
            return 0, 0, 1
        return _ADDRESS[self.wrap_s], _ADDRESS[self.wrap_t], 0 if self.min_filter == GL_NEAREST else 1


@dataclass
class Mesh:
    name: str
    pos: np.ndarray          # (V, 3) float32
    idx: np.ndarray          # (T, 3) uint32
    col: np.ndarray | None   # (V, 4) float32, or None (material colour / texture)
    base_colour: tuple = (1.0, 1.0, 1.0, 1.0)
    texture: Texture | None = None   # base-colour texture (used when col is None)
    uv: np.ndarray | None = None     # (V, 2) float32 TEXCOORD_0; None: barycentric UVs

    def textured(self) -> bool:
        return self.col is None and self.texture is not None

    def tri_pos(self) -> np.ndarray:
        return self.pos[self.idx].reshape(-1, 9).astype(np.float32)

    def tri_col(self) -> np.ndarray:
        if self.textured():
            t = self.idx.shape[0]
            out = np.zeros((t, 9), np.float32)
            if self.uv is None:
                out[:, 0:6] = np.array([0, 0, 1, 0, 0, 1], np.float32)
            else:
                out[:, 0:6] = self.uv.astype(np.float32)[self.idx].reshape(-1, 6)
            return out
        if self.col is None:
            c = np.asarray(self.base_colour[:3], dtype=np.float32)
            return np.tile(np.concatenate([c, c, c]), (self.idx.shape[0], 1))
        return self.col[self.idx][:, :, :3].reshape(-1, 9).astype(np.float32)

    def tri_mode(self, texture_index: int = 0) -> np.ndarray:
        if self.textured():
            return np.full(self.idx.shape[0], 2 | (texture_index << 2), dtype=np.uint32)
        return np.full(self.idx.shape[0], 0 if self.col is not None else 1, dtype=np.uint32)


def encode_png(rgba: np.ndarray, filters=(0, 1, 2, 3, 4)) -> bytes:
    """RGBA8 PNG; row filter types cycle through `filters`."""
    import struct
    import zlib
    img = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = img.shape[:2]
    rows = img.reshape(h, w * 4).astype(np.int32)
    raw = bytearray()
    for y in range(h):
        ft = filters[y % len(filters)]
        cur = rows[y]
        prev = rows[y - 1] if y > 0 else np.zeros_like(cur)
        left = np.concatenate([np.zeros(4, np.int32), cur[:-4]])
        upleft = np.concatenate([np.zeros(4, np.int32), prev[:-4]])
        if ft == 0:
            pred = np.zeros_like(cur)
        elif ft == 1:
            pred = left
        elif ft == 2:
            pred = prev
        elif ft == 3:
            pred = (left + prev) >> 1
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
        raw.append(ft)
        raw.extend(((cur - pred) & 255).astype(np.uint8).tobytes())

    def chunk(kind: bytes, body: bytes) -> bytes:
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xffffffff)

    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)
    return b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(raw), 6)) + \
        chunk(b"IEND", b"")


@dataclass
class Eye:
    name: str
    omm: np.ndarray          # (N, 8) float32: position, axis, acceptance, focal offset
    projection: str = "single_dimension_fast"
    translation: tuple = (0.0, 0.0, 0.0)
    yaw: float = 0.0


@dataclass
class Camera:
    """Regular glTF camera: "perspective" (yfov radians), "orthographic"
    (xmag, ymag) or "panoramic" (a perspective camera with the panoramic extra)."""
    name: str
    kind: str = "perspective"
    yfov: float = 1.0
    xmag: float = 1.0
    ymag: float = 1.0
    translation: tuple = (0.0, 0.0, 0.0)
    yaw: float = 0.0


@dataclass
class Scene:
    meshes: list
    eyes: list = field(default_factory=list)
    cameras: list = field(default_factory=list)   # regular cameras, listed after the eyes
    background: str = "simple_sky"
    heightfield: tuple | None = None   # (xs, zs, heights) for ground queries

    def tri_pos(self) -> np.ndarray:
        return np.concatenate([m.tri_pos() for m in self.meshes]) if self.meshes else np.zeros((0, 9), np.float32)

    def tri_col(self) -> np.ndarray:
        return np.concatenate([m.tri_col() for m in self.meshes]) if self.meshes else np.zeros((0, 9), np.float32)

    def textures(self) -> list:
        """Distinct textures of textured meshes, in first-use order."""
        out = []
        for m in self.meshes:
            if m.textured() and not any(t is m.texture for t in out):
                out.append(m.texture)
        return out

    def tri_mode(self) -> np.ndarray:
        if not self.meshes:
            return np.zeros(0, np.uint32)
        texs = self.textures()
        parts = []
        for m in self.meshes:
            ti = next((k for k, t in enumerate(texs) if t is m.texture), 0) if m.textured() else 0
            parts.append(m.tri_mode(ti))
        return np.concatenate(parts)

    def texture_arrays(self):
        """(descriptors (n, 6) int32: width, height, wrap s, wrap t, filter,
        first texel; texels uint32 RGBA8 with red in the low byte)."""
        desc, texels, first = [], [], 0
        for t in self.textures():
            h, w = t.rgba.shape[:2]
            desc.append((w, h) + t.desc() + (first,))
            px = t.rgba.astype(np.uint32).reshape(-1, 4)
            texels.append(px[:, 0] | px[:, 1] << 8 | px[:, 2] << 16 | px[:, 3] << 24)
            first += w * h
        if not desc:
            return np.zeros((0, 6), np.int32), np.zeros(0, np.uint32)
        return np.array(desc, np.int32), np.concatenate(texels).astype(np.uint32)

    def tri_count(self) -> int:
        return int(sum(m.idx.shape[0] for m in self.meshes))

    def sky_mode(self) -> int:
        return 1 if self.background == "simple_sky" else 0

    def ground_height(self, x: float, z: float) -> float:
        xs, zs, h = self.heightfield
        fx = np.interp(x, xs, np.arange(len(xs)))
        fz = np.interp(z, zs, np.arange(len(zs)))
        i0, j0 = int(min(math.floor(fz), len(zs) - 2)), int(min(math.floor(fx), len(xs) - 2))
        tz, tx = fz - i0, fx - j0
        return float((1 - tz) * ((1 - tx) * h[i0, j0] + tx * h[i0, j0 + 1]) +
                     tz * ((1 - tx) * h[i0 + 1, j0] + tx * h[i0 + 1, j0 + 1]))


# ------------------------------------------------------------------ noise ----

def value_noise(x: np.ndarray, z: np.ndarray, rng: np.random.Generator, octaves: int, base_freq: float,
                gain: float = 0.5) -> np.ndarray:
    """Sum of `octaves` smooth value-noise layers in [-1, 1] (normalised)."""
    total = np.zeros(np.broadcast(x, z).shape, dtype=np.float64)
    amp, norm, freq = 1.0, 0.0, base_freq
    for _ in range(octaves):
        u, v = x * freq, z * freq
        u0, v0 = np.floor(u), np.floor(v)
        iu, iv = u0.astype(np.int64), v0.astype(np.int64)
        iu -= iu.min()
        iv -= iv.min()
        lattice = rng.uniform(-1.0, 1.0, size=(iv.max() + 2, iu.max() + 2))
        fu, fv = u - u0, v - v0
        su, sv = fu * fu * (3 - 2 * fu), fv * fv * (3 - 2 * fv)
        a = lattice[iv, iu]
        b = lattice[iv, iu + 1]
        c = lattice[iv + 1, iu]
        d = lattice[iv + 1, iu + 1]
        total += amp * ((1 - sv) * ((1 - su) * a + su * b) + sv * ((1 - su) * c + su * d))
        norm += amp
        amp *= gain
        freq *= 2.0
    return total / norm


def heightfield(n: int, extent: float, amplitude: float, octaves: int, base_freq: float,
                rng: np.random.Generator, name: str = "terrain", colour: str = "uniform"):
    """n x n vertex grid centred on the origin, 2 (n-1)^2 triangles."""
    xs = np.linspace(-extent / 2, extent / 2, n)
    zs = np.linspace(-extent / 2, extent / 2, n)
    X, Z = np.meshgrid(xs, zs)
    H = amplitude * value_noise(X, Z, rng, octaves, base_freq)
    pos = np.stack([X, H, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (i * n + j).ravel()
    v01 = v00 + 1
    v10 = v00 + n
    v11 = v10 + 1
    idx = np.empty((2 * v00.size, 3), dtype=np.uint32)
    idx[0::2] = np.stack([v00, v10, v11], axis=1)
    idx[1::2] = np.stack([v00, v11, v01], axis=1)
    if colour == "uniform":
        rgb = rng.uniform(0.0, 1.0, size=(pos.shape[0], 3))
    else:  # earthy colours with fBm variation
        t = 0.5 + 0.5 * value_noise(X, Z, rng, 4, 0.5).reshape(-1)
        base = np.array([0.35, 0.30, 0.18]) * (1 - t[:, None]) + np.array([0.25, 0.45, 0.15]) * t[:, None]
        rgb = np.clip(base + rng.normal(0.0, 0.06, size=base.shape), 0.0, 1.0)
    col = np.concatenate([rgb, np.ones((pos.shape[0], 1))], axis=1).astype(np.float32)
    return Mesh(name, pos, idx, col), (xs, zs, H)


def cylinders(count: int, scene_hf, rng: np.random.Generator, extent: float, segments: int = 16) -> Mesh:
    """`count` upright capped cylinders (side + top cap: 3*segments triangles each)."""
    xs, zs, H = scene_hf
    verts, idxs, cols = [], [], []
    base_v = 0
    ang = np.arange(segments) * (2 * math.pi / segments)
    ring = np.stack([np.cos(ang), np.zeros(segments), np.sin(ang)], axis=1)
    for k in range(count):
        cx, cz = rng.uniform(-extent / 2 * 0.9, extent / 2 * 0.9, size=2)
        r = rng.uniform(0.02, 0.1)
        h = rng.uniform(0.2, 2.0)
        gi = int(np.clip(np.searchsorted(zs, cz), 0, len(zs) - 1))
        gj = int(np.clip(np.searchsorted(xs, cx), 0, len(xs) - 1))
        y0 = H[gi, gj] - 0.05
        bottom = ring * r + np.array([cx, y0, cz])
        top = ring * r + np.array([cx, y0 + h, cz])
        centre = np.array([[cx, y0 + h, cz]])
        v = np.concatenate([bottom, top, centre])
        s = np.arange(segments)
        s1 = (s + 1) % segments
        side_a = np.stack([s, s1, segments + s1], axis=1)
        side_b = np.stack([s, segments + s1, segments + s], axis=1)
        cap = np.stack([segments + s, segments + s1, np.full(segments, 2 * segments)], axis=1)
        tri = np.concatenate([side_a, side_b, cap]) + base_v
        green = np.array([0.15, 0.45, 0.12]) + rng.normal(0, 0.05, size=3)
        c = np.clip(green + rng.normal(0.0, 0.08, size=(v.shape[0], 3)), 0.0, 1.0)
        verts.append(v)
        idxs.append(tri)
        cols.append(np.concatenate([c, np.ones((v.shape[0], 1))], axis=1))
        base_v += v.shape[0]
    return Mesh("vegetation", np.concatenate(verts).astype(np.float32), np.concatenate(idxs).astype(np.uint32),
                np.concatenate(cols).astype(np.float32))


# ------------------------------------------------------------------- eyes ----

def icosphere(subdivisions: int):
    """Unit icosphere: (vertices (V, 3) float64, faces (F, 3) int64), F = 20 * 4^s."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    faces = np.array(f, np.int64)
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = np.array(nf, np.int64)
    return np.array(verts), faces


def rocks(count: int, scene_hf, rng: np.random.Generator, extent: float, subdivisions: int = 5) -> Mesh:
    """`count` noisy icosphere rocks (radius 0.1-0.6 m) half sunk into the terrain."""
    xs, zs, H = scene_hf
    sv, sf = icosphere(subdivisions)
    verts, idxs, cols = [], [], []
    for k in range(count):
        cx, cz = rng.uniform(-extent / 2 * 0.9, extent / 2 * 0.9, size=2)
        r = rng.uniform(0.1, 0.6)
        gi = int(np.clip(np.searchsorted(zs, cz), 0, len(zs) - 1))
        gj = int(np.clip(np.searchsorted(xs, cx), 0, len(xs) - 1))
        bump = 1.0 + 0.25 * value_noise(sv[:, 0] * 3 + k, sv[:, 2] * 3 + sv[:, 1], rng, 3, 1.0)
        v = sv * (r * bump)[:, None] * np.array([1.0, rng.uniform(0.5, 0.9), 1.0]) + np.array([cx, H[gi, gj], cz])
        grey = rng.uniform(0.3, 0.6)
        c = np.clip(grey + rng.normal(0.0, 0.05, size=(v.shape[0], 3)), 0.0, 1.0)
        verts.append(v)
        idxs.append(sf + sum(len(x) for x in verts[:-1]))
        cols.append(np.concatenate([c, np.ones((v.shape[0], 1))], axis=1))
    return Mesh("rocks", np.concatenate(verts).astype(np.float32), np.concatenate(idxs).astype(np.uint32),
                np.concatenate(cols).astype(np.float32))


def world_ommatidia(omm: np.ndarray, pose12: np.ndarray) -> np.ndarray:
    """Ommatidia moved into world space by a pose (position, x, y, z axes): an
    eye rendered at the identity pose with these ommatidia sees what the
    original eye sees at `pose` (up to rounding)."""
    P = np.asarray(pose12, np.float64)
    R = np.stack([P[3:6], P[6:9], P[9:12]], axis=1)
    out = omm.astype(np.float64).copy()
    out[:, 0:3] = omm[:, 0:3].astype(np.float64) @ R.T + P[0:3]
    out[:, 3:6] = omm[:, 3:6].astype(np.float64) @ R.T
    return out.astype(np.float32)


def fibonacci_eye(n: int, radius: float, acceptance: float) -> np.ndarray:
    """n ommatidia on a Fibonacci sphere, axes radial, focal offset 0."""
    k = np.arange(n, dtype=np.float64)
    y = 1.0 - 2.0 * (k + 0.5) / n
    rad = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    phi = k * (math.pi * (3.0 - math.sqrt(5.0)))
    d = np.stack([np.cos(phi) * rad, y, np.sin(phi) * rad], axis=1)
    out = np.zeros((n, 8), dtype=np.float32)
    out[:, 0:3] = (d * radius).astype(np.float32)
    out[:, 3:6] = d.astype(np.float32)
    out[:, 6] = np.float32(acceptance)
    out[:, 7] = 0.0
    return out


def ellipsoid_patch_eye(n: int, semi_axes, azimuth_span: float, elevation_span: float, centre_azimuth: float,
                        acceptance_range, rng: np.random.Generator, offset=(0.0, 0.0, 0.0)) -> np.ndarray:
    """n ommatidia on an ellipsoidal patch facing local +z (the eye's forward
    axis), uniform in solid angle over azimuth x elevation; axes are the
    surface normals; acceptance angles uniform in acceptance_range."""
    a, b, c = semi_axes
    k = np.arange(n, dtype=np.float64)
    se = math.sin(elevation_span / 2)
    sin_el = -se + 2 * se * (k + 0.5) / n
    el = np.arcsin(sin_el)
    frac = np.mod(k * 0.6180339887498949, 1.0)
    az = centre_azimuth + (frac - 0.5) * azimuth_span
    u = np.stack([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)], axis=1)
    # point of the ellipsoid along u and its outward normal
    s = 1.0 / np.sqrt((u[:, 0] / a) ** 2 + (u[:, 1] / b) ** 2 + (u[:, 2] / c) ** 2)
    p = u * s[:, None]
    nrm = np.stack([p[:, 0] / a ** 2, p[:, 1] / b ** 2, p[:, 2] / c ** 2], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    out = np.zeros((n, 8), dtype=np.float32)
    out[:, 0:3] = (p + np.asarray(offset)).astype(np.float32)
    out[:, 3:6] = nrm.astype(np.float32)
    out[:, 6] = rng.uniform(acceptance_range[0], acceptance_range[1], size=n).astype(np.float32)
    out[:, 7] = 0.0
    return out


def write_eye_file(path, omm: np.ndarray) -> None:
    with open(path, "w") as f:
        for row in np.asarray(omm, dtype=np.float32):
            f.write(" ".join(repr(float(v)) for v in row) + "\n")


# ------------------------------------------------------------------- glTF ----

def write_gltf(directory, name: str, scene: Scene) -> pathlib.Path:
    """Writes name.gltf + name.bin (+ one .eye per eye) into `directory`."""
    d = pathlib.Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    blob = bytearray()
    views, accessors, meshes, materials, nodes, cameras = [], [], [], [], [], []

    def add_view(arr: np.ndarray, target: int) -> int:
        while len(blob) % 4:
            blob.append(0)
        off = len(blob)
        data = np.ascontiguousarray(arr).tobytes()
        blob.extend(data)
        views.append({"buffer": 0, "byteOffset": off, "byteLength": len(data), "target": target})
        return len(views) - 1

    images, samplers, textures = [], [], []
    tex_index = {}

    def add_texture(t: Texture) -> int:
        if id(t) in tex_index:
            return tex_index[id(t)]
        png = encode_png(t.rgba)
        if t.embed:
            while len(blob) % 4:
                blob.append(0)
            off = len(blob)
            blob.extend(png)
            views.append({"buffer": 0, "byteOffset": off, "byteLength": len(png)})
            images.append({"bufferView": len(views) - 1, "mimeType": "image/png"})
        else:
            fname = f"{name}-tex{len(images)}.png"
            (d / fname).write_bytes(png)
            images.append({"uri": fname})
        tex = {"source": len(images) - 1}
        if t.min_filter is not None:
            samplers.append({"wrapS": t.wrap_s, "wrapT": t.wrap_t, "minFilter": t.min_filter,
                             "magFilter": GL_LINEAR})
            tex["sampler"] = len(samplers) - 1
        textures.append(tex)
        tex_index[id(t)] = len(textures) - 1
        return tex_index[id(t)]

    for m in scene.meshes:
        pv = add_view(m.pos.astype(np.float32), 34962)
        accessors.append({"bufferView": pv, "componentType": 5126, "count": int(m.pos.shape[0]), "type": "VEC3",
                          "min": [float(v) for v in m.pos.min(axis=0)], "max": [float(v) for v in m.pos.max(axis=0)]})
        attrs = {"POSITION": len(accessors) - 1}
        if m.col is not None:
            cv = add_view(m.col.astype(np.float32), 34962)
            accessors.append({"bufferView": cv, "componentType": 5126, "count": int(m.col.shape[0]), "type": "VEC4"})
            attrs["COLOR_0"] = len(accessors) - 1
        if m.uv is not None:
            uvv = add_view(m.uv.astype(np.float32), 34962)
            accessors.append({"bufferView": uvv, "componentType": 5126, "count": int(m.uv.shape[0]), "type": "VEC2"})
            attrs["TEXCOORD_0"] = len(accessors) - 1
        iv = add_view(m.idx.astype(np.uint32).reshape(-1), 34963)
        accessors.append({"bufferView": iv, "componentType": 5125, "count": int(m.idx.size), "type": "SCALAR"})
        pbr = {"baseColorFactor": [float(v) for v in m.base_colour]}
        if m.texture is not None:
            pbr["baseColorTexture"] = {"index": add_texture(m.texture)}
        materials.append({"name": m.name + "-material", "pbrMetallicRoughness": pbr})
        meshes.append({"name": m.name, "primitives": [{"attributes": attrs, "indices": len(accessors) - 1,
                                                       "material": len(materials) - 1, "mode": 4}]})
        nodes.append({"name": m.name, "mesh": len(meshes) - 1})
    for e in scene.eyes:
        eye_file = f"{name}-{e.name}.eye"
        write_eye_file(d / eye_file, e.omm)
        cameras.append({"name": e.name, "type": "perspective",
                        "perspective": {"yfov": 1.0, "znear": 0.001, "aspectRatio": 1.0},
                        "extras": {"compound-eye": "true", "compound-structure": eye_file,
                                   "compound-projection": e.projection}})
        nodes.append({"name": e.name, "camera": len(cameras) - 1, "translation": [float(v) for v in e.translation],
                      "rotation": [0.0, math.sin(e.yaw / 2), 0.0, math.cos(e.yaw / 2)]})
    for c in scene.cameras:
        if c.kind == "orthographic":
            cam = {"name": c.name, "type": "orthographic",
                   "orthographic": {"xmag": c.xmag, "ymag": c.ymag, "znear": 0.01, "zfar": 100.0}}
        else:
            cam = {"name": c.name, "type": "perspective",
                   "perspective": {"yfov": c.yfov, "znear": 0.01, "aspectRatio": 1.0}}
            if c.kind == "panoramic":
                cam["extras"] = {"panoramic": "true"}
        cameras.append(cam)
        nodes.append({"name": c.name, "camera": len(cameras) - 1, "translation": [float(v) for v in c.translation],
                      "rotation": [0.0, math.sin(c.yaw / 2), 0.0, math.cos(c.yaw / 2)]})
    bin_name = f"{name}.bin"
    (d / bin_name).write_bytes(bytes(blob))
    doc = {
        "asset": {"version": "2.0", "generator": "compound-ray_amd synthetic scenes"},
        "scene": 0,
        "scenes": [{"name": name, "nodes": list(range(len(nodes))), "extras": {"background-shader": scene.background}}],
        "nodes": nodes, "meshes": meshes, "materials": materials, "cameras": cameras,
        "accessors": accessors, "bufferViews": views,
        "buffers": [{"uri": bin_name, "byteLength": len(blob)}],
    }
    if textures:
        doc.update({"images": images, "textures": textures})
        if samplers:
            doc["samplers"] = samplers
    path = d / f"{name}.gltf"
    path.write_text(json.dumps(doc))
    return path


# ---------------------------------------------------------------- configs ----

def config1_scene(seed: int = 1) -> Scene:
    rng = np.random.default_rng(seed)
    terrain, hf = heightfield(100, 10.0, 0.5, 4, 0.3, rng, colour="uniform")
    return Scene([terrain], heightfield=hf)


def config1_eye(n: int = 1000, acceptance_deg: float = 2.0) -> np.ndarray:
    return fibonacci_eye(n, 0.001, acceptance_deg * DEG)


def config1_pose(scene: Scene, seed: int = 1) -> np.ndarray:
    """(x, y, z, rotX, rotY, rotZ): eye 0.3 m above the terrain centre, seeded yaw."""
    rng = np.random.default_rng(seed + 1000)
    yaw = float(rng.uniform(0.0, 2 * math.pi))
    return np.array([0.0, scene.ground_height(0.0, 0.0) + 0.3, 0.0, 0.0, yaw, 0.0], dtype=np.float32)


def config2_scene(seed: int = 2, grid: int = 1024, extent: float = 40.0, n_cyl: int = 2000) -> Scene:
    rng = np.random.default_rng(seed)
    terrain, hf = heightfield(grid, extent, 1.5, 6, 0.08, rng, colour="earth")
    veg = cylinders(n_cyl, hf, rng, extent)
    return Scene([terrain, veg], heightfield=hf)


def textured_scene(seed: int = 11) -> Scene:
    """Textured test scene over 10 m: a heightfield whose repeating base-colour
    texture (linear filter) is mapped by world x/z, cylinders with a second
    texture (point filter, mirrored/clamped addressing, barycentric UVs), and
    vertex-coloured cylinders."""
    rng = np.random.default_rng(seed)
    terrain, hf = heightfield(60, 10.0, 0.4, 3, 0.3, rng)
    gy, gx = np.mgrid[0:48, 0:64]
    img = np.zeros((48, 64, 4), np.uint8)
    img[..., 0] = (128 + 100 * np.sin(gx * 0.35) * np.cos(gy * 0.2)).astype(np.uint8)
    img[..., 1] = rng.integers(40, 220, size=(48, 64))
    img[..., 2] = ((gx * 4 + gy * 3) % 256).astype(np.uint8)
    img[..., 3] = 255
    terrain.col = None
    terrain.texture = Texture(img, GL_REPEAT, GL_REPEAT, GL_LINEAR)
    terrain.uv = np.stack([terrain.pos[:, 0] * 0.3, terrain.pos[:, 2] * 0.3], axis=1).astype(np.float32)
    stripes = np.zeros((8, 5, 4), np.uint8)
    stripes[..., 0] = (np.arange(5) * 60)[None, :]
    stripes[..., 1] = (np.arange(8) * 30)[:, None]
    stripes[..., 2] = 200
    stripes[..., 3] = 255
    veg = cylinders(60, hf, rng, 10.0)
    veg.col = None
    veg.texture = Texture(stripes, GL_MIRRORED_REPEAT, GL_CLAMP_TO_EDGE, GL_NEAREST, embed=False)
    veg2 = cylinders(40, hf, rng, 10.0)
    veg2.name = "vegetation-coloured"
    return Scene([terrain, veg, veg2], heightfield=hf)


def config2_eyes(seed: int = 2, n: int = 5500):
    rng = np.random.default_rng(seed + 77)
    axes = (0.0006, 0.0010, 0.0008)
    left = ellipsoid_patch_eye(n, axes, 160 * DEG, 120 * DEG, 50 * DEG, (1.5 * DEG, 3.0 * DEG), rng,
                               offset=(0.0015, 0.0, 0.0))
    right = ellipsoid_patch_eye(n, axes, 160 * DEG, 120 * DEG, -50 * DEG, (1.5 * DEG, 3.0 * DEG), rng,
                                offset=(-0.0015, 0.0, 0.0))
    return left, right


def config2_pose(scene: Scene, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed + 1000)
    yaw = float(rng.uniform(0.0, 2 * math.pi))
    return np.array([0.0, scene.ground_height(0.0, 0.0) + 0.3, 0.0, 0.0, yaw, 0.0], dtype=np.float32)


def config5_scene(seed: int = 5, grid: int = 2048, extent: float = 80.0, n_cyl: int = 20000,
                  n_rocks: int = 200, cyl_segments: int = 64) -> Scene:
    """2048^2 heightfield (8,380,418 triangles) + 20,000 64-segment cylinders
    (3,840,000) + 200 subdivision-5 icosphere rocks (4,096,000): 16,316,418
    triangles."""
    rng = np.random.default_rng(seed)
    terrain, hf = heightfield(grid, extent, 3.0, 7, 0.04, rng, colour="earth")
    veg = cylinders(n_cyl, hf, rng, extent, segments=cyl_segments)
    stones = rocks(n_rocks, hf, rng, extent)
    return Scene([terrain, veg, stones], heightfield=hf)


def config5_eyes(seed: int = 5, count: int = 256, lo: int = 1000, hi: int = 30000):
    """`count` eyes, ommatidium counts log-uniform in [lo, hi], alternating
    Fibonacci-sphere and ellipsoidal-patch layouts, acceptance U[1, 5] deg."""
    rng = np.random.default_rng(seed + 500)
    sizes = log_uniform_counts(count, lo, hi, seed + 501)
    eyes = []
    for k, n in enumerate(sizes):
        if k % 2 == 0:
            e = fibonacci_eye(int(n), 0.001, 1.0 * DEG)
            e[:, 6] = rng.uniform(1.0 * DEG, 5.0 * DEG, size=int(n)).astype(np.float32)
        else:
            axes = tuple(rng.uniform(0.0005, 0.0012, size=3))
            e = ellipsoid_patch_eye(int(n), axes, rng.uniform(90, 200) * DEG, rng.uniform(60, 150) * DEG,
                                    rng.uniform(-60, 60) * DEG, (1.0 * DEG, 5.0 * DEG), rng)
        eyes.append(e)
    return eyes


def batch_poses(scene: Scene, count: int, seed: int, extent: float, height=(0.05, 1.0),
                pitch: float = 10 * DEG) -> np.ndarray:
    """`count` poses over the terrain: height above ground U[height], yaw
    U[0, 2pi), pitch U[-pitch, pitch]."""
    rng = np.random.default_rng(seed)
    out = np.zeros((count, 6), dtype=np.float32)
    for k in range(count):
        x, z = rng.uniform(-extent / 2 * 0.8, extent / 2 * 0.8, size=2)
        h = rng.uniform(*height)
        out[k] = (x, scene.ground_height(x, z) + h, z, rng.uniform(-pitch, pitch), rng.uniform(0, 2 * math.pi), 0.0)
    return out


def log_uniform_counts(n: int, lo: int, hi: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi + 1), size=n))).clip(lo, hi).astype(np.uint32)


def build_config(cfg: int, directory, seed: int | None = None) -> dict:
    """Generates config `cfg` (1, 2 or 3) into `directory`; returns paths and arrays."""
    directory = pathlib.Path(directory)
    if cfg in (1, 3):
        s = 1 if seed is None else seed
        scene = config1_scene(s)
        n = 1000 if cfg == 1 else 5000
        eye = config1_eye(n)
        pose = config1_pose(scene, s)
        scene.eyes = [Eye("eye", eye, translation=(float(pose[0]), float(pose[1]), float(pose[2])), yaw=float(pose[4]))]
        path = write_gltf(directory, f"config{cfg}", scene)
        out = {"gltf": path, "scene": scene, "eyes": [eye], "pose": pose, "spp": 32}
        if cfg == 3:
            out["counts"] = log_uniform_counts(n, 8, 256, s + 3)
            out["poses"] = batch_poses(scene, 4096, s + 5, 10.0)
        return out
    if cfg == 2:
        s = 2 if seed is None else seed
        scene = config2_scene(s)
        left, right = config2_eyes(s)
        pose = config2_pose(scene, s)
        tr = (float(pose[0]), float(pose[1]), float(pose[2]))
        scene.eyes = [Eye("left-eye", left, translation=tr, yaw=float(pose[4])),
                      Eye("right-eye", right, translation=tr, yaw=float(pose[4]))]
        path = write_gltf(directory, "config2", scene)
        return {"gltf": path, "scene": scene, "eyes": [left, right], "pose": pose, "spp": 128}
    raise ValueError(f"unknown config {cfg}")


def pose_matrix12(pose6) -> np.ndarray:
    """Float64 reference of setCameraPose's local space (for sanity checks only)."""
    px, py, pz, rx, ry, rz = [float(v) for v in pose6]

    def rot(axis, ang):
        k = np.asarray(axis, dtype=float)
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K

    R = rot((0, 0, 1), rz) @ rot((0, 1, 0), ry) @ rot((1, 0, 0), rx)
    return np.concatenate([[px, py, pz], R[:, 0], R[:, 1], R[:, 2]])
