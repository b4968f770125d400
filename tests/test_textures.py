"""Base-colour textures: PNG decoding of the loader, texture descriptors of a
glTF, and the oracle's texture sampler (pinned by closed forms)."""
import base64
import json
import pathlib
import struct
import zlib

import numpy as np
import pytest

import oracle
from compound_ray_amd import eyerenderer as er
from compound_ray_amd import scenes

import standin


# ------------------------------------------------------------ PNG writer ----

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def pack_row(samples: np.ndarray, depth: int) -> bytes:
    """One row of samples (already flattened over channels) at `depth` bits."""
    s = samples.astype(np.uint32)
    if depth == 16:
        return s.astype(">u2").tobytes()
    if depth == 8:
        return s.astype(np.uint8).tobytes()
    per = 8 // depth
    out = bytearray((len(s) * depth + 7) // 8)
    for k, v in enumerate(s):
        out[k // per] |= int(v) << (8 - depth - (k % per) * depth)
    return bytes(out)


def filter_row(row: bytes, prev: bytes, bpp: int, ft: int) -> bytes:
    cur = np.frombuffer(row, np.uint8).astype(np.int32)
    up = np.frombuffer(prev, np.uint8).astype(np.int32) if prev else np.zeros_like(cur)
    left = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])[:len(cur)]
    ul = np.concatenate([np.zeros(bpp, np.int32), up[:-bpp]])[:len(cur)]
    if ft == 0:
        pred = np.zeros_like(cur)
    elif ft == 1:
        pred = left
    elif ft == 2:
        pred = up
    elif ft == 3:
        pred = (left + up) >> 1
    else:
        p = left + up - ul
        pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
        pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
    return bytes([ft]) + ((cur - pred) & 255).astype(np.uint8).tobytes()


def write_png(samples: np.ndarray, ctype: int, depth: int, interlace: bool = False, palette=None, trns=None,
              level: int = 6) -> bytes:
    """samples: (h, w, channels) integers at `depth` bits (palette indices for ctype 3)."""
    h, w = samples.shape[:2]
    ch = CHANNELS[ctype]
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    ft = 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = b""
        for y in range(sub.shape[0]):
            row = pack_row(sub[y].reshape(-1), depth)
            raw += filter_row(row, prev, bpp, ft % 5)
            ft += 1
            prev = row

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xffffffff)

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if palette is not None:
        out += chunk(b"PLTE", np.asarray(palette, np.uint8).tobytes())
    if trns is not None:
        out += chunk(b"tRNS", trns)
    data = zlib.compress(bytes(raw), level)
    # split IDAT across chunks to exercise concatenation
    half = len(data) // 2
    return out + chunk(b"IDAT", data[:half]) + chunk(b"IDAT", data[half:]) + chunk(b"IEND", b"")


def to8(v: np.ndarray, depth: int) -> np.ndarray:
    v = v.astype(np.uint64)
    if depth == 16:
        return ((v * 255 + 32767) // 65535).astype(np.uint8)
    return (v * 255 // (2 ** depth - 1)).astype(np.uint8)


def textured_gltf(tmp_path, rgba, **tex_kwargs) -> pathlib.Path:
    """A one-quad scene whose texture is written as an external PNG file."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    tex = scenes.Texture(rgba, embed=False, **tex_kwargs)
    scene = scenes.Scene([scenes.Mesh("quad", pos, idx, None, texture=tex, uv=uv)])
    return scenes.write_gltf(tmp_path, "quad", scene)


# ------------------------------------------------------------------ tests ----

@pytest.mark.parametrize("ctype,depth", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2),
                                         (3, 4), (3, 8), (4, 8), (4, 16), (6, 8), (6, 16)])
@pytest.mark.parametrize("interlace", [False, True])
def test_png_decoder(lib, tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + interlace)
    h, w = 13, 11   # odd sizes leave partial Adam7 passes and partial bytes
    ch = CHANNELS[ctype]
    samples = rng.integers(0, 2 ** depth, size=(h, w, ch))
    palette = trns = None
    if ctype == 3:
        ncol = 2 ** depth
        palette = rng.integers(0, 256, size=(ncol, 3))
        alpha = rng.integers(0, 256, size=max(1, ncol // 2)).astype(np.uint8)
        trns = alpha.tobytes()
        want = np.zeros((h, w, 4), np.uint8)
        want[..., :3] = palette[samples[..., 0]]
        a = np.full(ncol, 255, np.uint8)
        a[:len(alpha)] = alpha
        want[..., 3] = a[samples[..., 0]]
    elif ctype == 0:
        key = int(samples[0, 0, 0])
        trns = struct.pack(">H", key)
        g = to8(samples[..., 0], depth)
        want = np.stack([g, g, g, np.where(samples[..., 0] == key, 0, 255).astype(np.uint8)], -1)
    elif ctype == 2:
        key = samples[1, 2]
        trns = struct.pack(">HHH", *[int(v) for v in key])
        hit = np.all(samples == key, axis=-1)
        want = np.concatenate([to8(samples, depth), np.where(hit, 0, 255).astype(np.uint8)[..., None]], -1)
    elif ctype == 4:
        g = to8(samples[..., 0], depth)
        want = np.stack([g, g, g, to8(samples[..., 1], depth)], -1)
    else:
        want = to8(samples, depth)
    path = textured_gltf(tmp_path, np.zeros((2, 2, 4), np.uint8))
    (tmp_path / "quad-tex0.png").write_bytes(write_png(samples, ctype, depth, interlace, palette, trns))
    desc, texels = er.parse_textures_np(lib, str(path))
    assert desc.shape == (1, 6) and tuple(desc[0, :2]) == (w, h)
    got = texels.view(np.uint8).reshape(h, w, 4)
    assert np.array_equal(got, want)


def test_png_stored_and_fixed_blocks(lib, tmp_path):
    """zlib level 0 (stored blocks) and a tiny image (fixed Huffman codes)."""
    rng = np.random.default_rng(3)
    for level, shape in ((0, (40, 50)), (9, (1, 1)), (1, (3, 2))):
        img = rng.integers(0, 256, size=shape + (4,)).astype(np.uint8)
        path = textured_gltf(tmp_path, img)
        (tmp_path / "quad-tex0.png").write_bytes(write_png(img, 6, 8, level=level))
        _, texels = er.parse_textures_np(lib, str(path))
        assert np.array_equal(texels.view(np.uint8).reshape(img.shape), img)


def test_corrupt_png_is_reported(lib, tmp_path):
    img = np.full((4, 4, 4), 7, np.uint8)
    path = textured_gltf(tmp_path, img)
    png = bytearray(write_png(img, 6, 8))
    png[-20] ^= 0xff   # inside the last IDAT: CRC mismatch
    (tmp_path / "quad-tex0.png").write_bytes(bytes(png))
    desc, _ = er.parse_textures_np(lib, str(path))
    assert desc.shape[0] == 0   # texture dropped with a warning, base colour used
    _, _, mode = er.parse_gltf_np(lib, str(path))
    assert np.all(mode == 1)


def test_texture_descriptors_and_uvs(lib, tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, size=(6, 9, 4)).astype(np.uint8)
    path = textured_gltf(tmp_path, img, wrap_s=scenes.GL_MIRRORED_REPEAT, wrap_t=scenes.GL_CLAMP_TO_EDGE,
                         min_filter=scenes.GL_NEAREST)
    desc, texels = er.parse_textures_np(lib, str(path))
    assert desc.tolist() == [[9, 6, 2, 1, 0, 0]]
    assert np.array_equal(texels.view(np.uint8).reshape(6, 9, 4), img)
    pos, col, mode = er.parse_gltf_np(lib, str(path))
    assert np.all(mode == 2)
    assert np.array_equal(col[0, :6], [0, 0, 1, 0, 1, 1]) and np.array_equal(col[1, :6], [0, 0, 1, 1, 0, 1])
    # without a sampler: repeat + linear (MulticamScene.cpp:606-610); without UVs: barycentrics
    path2 = textured_gltf(tmp_path / "b", img, min_filter=None)
    assert er.parse_textures_np(lib, str(path2))[0].tolist() == [[9, 6, 0, 0, 1, 0]]
    doc = json.loads(path2.read_text())
    del doc["meshes"][0]["primitives"][0]["attributes"]["TEXCOORD_0"]
    path2.write_text(json.dumps(doc))
    _, col2, _ = er.parse_gltf_np(lib, str(path2))
    assert np.array_equal(col2[:, :6], np.tile([0, 0, 1, 0, 0, 1], (2, 1)))


def python_png_rgba(data: bytes) -> np.ndarray:
    """Minimal independent decoder (zlib from the Python standard library) for
    8-bit non-interlaced RGB / RGBA PNGs."""
    pos, idat = 8, b""
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, lace = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    assert depth == 8 and ctype in (2, 6) and lace == 0
    ch = 3 if ctype == 2 else 4
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * ch)
    out = np.zeros((h, w * ch), np.int32)
    for y in range(h):
        ft, cur = raw[y, 0], raw[y, 1:].astype(np.int32)
        up = out[y - 1] if y else np.zeros(w * ch, np.int32)
        row = np.zeros(w * ch, np.int32)
        for i in range(w * ch):
            a = row[i - ch] if i >= ch else 0
            b = up[i]
            c = up[i - ch] if i >= ch else 0
            if ft == 0:
                p = 0
            elif ft == 1:
                p = a
            elif ft == 2:
                p = b
            elif ft == 3:
                p = (a + b) >> 1
            else:
                q = a + b - c
                pa, pb, pc = abs(q - a), abs(q - b), abs(q - c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            row[i] = (cur[i] + p) & 255
        out[y] = row
    img = out.reshape(h, w, ch).astype(np.uint8)
    if ch == 3:
        img = np.concatenate([img, np.full((h, w, 1), 255, np.uint8)], -1)
    return img


def test_reference_scene_texture(lib, tmp_path):
    """The reference's textured scene (natural-standin-sky.gltf): its embedded
    PNG decodes to the same texels as an independent decoder, and its sampler
    (LINEAR / LINEAR_MIPMAP_LINEAR, no wrap given) maps to repeat + linear."""
    REF_SCENE = standin.unpack(tmp_path)["gltf"]
    desc, texels = er.parse_textures_np(lib, str(REF_SCENE))
    doc = json.loads(REF_SCENE.read_text())
    bv = doc["bufferViews"][doc["images"][0]["bufferView"]]
    uri = doc["buffers"][bv["buffer"]]["uri"]
    blob = base64.b64decode(uri.split(",", 1)[1]) if uri.startswith("data:") else \
        (REF_SCENE.parent / uri).read_bytes()
    png = blob[bv.get("byteOffset", 0):bv.get("byteOffset", 0) + bv["byteLength"]]
    want = python_png_rgba(png)
    assert desc.shape[0] == 1 and tuple(desc[0]) == (want.shape[1], want.shape[0], 0, 0, 1, 0)
    sub = slice(0, 64)
    got = texels.view(np.uint8).reshape(want.shape)
    assert np.array_equal(got[sub], want[sub])
    pos, col, mode = er.parse_gltf_np(lib, str(REF_SCENE))
    assert np.count_nonzero((mode & 3) == 2) > 0


# ------------------------------------------- oracle sampler, closed forms ----

def sampler_scene(rgba, wrap=(0, 0), filt=1):
    h, w = rgba.shape[:2]
    px = rgba.astype(np.uint32).reshape(-1, 4)
    texels = px[:, 0] | px[:, 1] << 8 | px[:, 2] << 16 | px[:, 3] << 24
    desc = np.array([[w, h, wrap[0], wrap[1], filt, 0]], np.int32)
    tri = np.zeros((1, 9), np.float32)
    tri[0, 3] = tri[0, 7] = 1
    return oracle.OracleScene(tri, np.zeros((1, 9), np.float32), np.array([2], np.uint32), 1, 4,
                              textures=(desc, texels))


def test_oracle_sampler_constant_and_centres():
    rng = np.random.default_rng(8)
    const = np.full((5, 7, 4), (10, 200, 77, 255), np.uint8)
    want = np.float32([10, 200, 77]) / np.float32(255)
    for filt in (0, 1):
        S = sampler_scene(const, (0, 0), filt)
        for u, v in rng.uniform(-3, 3, size=(50, 2)):
            got = S.texture_sample(0, u, v)
            # point: the texel itself; linear: weights summing to 1, within rounding
            assert np.array_equal(got, want) if filt == 0 else np.allclose(got, want, rtol=0, atol=3e-7)
    img = rng.integers(0, 256, size=(4, 6, 4)).astype(np.uint8)
    for filt in (0, 1):
        for wrap in ((0, 0), (1, 1), (2, 2)):
            S = sampler_scene(img, wrap, filt)
            for j in range(4):
                for i in range(6):
                    got = S.texture_sample(0, (i + 0.5) / 6, (j + 0.5) / 4)
                    assert np.array_equal(got, img[j, i, :3].astype(np.float32) / np.float32(255)), (filt, wrap)


def test_oracle_sampler_addressing():
    """Point sampling outside [0, 1): repeat is periodic, clamp holds the edge
    texel, mirrored repeat reflects odd periods."""
    img = np.zeros((1, 4, 4), np.uint8)
    img[0, :, 0] = [10, 20, 30, 40]
    col = lambda S, u: int(round(S.texture_sample(0, u, 0.5)[0] * 255))   # noqa: E731
    rep, clamp, mirror = (sampler_scene(img, (m, m), 0) for m in (0, 1, 2))
    for k, u in enumerate([0.125, 0.375, 0.625, 0.875]):
        for p in (-2, -1, 1, 3):
            assert col(rep, u + p) == img[0, k, 0]
        assert col(clamp, u - 5) == 10 and col(clamp, u + 5) == 40
        assert col(mirror, u + 1) == img[0, 3 - k, 0]      # odd period: reflected
        assert col(mirror, u - 1) == img[0, 3 - k, 0]
        assert col(mirror, u + 2) == img[0, k, 0]


def test_oracle_sampler_linear_ramp():
    """Bilinear filtering of a ramp reproduces the ramp between texel centres
    (to the 1/256 weight step), and blends across the seam per address mode."""
    w = 16
    img = np.zeros((2, w, 4), np.uint8)
    img[:, :, 0] = (np.arange(w) * 16)[None, :]
    img[:, :, 1] = 255
    S = sampler_scene(img, (1, 1), 1)
    for u in np.linspace(0.5 / w, 1 - 0.5 / w, 97):
        x = u * w - 0.5
        want = 16 * x / 255.0
        got = S.texture_sample(0, float(u), 0.25)
        assert abs(got[0] - want) <= 16 / 255 / 256 + 1e-6
        assert got[1] == 1.0
    # at u = 0 the footprint straddles the left edge: repeat blends texels w-1 and 0
    # equally, clamp and mirrored repeat both give texel 0
    rep, clamp, mirror = (sampler_scene(img, (m, m), 1) for m in (0, 1, 2))
    assert abs(rep.texture_sample(0, 0.0, 0.25)[0] - 0.5 * (16 * (w - 1)) / 255) < 1e-6
    assert clamp.texture_sample(0, 0.0, 0.25)[0] == 0.0
    assert mirror.texture_sample(0, 0.0, 0.25)[0] == 0.0
