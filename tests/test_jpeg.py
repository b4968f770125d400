"""JPEG base-colour textures (host decoder, no GPU).

The reference loads glTF images through tinygltf's stb_image
(libEyeRenderer3/MulticamScene.cpp:583-599), which reads JPEG as well as PNG;
the reference's own JPEG textures (data/cube.jpg,
data/ofstad-arena/ofstad_patterning.jpg) are baseline JFIF at 4:2:0.

Two independent checks:
  * files from the test JPEG writer (tests/jpeg_writer.py) holding known
    coefficients: RGB and greyscale output must equal scipy's orthonormal
    inverse DCT-II of the dequantised coefficients (T.81's IDCT), to one
    level where float32 and float64 round a .5 tie apart;
  * Pillow's libjpeg decode of Pillow-encoded and writer-encoded files
    (YCbCr, chroma subsampling, restart markers, optimised tables): within
    3 levels per channel and a small mean difference, the spread between
    integer and float IDCTs and fixed- and float-point colour conversion.
"""
import io

import numpy as np
import pytest

from compound_ray_amd import eyerenderer as er
from jpeg_writer import ZIGZAG, idct_blocks, write_jpeg
from test_textures import textured_gltf

PIL = pytest.importorskip("PIL.Image")


@pytest.fixture()
def decode(lib, tmp_path):
    path = textured_gltf(tmp_path, np.zeros((2, 2, 4), np.uint8))

    def run(data: bytes):
        (tmp_path / "quad-tex0.png").write_bytes(data)   # the loader sniffs the format
        desc, texels = er.parse_textures_np(lib, str(path))
        if desc.shape[0] == 0:
            return None
        w, h = int(desc[0, 0]), int(desc[0, 1])
        return texels.view(np.uint8).reshape(h, w, 4)

    run.path = path
    return run


def smooth_image(h, w, seed, noise=20.0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.stack([128 + 100 * np.sin(x / 7 + y / 11), 128 + 90 * np.cos(x / 5 - y / 9), (x * 3 + y * 2) % 256], -1)
    return np.clip(img + rng.normal(0, noise, img.shape), 0, 255).astype(np.uint8)


def pillow_jpeg(img, **kw) -> bytes:
    b = io.BytesIO()
    PIL.fromarray(img).save(b, "JPEG", **kw)
    return b.getvalue()


def pillow_decode(data: bytes) -> np.ndarray:
    return np.asarray(PIL.open(io.BytesIO(data)).convert("RGB"))


def close_to_pillow(got, data, mean_tol):
    want = pillow_decode(data)
    assert got.shape[:2] == want.shape[:2]
    assert np.all(got[..., 3] == 255)
    d = np.abs(got[..., :3].astype(int) - want.astype(int))
    assert d.max() <= 3, d.max()
    assert d.mean() <= mean_tol, d.mean()


def test_zigzag_table():
    """The writer's zig-zag order (derived from the anti-diagonal walk) is a
    permutation starting 0, 1, 8, 16, 9, 2 and ending 47, 55, 62, 63."""
    assert sorted(ZIGZAG) == list(range(64))
    assert ZIGZAG[:6] == [0, 1, 8, 16, 9, 2] and ZIGZAG[-4:] == [47, 55, 62, 63]


@pytest.mark.parametrize("shape", [(37, 53), (64, 64), (17, 9), (1, 1)])
@pytest.mark.parametrize("subsampling", [0, 1, 2])
@pytest.mark.parametrize("quality", [50, 95])
def test_pillow_baseline(decode, shape, subsampling, quality):
    """4:4:4, 4:2:2 and 4:2:0 JFIF at two qualities, sizes that leave partial
    MCUs, against Pillow's decode."""
    img = smooth_image(*shape, seed=subsampling * 10 + quality)
    data = pillow_jpeg(img, subsampling=subsampling, quality=quality)
    close_to_pillow(decode(data), data, 0.1 if subsampling == 0 else 0.3)


@pytest.mark.parametrize("kw", [dict(subsampling=2, restart_marker_blocks=3),
                                dict(subsampling=1, restart_marker_rows=1),
                                dict(subsampling=0, optimize=True),
                                dict(subsampling=2, quality=100)])
def test_pillow_restarts_and_optimised_tables(decode, kw):
    img = smooth_image(45, 61, seed=7)
    data = pillow_jpeg(img, **{"quality": 90, **kw})
    close_to_pillow(decode(data), data, 0.3)


def test_pillow_greyscale(decode):
    img = smooth_image(40, 30, seed=3)[..., 0]
    b = io.BytesIO()
    PIL.fromarray(img, "L").save(b, "JPEG", quality=80)
    got = decode(b.getvalue())
    want = np.asarray(PIL.open(io.BytesIO(b.getvalue())))
    assert np.array_equal(got[..., 0], got[..., 1]) and np.array_equal(got[..., 0], got[..., 2])
    assert np.abs(got[..., 0].astype(int) - want).max() <= 1


def test_reference_format_texture(decode):
    """A 1024 x 1024 4:2:0 JFIF texture, the format of the reference's
    cube.jpg, decodes to within a level of Pillow nearly everywhere."""
    img = smooth_image(1024, 1024, seed=11, noise=4.0)
    data = pillow_jpeg(img, subsampling=2, quality=75)
    got = decode(data)
    want = pillow_decode(data)
    d = np.abs(got[..., :3].astype(int) - want.astype(int))
    assert d.max() <= 3 and d.mean() < 0.15


def random_coefficients(rng, rows, cols, scale=40.0):
    """Quantised coefficients whose magnitude falls with frequency, DC spread
    enough that some samples clip."""
    u = np.arange(64) % 8
    v = np.arange(64) // 8
    sd = scale / (1.0 + u + v)
    c = np.round(rng.normal(0, 1, (rows, cols, 64)) * sd).astype(np.int64)
    c[..., 0] = rng.integers(-70, 70, size=(rows, cols))
    c[..., rng.random(64) < 0.25] = 0          # runs of zeros, ZRL codes
    c[rng.random((rows, cols)) < 0.2, 1:] = 0  # DC-only blocks
    return c


def assert_idct_equal(got, want):
    d = np.abs(got.astype(int) - want.astype(int))
    assert d.max() <= 1
    assert (d == 0).mean() >= 0.995


@pytest.mark.parametrize("wide,extended,restart,separate", [(False, False, 0, False), (True, True, 2, False),
                                                            (False, False, 5, True), (True, True, 1, True)])
def test_rgb_coefficients_exact(decode, wide, extended, restart, separate):
    """Adobe RGB (no colour transform), 4:4:4: every channel is the IDCT of
    its coefficients.  Covers 16-bit quantisation tables under SOF1, restart
    intervals and components in separate (non-interleaved) scans."""
    rng = np.random.default_rng(int(wide) * 8 + restart * 2 + int(separate))
    W, H = 29, 21
    rows, cols = -(-H // 8), -(-W // 8)
    hi = 600 if wide else 12
    q = {0: rng.integers(1, hi, 64), 1: rng.integers(1, 12, 64)}
    comps = [dict(id=i + 1, h=1, v=1, tq=min(i, 1), coef=random_coefficients(rng, rows, cols, 40.0 if i else 4.0))
             for i in range(3)]
    if wide:
        comps[0]["coef"] //= 20   # keep the wide table's samples mostly in range
    scans = [[0], [1], [2]] if separate else [[0, 1, 2]]
    data = write_jpeg(W, H, comps, q, scans, restart=restart, extended=extended, adobe_transform=0)
    got = decode(data)
    assert got is not None and got.shape == (H, W, 4)
    for i, c in enumerate(comps):
        want = idct_blocks(c["coef"], q[c["tq"]])[:H, :W]
        assert_idct_equal(got[..., i], want)
    assert np.all(got[..., 3] == 255)


@pytest.mark.parametrize("W,H,restart", [(8, 8, 0), (13, 30, 3), (1, 1, 0), (300, 2, 7)])
def test_greyscale_coefficients_exact(decode, W, H, restart):
    rng = np.random.default_rng(W * H + restart)
    rows, cols = -(-H // 8), -(-W // 8)
    q = {2: rng.integers(1, 20, 64)}
    c = dict(id=7, h=1, v=1, tq=2, coef=random_coefficients(rng, rows, cols))
    got = decode(write_jpeg(W, H, [c], q, [[0]], restart=restart))
    want = idct_blocks(c["coef"], q[2])[:H, :W]
    assert_idct_equal(got[..., 0], want)
    assert np.array_equal(got[..., 0], got[..., 2])


@pytest.mark.parametrize("factors,scans,restart", [
    (((2, 1), (1, 1), (1, 1)), [[0, 1, 2]], 0),           # 4:2:2
    (((1, 2), (1, 1), (1, 1)), [[0, 1, 2]], 4),           # 4:4:0
    (((2, 2), (1, 1), (1, 1)), [[0], [1, 2]], 3),         # 4:2:0, luma in a scan of its own
    (((2, 2), (1, 1), (1, 1)), [[0], [1], [2]], 0),       # 4:2:0, three scans
    (((4, 1), (1, 1), (1, 1)), [[0, 1, 2]], 2),           # 4:1:1 (replicated chroma)
    (((2, 2), (2, 1), (1, 1)), [[0, 1, 2]], 0),           # mixed chroma factors
])
def test_sampling_factors_against_pillow(decode, factors, scans, restart):
    """YCbCr files with chroma subsampled in various ways and scan layouts,
    coefficients taken from a smooth image so colours stay in gamut."""
    from scipy.fft import dctn

    W, H = 45, 38
    img = smooth_image(H, W, seed=sum(sum(f) for f in factors), noise=6.0).astype(np.float64)
    r, g, b = img[..., 0], img[..., 1], img[..., 2]
    planes = [0.299 * r + 0.587 * g + 0.114 * b,
              -0.168736 * r - 0.331264 * g + 0.5 * b + 128,
              0.5 * r - 0.418688 * g - 0.081312 * b + 128]
    hmax = max(f[0] for f in factors)
    vmax = max(f[1] for f in factors)
    mcux, mcuy = -(-W // (8 * hmax)), -(-H // (8 * vmax))
    q = {0: np.full(64, 2), 1: np.full(64, 3)}
    comps = []
    for i, ((h, v), plane) in enumerate(zip(factors, planes)):
        fx, fy = hmax // h, vmax // v
        rows, cols = mcuy * v, mcux * h
        # box-average to the component's resolution, edge-replicated padding
        ph, pw = rows * 8 * fy, cols * 8 * fx
        p = np.pad(plane, ((0, max(0, ph - H)), (0, max(0, pw - W))), mode="edge")[:ph, :pw]
        p = p.reshape(rows * 8, fy, cols * 8, fx).mean(axis=(1, 3)) - 128.0
        blocks = p.reshape(rows, 8, cols, 8).transpose(0, 2, 1, 3)
        coef = dctn(blocks, type=2, norm="ortho", axes=(2, 3)).reshape(rows, cols, 64)
        tq = 0 if i == 0 else 1
        comps.append(dict(id=i + 1, h=h, v=v, tq=tq, coef=np.round(coef / q[tq]).astype(np.int64)))
    data = write_jpeg(W, H, comps, q, scans, restart=restart)
    got = decode(data)
    assert got is not None and got.shape == (H, W, 4)
    close_to_pillow(got, data, 0.4)


def test_unsupported_and_corrupt_files_are_reported(lib, decode):
    img = smooth_image(24, 24, seed=5)
    good = pillow_jpeg(img, quality=90)
    assert decode(good) is not None
    progressive = pillow_jpeg(img, quality=90, progressive=True)
    for bad in (progressive, good[: len(good) // 2], good[:2] + b"\xff\xc4\x00\x03\x00" + good[2:],
                b"\xff\xd8\xff\xd9"):
        assert decode(bad) is None   # texture dropped with a warning, base colour used
    _, _, mode = er.parse_gltf_np(lib, str(decode.path))
    assert np.all(mode == 1)
