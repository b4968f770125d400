"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host code that
reads untrusted files (glTF/JSON, PNG/DEFLATE, JPEG, .eye), the neighbour table and
the CPU oracle (CPU only; tests/sanitize/).  Well-formed inputs, the
reference's own scene files, and seeded truncations and byte corruptions of
them must load or be rejected with an error, never touch memory out of
bounds, leak, or hit undefined behaviour."""
import gzip
import io
import os
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from compound_ray_amd import scenes
from jpeg_writer import write_jpeg
from test_textures import write_png

Image = pytest.importorskip("PIL.Image")

HERE = pathlib.Path(__file__).resolve().parent
SAN_DIR = HERE / "sanitize"
GOLDEN = HERE / "golden"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")


@pytest.fixture(scope="module")
def driver():
    if shutil.which("make") is None or shutil.which("g++") is None:
        pytest.skip("make / g++ not available")
    res = subprocess.run(["make", "-s", "-j8", "-C", str(SAN_DIR)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
    return SAN_DIR / "_build" / "driver"


def run(driver, *args):
    res = subprocess.run([str(driver), *map(str, args)], capture_output=True, text=True, env=ENV, timeout=600)
    report = res.stderr[-4000:]
    assert "AddressSanitizer" not in res.stderr and "LeakSanitizer" not in res.stderr, report
    assert "runtime error" not in res.stderr, report
    assert res.returncode == 0, (res.returncode, report)
    return res.stdout


def mutants(data: bytes, seed: int, n: int):
    """Seeded truncations and byte corruptions of `data`."""
    rng = np.random.default_rng(seed)
    out = [data[:k] for k in sorted(set(int(v) for v in rng.integers(0, len(data), size=n // 2)))]
    for _ in range(n - len(out)):
        b = bytearray(data)
        for p in rng.integers(0, len(b), size=int(rng.integers(1, 6))):
            b[p] = int(rng.integers(0, 256))
        out.append(bytes(b))
    return out


@pytest.fixture(scope="module")
def corpus(tmp_path_factory):
    d = tmp_path_factory.mktemp("sancorpus")
    files = {}
    for src, dst in (("test_scene.gltf.gz", "test-scene.gltf"), ("test_scene_eye.eye.gz", "test.eye"),
                     ("test_scene_eye100.eye.gz", "test100.eye"),
                     ("natural_standin_sky.gltf.gz", "natural-standin-sky.gltf")):
        (d / dst).write_bytes(gzip.decompress((GOLDEN / src).read_bytes()))
    eye = gzip.decompress((GOLDEN / "standin_eye.eye.gz").read_bytes())
    (d / "1000-horizontallyAcute-variableDegree.eye").write_bytes(eye)
    (d / "eyes").mkdir()
    (d / "eyes" / "1000-horizontallyAcute-variableDegree.eye").write_bytes(eye)
    files["test_scene"] = d / "test-scene.gltf"
    files["standin"] = d / "natural-standin-sky.gltf"
    files["eyes"] = [d / "test.eye", d / "test100.eye"]
    tex = scenes.textured_scene()
    tex.eyes = [scenes.Eye("eye", scenes.fibonacci_eye(80, 0.001, 0.05), translation=(0.0, 1.0, 0.0))]
    files["textured"] = scenes.write_gltf(d / "textured", "textured", tex)
    rng = np.random.default_rng(4)
    pngs = []
    for k, (ctype, depth) in enumerate([(0, 1), (0, 16), (2, 8), (3, 4), (4, 16), (6, 8)]):
        ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
        hi = 16 if ctype == 3 else 2 ** depth
        px = rng.integers(0, hi, size=(9, 13, ch))
        pal = rng.integers(0, 256, size=(16, 3)) if ctype == 3 else None
        p = d / f"img{k}.png"
        p.write_bytes(write_png(px, ctype, depth, interlace=k % 2 == 1, palette=pal))
        pngs.append(p)
    files["pngs"] = pngs
    jpegs = []
    img = rng.integers(0, 256, size=(19, 27, 3)).astype(np.uint8)
    for k, kw in enumerate([dict(subsampling=2), dict(subsampling=1, restart_marker_blocks=2), dict(subsampling=0),
                            dict(progressive=True)]):
        b = io.BytesIO()
        Image.fromarray(img).save(b, "JPEG", quality=85, **kw)
        p = d / f"img{k}.jpg"
        p.write_bytes(b.getvalue())
        jpegs.append(p)
    q = {0: rng.integers(1, 400, 64)}
    comp = [dict(id=1, h=2, v=2, tq=0, coef=rng.integers(-5, 5, size=(4, 4, 64))),
            dict(id=2, h=1, v=1, tq=0, coef=rng.integers(-5, 5, size=(2, 2, 64))),
            dict(id=3, h=1, v=1, tq=0, coef=rng.integers(-5, 5, size=(2, 2, 64)))]
    p = d / "writer.jpg"
    p.write_bytes(write_jpeg(30, 25, comp, q, [[0], [1, 2]], restart=1, extended=True))
    jpegs.append(p)
    files["jpegs"] = jpegs
    return d, files


def test_parsers_on_well_formed_files(driver, corpus):
    d, files = corpus
    out = run(driver, "gltf", files["test_scene"], files["standin"], files["textured"])
    assert out.count(": ok (") == 3
    out = run(driver, "png", *files["pngs"])
    assert out.count(": ok ") == len(files["pngs"])
    out = run(driver, "jpeg", *files["jpegs"])
    assert out.count(": ok ") == len(files["jpegs"]) - 1 and "progressive JPEG is not supported" in out
    out = run(driver, "eye", *files["eyes"])
    assert "(1000 ommatidia)" in out and "(100 ommatidia)" in out


@pytest.mark.parametrize("n,k", [(1, 6), (7, 6), (300, 6), (5000, 16)])
def test_neighbour_table(driver, n, k):
    run(driver, "knn", n, k, 11)


@pytest.mark.parametrize("scene,spp", [("test_scene", 8), ("standin", 4), ("textured", 6)])
def test_oracle_render(driver, corpus, scene, spp):
    out = run(driver, "render", corpus[1][scene], spp)
    assert "agree: yes" in out


def test_parsers_on_corrupted_files(driver, corpus, tmp_path):
    d, files = corpus
    png_bytes = files["pngs"][2].read_bytes()
    cases = []
    for k, m in enumerate(mutants(png_bytes, 1, 120)):
        p = tmp_path / f"png{k}.png"
        p.write_bytes(m)
        cases.append(p)
    run(driver, "png", *cases)
    cases = []
    for j, src in enumerate((files["jpegs"][0], files["jpegs"][1], files["jpegs"][4])):
        for k, m in enumerate(mutants(src.read_bytes(), 10 + j, 80)):
            p = tmp_path / f"jpeg{j}_{k}.jpg"
            p.write_bytes(m)
            cases.append(p)
    run(driver, "jpeg", *cases)
    # glTF documents with an embedded buffer: JSON and accessor corruption
    gltf = files["test_scene"].read_bytes()
    cases = []
    for k, m in enumerate(mutants(gltf, 2, 160)):
        p = d / f"mut{k}.gltf"   # beside the eye files the scene names
        p.write_bytes(m)
        cases.append(p)
    run(driver, "gltf", *cases)
    # the textured scene's glTF (external PNGs beside it)
    tex = files["textured"]
    cases = []
    for k, m in enumerate(mutants(tex.read_bytes(), 3, 80)):
        p = tex.parent / f"mut{k}.gltf"
        p.write_bytes(m)
        cases.append(p)
    run(driver, "gltf", *cases)
    # .eye files
    eye = files["eyes"][1].read_bytes()
    cases = []
    for k, m in enumerate(mutants(eye, 4, 40)):
        p = tmp_path / f"eye{k}.eye"
        p.write_bytes(m)
        cases.append(p)
    run(driver, "eye", *cases)
