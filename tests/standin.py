"""The reference's shipped textured scene as a test input.

natural_standin_sky.gltf.gz is data/natural-standin-sky.gltf of the reference
(a Blender export: a 24,200-triangle terrain with u16 indices under a scaled
node, one embedded 1024x1024 PNG base-colour texture, the simple_sky
background and four cameras whose eye paths are relative to the glTF), and
standin_eye.eye.gz is data/eyes/1000-horizontallyAcute-variableDegree.eye
(1,000 ommatidia), both gzip-compressed by tests/golden/make_standin_fixture.py.
"""
import gzip
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
EYE_NAME = "1000-horizontallyAcute-variableDegree.eye"
VECTOR_CAMERA = b"insect-eye-fast-vector"


def unpack(directory) -> dict:
    """Writes the scene and its eye file into `directory`; the eye is written
    both beside the glTF and under eyes/, the two relative paths the scene's
    cameras name."""
    d = pathlib.Path(directory)
    gltf = d / "natural-standin-sky.gltf"
    gltf.write_bytes(gzip.decompress((GOLDEN / "natural_standin_sky.gltf.gz").read_bytes()))
    eye = gzip.decompress((GOLDEN / "standin_eye.eye.gz").read_bytes())
    (d / "eyes").mkdir(exist_ok=True)
    (d / EYE_NAME).write_bytes(eye)
    (d / "eyes" / EYE_NAME).write_bytes(eye)
    return {"gltf": gltf, "eye": d / EYE_NAME}


def read_eye(path) -> np.ndarray:
    """(n, 8) float32 rows of a .eye file: position, direction, acceptance, focal offset."""
    rows = [[float(t) for t in line.split()] for line in pathlib.Path(path).read_text().splitlines() if line.strip()]
    return np.asarray(rows, dtype=np.float32)
