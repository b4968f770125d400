"""Writes the test-scene fixtures of tests/golden/ from the reference's data
files (all data; no reference code is copied or run):

test_scene.gltf.gz       data/test-scene/test-scene.gltf, gzip-compressed: a
                         cube and a Suzanne head (980 triangles, u16 indices,
                         one material with baseColorFactor 0.8 on the cube and
                         none on Suzanne) under the default background (the
                         scene names no background shader), with six cameras:
                         two pinhole, one orthographic, one panoramic and two
                         compound eyes.
test_scene_eye.eye.gz    data/test-scene/test.eye (1,000 ommatidia).
test_scene_eye100.eye.gz data/test-scene/test100.eye (100 ommatidia).

Run from the repository root: python tests/golden/make_test_scene_fixture.py
"""
import gzip
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
REF = pathlib.Path("/root/reference/data/test-scene")


def main() -> None:
    for src, dst in (("test-scene.gltf", "test_scene.gltf.gz"), ("test.eye", "test_scene_eye.eye.gz"),
                     ("test100.eye", "test_scene_eye100.eye.gz")):
        (HERE / dst).write_bytes(gzip.compress((REF / src).read_bytes(), 9, mtime=0))


if __name__ == "__main__":
    main()
