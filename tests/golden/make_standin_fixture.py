"""Writes the natural-standin-sky fixtures of tests/golden/ from the reference's
data files (all data; no reference code is copied or run):

natural_standin_sky.gltf.gz   data/natural-standin-sky.gltf, gzip-compressed:
                              a 24,200-triangle textured terrain under the
                              simple_sky background, with four cameras.
standin_eye.eye.gz            data/eyes/1000-horizontallyAcute-variableDegree.eye,
                              gzip-compressed (1,000 ommatidia).
equidistant_eye.eye.gz        data/eyes/1000-equidistant.eye, gzip-compressed
                              (1,000 ommatidia): the eye of the reference's
                              speed test (python-examples/speed-test/speedTest.py:81).

Run from the repository root: python tests/golden/make_standin_fixture.py
"""
import gzip
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
REF = pathlib.Path("/root/reference")


def main() -> None:
    gltf = (REF / "data" / "natural-standin-sky.gltf").read_bytes()
    (HERE / "natural_standin_sky.gltf.gz").write_bytes(gzip.compress(gltf, 9, mtime=0))
    eye = (REF / "data" / "eyes" / "1000-horizontallyAcute-variableDegree.eye").read_bytes()
    (HERE / "standin_eye.eye.gz").write_bytes(gzip.compress(eye, 9, mtime=0))
    eq = (REF / "data" / "eyes" / "1000-equidistant.eye").read_bytes()
    (HERE / "equidistant_eye.eye.gz").write_bytes(gzip.compress(eq, 9, mtime=0))


if __name__ == "__main__":
    main()
