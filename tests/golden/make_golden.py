"""Writes the committed fixtures of tests/golden/.

philox_rocrand.json  Philox4x32-10 outputs of AMD's open-source rocRAND engine
                     (host build of philox_rocrand.cpp with hipcc): independent
                     known answers for the oracle's and the GPU's generator.
oracle_c1.npz        Regression anchors of the CPU oracle on configuration 1
                     (rays, hit ids, colours of a few ommatidia).  They pin the
                     oracle against unnoticed drift; correctness itself is pinned
                     by the known-answer, closed-form and brute-force tests.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import pathlib
import subprocess
import sys
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))


def philox_vectors():
    with tempfile.TemporaryDirectory() as d:
        exe = pathlib.Path(d) / "kat"
        subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1",
                        str(HERE / "philox_rocrand.cpp"), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    (HERE / "philox_rocrand.json").write_text(json.dumps(data, indent=1) + "\n")


def oracle_fixture():
    import __graft_entry__
    __graft_entry__.import_package()
    from compound_ray_amd import scenes
    import oracle

    with tempfile.TemporaryDirectory() as d:
        cfg = scenes.build_config(1, d)
    O = oracle.OracleScene.from_scene(cfg["scene"])
    pose = oracle.pose12(cfg["pose"])
    omm = cfg["eyes"][0][:64]
    rays = oracle.generate_rays(omm, pose, spp=32)
    ids, tbb = O.trace(rays, bvh=False)
    r = O.render(omm, pose, spp=32)
    np.savez_compressed(HERE / "oracle_c1.npz", pose6=cfg["pose"], pose12=pose, omm=omm, rays=rays, ids=ids,
                        tbb=tbb, colour_gpu_order=r["colour_gpu_order"], colour_seq=r["colour_seq"])


if __name__ == "__main__":
    philox_vectors()
    oracle_fixture()
