"""Textured scenes on the GPU: per-sample and per-ommatidium colours of a
scene mixing linear/point textures, all three address modes, barycentric UVs
and vertex colours are bit-identical to the oracle."""
import numpy as np
import pytest

import oracle
from compound_ray_amd import eyerenderer as er
from compound_ray_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def textured(tmp_path_factory):
    scene = scenes.textured_scene()
    eye = scenes.fibonacci_eye(600, 0.001, 3.0 * scenes.DEG)
    pose = np.array([0.3, scene.ground_height(0.3, -0.2) + 0.4, -0.2, 0.15, 0.7, 0.0], np.float32)
    scene.eyes = [scenes.Eye("eye", eye)]
    path = scenes.write_gltf(tmp_path_factory.mktemp("tex"), "textured", scene)
    return {"gltf": path, "scene": scene, "eye": eye, "pose": pose, "oracle": oracle.OracleScene.from_scene(scene)}


@pytest.mark.parametrize("spp", [16, 64])
def test_textured_render_bit_exact(lib, textured, spp):
    lib.setSampleRecording(True)
    try:
        lib.loadGlTFscene(str(textured["gltf"]).encode())
        assert er.last_error(lib) == ""
        er.gotoFirstCompoundEye(lib)
        lib.setCurrentEyeSamplesPerOmmatidium(spp)
        lib.setCameraPose(*[float(v) for v in textured["pose"]])
        assert lib.renderFrame() >= 0, er.last_error(lib)
        ref = textured["oracle"].render(textured["eye"], er.pose_np(lib), spp=spp, want_samples=True)
        samples = er.samples_np(lib)
        assert np.array_equal(samples.view(np.uint32), ref["samples"].view(np.uint32))
        assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
        # the scene really exercises every colour source
        # the view hits vertex colours (0), texture 0 (2) and texture 1 (6)
        assert {0, 2, 6}.issubset(oracle_hit_modes(textured, lib))
    finally:
        lib.setSampleRecording(False)


def oracle_hit_modes(textured, lib):
    ids, _ = textured["oracle"].trace(er.rays_np(lib))
    mode = textured["scene"].tri_mode()
    return set(int(m) for m in np.unique(mode[ids[ids >= 0]]))


def test_textured_batch_matches_oracle(lib, textured):
    lib.loadGlTFscene(str(textured["gltf"]).encode())
    er.gotoFirstCompoundEye(lib)
    lib.setCurrentEyeSamplesPerOmmatidium(24)
    poses6 = scenes.batch_poses(textured["scene"], 6, seed=12, extent=10.0)
    _, rgb = er.render_batch_np(lib, poses6)
    poses12 = np.stack([oracle.pose12(p) for p in poses6])
    ref = textured["oracle"].render(textured["eye"], poses12, spp=24, frame=lib.getCurrentEyeFrameIndex())
    assert np.array_equal(rgb, ref["colour_gpu_order"])
