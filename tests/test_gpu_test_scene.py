"""The reference's test scene (data/test-scene/test-scene.gltf) on the GPU.

The scene names no background shader, so its misses take
__miss__default_background (shaders.cu:740-747), the default miss program of
MulticamScene.h:187.  Its two compound eyes (test.eye, 1,000 ommatidia, and
test100.eye, 100 ommatidia, both with 2-radian acceptance cones) see the cube,
the Suzanne head and a large share of that background, so the default miss
colour, its black border bands and the material colours all reach the
per-sample records, the ommatidial colours and the uint8 frames, which are
bit-identical to the oracle fed the same triangles, eye file and pose.  The
regular cameras of the same file (pinhole, orthographic, panoramic) are
checked the same way.
"""
import gzip
import pathlib

import numpy as np
import pytest

import oracle
from compound_ray_amd import eyerenderer as er
from standin import read_eye
from test_gpu_cameras import pinhole_scale

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
SKY_DEFAULT = 0


@pytest.fixture(scope="module")
def test_scene(tmp_path_factory, lib):
    d = tmp_path_factory.mktemp("testscene")
    for src, dst in (("test_scene.gltf.gz", "test-scene.gltf"), ("test_scene_eye.eye.gz", "test.eye"),
                     ("test_scene_eye100.eye.gz", "test100.eye")):
        (d / dst).write_bytes(gzip.decompress((GOLDEN / src).read_bytes()))
    pos, col, mode = er.parse_gltf_np(lib, str(d / "test-scene.gltf"))
    assert pos.shape == (980, 9)
    return {"gltf": d / "test-scene.gltf",
            "eyes": {b"insect-cam-1": read_eye(d / "test.eye"), b"insect-cam-2": read_eye(d / "test100.eye")},
            "oracle": oracle.OracleScene(pos, col, mode, SKY_DEFAULT, 4)}


def load(lib, test_scene):
    lib.loadGlTFscene(str(test_scene["gltf"]).encode())
    assert er.last_error(lib) == ""
    assert lib.getCameraCount() == 6


@pytest.mark.parametrize("camera,n", [(b"insect-cam-1", 1000), (b"insect-cam-2", 100)])
@pytest.mark.parametrize("spp,record", [(16, True), (128, False)])
def test_compound_eyes_default_background(lib, test_scene, camera, n, spp, record):
    load(lib, test_scene)
    assert lib.gotoCameraByName(camera)
    assert lib.isCompoundEyeActive()
    assert lib.getCurrentEyeOmmatidialCount() == n
    omm = test_scene["eyes"][camera]
    osc = test_scene["oracle"]
    lib.setSampleRecording(record)
    lib.setCurrentEyeSamplesPerOmmatidium(spp)
    w, h = 64, 32
    er.setRenderSize(lib, w, h)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        if spp > 64 and not record:
            # uniform counts above one wave: chunks of 64 samples as separate tasks
            assert lib.getLastRenderKernel() == b"k_render_tasks"
        frame_index = lib.getCurrentEyeFrameIndex() - 1
        pose = er.pose_np(lib)
        eye = lib.getCurrentEyeIndex()
        ref = osc.render(omm, pose, spp=spp, eye=eye, frame=frame_index, want_samples=record)
        colours = er.colours_np(lib)
        assert np.array_equal(colours, ref["colour_gpu_order"][0])
        if record:
            assert np.array_equal(er.samples_np(lib).view(np.uint32), ref["samples"].view(np.uint32))
        # the eye's projection (spherical_orientationwise, from the camera's extras)
        frame = np.array(lib.getFramePointer(), copy=True).reshape(h, w, 4)
        ids = oracle.spherical_ids(omm, w, h, 1)
        assert np.array_equal(frame, oracle.make_color(colours)[ids])
        # the default background is on the tested path: misses with its
        # gradient colour and misses in its black border bands
        rays = oracle.generate_rays(omm, pose, spp=spp, eye=eye, frame=frame_index)
        hit, _ = osc.trace(rays)
        miss = rays[hit < 0]
        assert (hit >= 0).any() and len(miss) > 0
        sky = osc.radiance(miss)
        black = np.all(sky == 0.0, axis=1)
        assert black.any() and (~black).any()
        assert np.all(sky[:, 2] == 0.0)
    finally:
        lib.setSampleRecording(False)
        er.setRenderSize(lib, 400, 400)


@pytest.mark.parametrize("camera,kind", [(b"Camera", 0), (b"ortho-camera", 1), (b"panoramic-camera", 2)])
def test_regular_cameras_default_background(lib, test_scene, camera, kind):
    load(lib, test_scene)
    assert lib.gotoCameraByName(camera)
    assert not lib.isCompoundEyeActive()
    w, h = 80, 60
    er.setRenderSize(lib, w, h)
    try:
        pose = er.pose_np(lib)
        if kind == 0:
            scale = pinhole_scale(0.39959652046304894)
        elif kind == 1:
            scale = np.float32([16.014286041259766, 16.014286041259766, 0.0])
        else:
            scale = np.zeros(3, np.float32)
        rays = oracle.camera_rays(kind, pose, scale, w, h)
        assert np.array_equal(er.camera_rays_np(lib).view(np.uint32), rays.view(np.uint32))
        assert lib.renderFrame() >= 0, er.last_error(lib)
        frame = np.array(lib.getFramePointer(), copy=True).reshape(h * w, 4)
        assert np.array_equal(frame, oracle.make_color(test_scene["oracle"].radiance(rays)))
        ids, _ = test_scene["oracle"].trace(rays)
        assert np.mean(ids >= 0) > 0.0
    finally:
        er.setRenderSize(lib, 400, 400)
