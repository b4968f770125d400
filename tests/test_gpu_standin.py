"""The reference's own shipped scene, end to end on the GPU: natural-standin-sky.gltf
(a Blender export with u16 indices, a scaled terrain node, camera nodes under
rotated parents, an embedded PNG texture and eye files named relative to the
glTF) loaded through loadGlTFscene and rendered through its
insect-eye-fast-vector camera.  Samples, ommatidial colours and the uint8
frame are bit-identical to the oracle fed the same triangles, texture, eye file
and pose."""
import numpy as np
import pytest

import oracle
import standin
from compound_ray_amd import eyerenderer as er

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def standin_scene(tmp_path_factory, lib):
    files = standin.unpack(tmp_path_factory.mktemp("standin"))
    pos, col, mode = er.parse_gltf_np(lib, str(files["gltf"]))
    tex = er.parse_textures_np(lib, str(files["gltf"]))
    return {"files": files, "eye": standin.read_eye(files["eye"]),
            "oracle": oracle.OracleScene(pos, col, mode, 1, 4, textures=tex)}


def load_vector_camera(lib, standin_scene):
    lib.loadGlTFscene(str(standin_scene["files"]["gltf"]).encode())
    assert er.last_error(lib) == ""
    assert lib.getCameraCount() == 4
    assert lib.gotoCameraByName(standin.VECTOR_CAMERA)
    assert lib.isCompoundEyeActive()
    assert lib.getCurrentEyeOmmatidialCount() == 1000


def test_standin_camera_axes(lib, standin_scene):
    """The camera node's world transform is the product of a +90 and a -90
    degree turn about x (identity up to rounding); the eye's local axes are
    the transform's x, y and -z (MulticamScene.cpp:215-217)."""
    load_vector_camera(lib, standin_scene)
    pose = er.pose_np(lib)
    want = np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, -1])
    assert np.allclose(pose, want, rtol=0, atol=2e-7)


@pytest.mark.parametrize("spp", [8, 96])
def test_standin_render_bit_exact(lib, standin_scene, spp):
    load_vector_camera(lib, standin_scene)
    lib.setCurrentEyeSamplesPerOmmatidium(spp)
    er.setRenderSize(lib, 1000, 1)
    lib.setSampleRecording(True)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        frame_index = lib.getCurrentEyeFrameIndex() - 1
        ref = standin_scene["oracle"].render(standin_scene["eye"], er.pose_np(lib), spp=spp,
                                             eye=lib.getCurrentEyeIndex(), frame=frame_index, want_samples=True)
        assert np.array_equal(er.samples_np(lib).view(np.uint32), ref["samples"].view(np.uint32))
        colours = er.colours_np(lib)
        assert np.array_equal(colours, ref["colour_gpu_order"][0])
        frame = np.array(lib.getFramePointer(), copy=True)
        assert np.array_equal(frame[0], oracle.make_color(colours))
        # the view holds both textured terrain and sky
        ids, _ = standin_scene["oracle"].trace(er.rays_np(lib))
        assert 0.2 < np.mean(ids >= 0) < 0.8
    finally:
        lib.setSampleRecording(False)
        er.setRenderSize(lib, 400, 400)
