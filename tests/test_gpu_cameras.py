"""Regular cameras (pinhole, orthographic, panoramic) on the GPU: primary rays
and frames are bit-identical to the oracle's."""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

import oracle
from compound_ray_amd import eyerenderer as er
from compound_ray_amd import scenes

pytestmark = pytest.mark.gpu

_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.tanf.argtypes = [ctypes.c_float]
_libm.tanf.restype = ctypes.c_float


def pinhole_scale(yfov_rad: float) -> np.ndarray:
    """PerspectiveCamera::setYFOV with aspect 1 (single precision, like the library)."""
    deg = np.float32(np.float32(yfov_rad) * np.float32(180.0) / np.float32(math.pi))
    rad = np.float32(deg / np.float32(180.0) * np.float32(math.pi))
    t = np.float32(_libm.tanf(float(np.float32(rad / np.float32(2.0)))))
    return np.array([t, t, 1.0], np.float32)


@pytest.fixture(scope="module")
def camera_scene(tmp_path_factory, config1):
    scene = config1["scene"]
    y = scene.ground_height(0.0, 0.0) + 0.6
    cams = [scenes.Camera("pin", "perspective", yfov=0.9, translation=(0.0, y, 0.0), yaw=0.4),
            scenes.Camera("ortho", "orthographic", xmag=2.5, ymag=1.5, translation=(0.5, y + 2.0, 0.3), yaw=-0.3),
            scenes.Camera("pano", "panoramic", translation=(-0.4, y, 0.2), yaw=1.1)]
    sc = scenes.Scene(scene.meshes, eyes=[], cameras=cams)
    path = scenes.write_gltf(tmp_path_factory.mktemp("cams"), "cams", sc)
    return {"gltf": path, "cams": cams}


@pytest.mark.parametrize("index,kind", [(0, 0), (1, 1), (2, 2)])
def test_camera_rays_and_frame(lib, camera_scene, oracle_c1, index, kind):
    lib.loadGlTFscene(str(camera_scene["gltf"]).encode())
    assert er.last_error(lib) == ""
    lib.gotoCamera(index)
    assert not lib.isCompoundEyeActive()
    cam = camera_scene["cams"][index]
    w, h = 96, 64
    er.setRenderSize(lib, w, h)
    try:
        # pitched so that every camera's view holds terrain and sky (the
        # orthographic camera looks along its +z axis, the others around it)
        pitch = 0.6 if kind == 1 else -0.35
        lib.setCameraPose(0.1 * index, float(cam.translation[1]), -0.2, pitch, 0.5 + index, 0.0)
        pose = er.pose_np(lib)
        scale = pinhole_scale(cam.yfov) if kind == 0 else np.array([cam.xmag, cam.ymag, 0.0], np.float32)
        want = oracle.camera_rays(kind, pose, scale, w, h)
        got = er.camera_rays_np(lib)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert lib.renderFrame() >= 0, er.last_error(lib)
        frame = np.array(lib.getFramePointer(), copy=True).reshape(h * w, 4)
        col = oracle_c1.radiance(want)
        assert np.array_equal(frame, oracle.make_color(col))
        # the view holds both terrain and sky
        ids, _ = oracle_c1.trace(want)
        assert 0.0 < np.mean(ids >= 0) < 1.0
    finally:
        er.setRenderSize(lib, 400, 400)
