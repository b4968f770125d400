"""The reference's camera calls through the C ABI, against a float64
restatement of DataRecordCamera.h:48-87 (libEyeRenderer.cpp:340-390):
resetPose, setPosition/getPosition, setLocalSpace, rotateAround (Rodrigues
rotation of the three axes about the normalised axis), rotateLocallyAround
(the axis first taken from local to world space), move and moveLocally.
Camera navigation wraps modulo the camera count (MulticamScene.cpp:928-945);
gotoCameraByName leaves camera 0 current when no name matches
(libEyeRenderer.cpp:330-340); the sample-count setters clamp at one sample and
are no-ops on a regular camera (CompoundEye.cpp:171-183); saveFrameAs writes a
P6 file of the frame's RGB bytes, rows bottom-up (sutil.cpp:387-418, 82-103)."""
import ctypes
import gzip
import math
import pathlib

import numpy as np
import pytest

from compound_ray_amd import eyerenderer as er

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def test_scene_path(tmp_path_factory):
    d = tmp_path_factory.mktemp("camcalls")
    for src, dst in (("test_scene.gltf.gz", "test-scene.gltf"), ("test_scene_eye.eye.gz", "test.eye"),
                     ("test_scene_eye100.eye.gz", "test100.eye")):
        (d / dst).write_bytes(gzip.decompress((GOLDEN / src).read_bytes()))
    return d / "test-scene.gltf"


def load(lib, path):
    lib.loadGlTFscene(str(path).encode())
    assert er.last_error(lib) == ""


def rotate_point(p, angle, axis):
    k = np.asarray(axis, np.float64)
    k = k / np.linalg.norm(k)
    return math.cos(angle) * p + math.sin(angle) * np.cross(k, p) + (1 - math.cos(angle)) * np.dot(k, p) * k


def model_of(pose12):
    P = np.asarray(pose12, np.float64)
    return {"pos": P[0:3].copy(), "x": P[3:6].copy(), "y": P[6:9].copy(), "z": P[9:12].copy()}


def as12(m):
    return np.concatenate([m["pos"], m["x"], m["y"], m["z"]])


def to_local(m, v):
    return v[0] * m["x"] + v[1] * m["y"] + v[2] * m["z"]


def op_rotate(m, angle, *axis):
    for k in ("x", "y", "z"):
        m[k] = rotate_point(m[k], angle, axis)


def op_rotate_locally(m, angle, *axis):
    op_rotate(m, angle, *to_local(m, np.asarray(axis, np.float64)))


def op_move(m, *step):
    m["pos"] = m["pos"] + np.asarray(step, np.float64)


def op_move_locally(m, *step):
    m["pos"] = m["pos"] + to_local(m, np.asarray(step, np.float64))


def test_camera_pose_calls_match_restatement(lib, test_scene_path):
    load(lib, test_scene_path)
    assert lib.gotoCameraByName(b"insect-cam-1")
    lib.resetCameraPose()
    assert np.array_equal(er.pose_np(lib), np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1]))
    lib.setCameraPosition(1.5, -2.25, 3.0)
    x, y, z = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    lib.getCameraPosition(ctypes.byref(x), ctypes.byref(y), ctypes.byref(z))
    assert (x.value, y.value, z.value) == (1.5, -2.25, 3.0)
    c, s = math.cos(0.4), math.sin(0.4)
    space = np.float32([c, 0, -s, 0, 1, 0, s, 0, c])
    lib.setCameraLocalSpace(*[float(v) for v in space])
    assert np.array_equal(er.pose_np(lib)[3:], space)
    assert np.array_equal(er.pose_np(lib)[:3], np.float32([1.5, -2.25, 3.0]))
    ops = [(lib.rotateCameraAround, op_rotate, (0.7, 1.0, 2.0, -0.5)),
           (lib.rotateCameraLocallyAround, op_rotate_locally, (-1.1, 0.0, 1.0, 0.3)),
           (lib.translateCamera, op_move, (0.5, 0.0, -1.0)),
           (lib.translateCameraLocally, op_move_locally, (1.0, 2.0, 3.0)),
           (lib.rotateCameraLocallyAround, op_rotate_locally, (2.9, 0.0, 0.0, 1.0)),
           (lib.rotateCameraAround, op_rotate, (-2.2, 0.0, -3.0, 0.0)),
           (lib.translateCameraLocally, op_move_locally, (-0.25, 0.0, 4.0))]
    for call, op, args in ops:
        m = model_of(er.pose_np(lib))
        call(*args)
        op(m, *args)
        got = er.pose_np(lib).astype(np.float64)
        assert np.allclose(got, as12(m), rtol=0, atol=4e-6), (call.__name__, got, as12(m))
    # setCameraPose = reset, rotate about world x, y, z, then move
    lib.setCameraPose(0.3, 1.2, -0.7, 0.2, -1.3, 0.45)
    m = model_of(np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1]))
    op_rotate(m, 0.2, 1, 0, 0)
    op_rotate(m, -1.3, 0, 1, 0)
    op_rotate(m, 0.45, 0, 0, 1)
    op_move(m, 0.3, 1.2, -0.7)
    assert np.allclose(er.pose_np(lib), as12(m), rtol=0, atol=4e-6)


def test_camera_navigation_and_sample_setters(lib, test_scene_path):
    load(lib, test_scene_path)
    n = lib.getCameraCount()
    assert n == 6
    names = []
    for i in range(n):
        lib.gotoCamera(i)
        assert lib.getCurrentCameraIndex() == i
        names.append(lib.getCurrentCameraName())
    assert set(names) == {b"Camera", b"far-camera", b"insect-cam-1", b"ortho-camera", b"panoramic-camera",
                          b"insect-cam-2"}
    lib.gotoCamera(n - 1)
    lib.nextCamera()
    assert lib.getCurrentCameraIndex() == 0
    lib.previousCamera()
    assert lib.getCurrentCameraIndex() == n - 1
    lib.gotoCamera(n + 2)
    assert lib.getCurrentCameraIndex() == 2
    lib.gotoCamera(-1)
    assert lib.getCurrentCameraIndex() == n - 1
    lib.gotoCamera(3)
    assert not lib.gotoCameraByName(b"no-such-camera")
    assert lib.getCurrentCameraIndex() == 0
    assert lib.gotoCameraByName(b"insect-cam-2")
    assert lib.getCurrentCameraName() == b"insect-cam-2"
    assert lib.isCompoundEyeActive()
    assert lib.getCurrentEyeDataPath().endswith(b"test100.eye")
    lib.setCurrentEyeSamplesPerOmmatidium(8)
    lib.changeCurrentEyeSamplesPerOmmatidiumBy(5)
    assert lib.getCurrentEyeSamplesPerOmmatidium() == 13
    lib.changeCurrentEyeSamplesPerOmmatidiumBy(-100)
    assert lib.getCurrentEyeSamplesPerOmmatidium() == 1
    # regular cameras: compound setters are no-ops
    assert lib.gotoCameraByName(b"ortho-camera")
    assert not lib.isCompoundEyeActive()
    lib.changeCurrentEyeSamplesPerOmmatidiumBy(5)
    lib.setCurrentEyeSamplesPerOmmatidium(20)
    assert lib.getCurrentEyeSamplesPerOmmatidium() == -1
    assert lib.getCurrentEyeOmmatidialCount() == 0
    assert lib.getCurrentEyeDataPath() == b""
    lib.gotoCameraByName(b"insect-cam-2")
    assert lib.getCurrentEyeSamplesPerOmmatidium() == 1


def test_save_frame_as_p6(lib, test_scene_path, tmp_path):
    load(lib, test_scene_path)
    assert lib.gotoCameraByName(b"Camera")
    w, h = 37, 11
    er.setRenderSize(lib, w, h)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        frame = np.array(lib.getFramePointer(), copy=True).reshape(h, w, 4)
        assert len(np.unique(frame.reshape(-1, 4), axis=0)) > 1
        out = tmp_path / "frame.ppm"
        lib.saveFrameAs(str(out).encode())
        assert er.last_error(lib) == ""
        data = out.read_bytes()
        header = f"P6\n{w} {h}\n255\n".encode()
        assert data[:len(header)] == header
        assert data[len(header):] == frame[::-1, :, :3].tobytes()
    finally:
        er.setRenderSize(lib, 400, 400)
