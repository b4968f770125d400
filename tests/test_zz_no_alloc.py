"""renderPoseBatchInto, the allocation-free batch render, through the C ABI:
device-resident poses in, device colours out, work counters in the library's
fixed scratch array.  Two calls in a row, the second with fewer poses (and
the counters left over by the first), give renderPoseBatch's colours bit for
bit, on the library's stream and on a caller's stream."""
import ctypes

import numpy as np
import pytest

from compound_ray_amd import eyerenderer as er
from compound_ray_amd import scenes

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(lib):
    """The HIP runtime the library itself uses (same soname, already loaded)."""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


class DeviceBuffer:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def pose12_of(lib, pose6):
    """The library's pose layout (getCameraPose) of a setCameraPose pose."""
    lib.setCameraPose(*[float(v) for v in pose6])
    return er.pose_np(lib)


def test_render_pose_batch_into_matches_render_pose_batch(lib, config1, hip):
    omm = config1["eyes"][0][:400]
    lib.setSampleRecording(False)
    lib.loadGlTFscene(str(config1["gltf"]).encode())
    assert er.last_error(lib) == ""
    er.gotoFirstCompoundEye(lib)
    er.set_ommatidia_np(lib, omm)
    lib.setCurrentEyeSamplesPerOmmatidium(24)
    counts = scenes.log_uniform_counts(len(omm), 1, 160, seed=21)
    er.set_sample_counts(lib, counts)
    poses_a = scenes.batch_poses(config1["scene"], 9, seed=61, extent=10.0)
    poses_b = scenes.batch_poses(config1["scene"], 3, seed=62, extent=10.0)
    # the existing entry point (it also uploads the eye)
    _, want_a = er.render_batch_np(lib, poses_a)
    _, want_b = er.render_batch_np(lib, poses_b)
    p12_a = np.stack([pose12_of(lib, p) for p in poses_a]).astype(np.float32)
    p12_b = np.stack([pose12_of(lib, p) for p in poses_b]).astype(np.float32)
    n = len(omm)
    d_poses = DeviceBuffer(hip, p12_a.nbytes)
    d_out = DeviceBuffer(hip, len(poses_a) * n * 16)
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # first call: 9 poses on the library's stream
        d_poses.upload(p12_a)
        assert lib.renderPoseBatchInto(d_poses.ptr, len(poses_a), d_out.ptr, None) == 0, er.last_error(lib)
        assert hip.hipDeviceSynchronize() == 0
        got = d_out.download((len(poses_a), n, 4), np.float32)
        assert np.array_equal(got[..., :3], want_a)
        assert np.array_equal(got[..., 3], np.broadcast_to(counts.astype(np.float32), (len(poses_a), n)))
        # second call: fewer poses, on a caller's stream, counters left by the first call
        d_poses.upload(p12_b)
        assert lib.renderPoseBatchInto(d_poses.ptr, len(poses_b), d_out.ptr, stream) == 0, er.last_error(lib)
        assert hip.hipStreamSynchronize(stream) == 0
        got = d_out.download((len(poses_b), n, 4), np.float32)
        assert np.array_equal(got[..., :3], want_b)
        # rejected: a changed eye not yet uploaded, and misaligned output
        er.set_sample_counts(lib, None)
        assert lib.renderPoseBatchInto(d_poses.ptr, len(poses_b), d_out.ptr, None) == -1
        assert "render the eye once" in er.last_error(lib)
        er.render_batch_np(lib, poses_b, want_rgb=False)
        assert lib.renderPoseBatchInto(d_poses.ptr, len(poses_b), ctypes.c_void_p(d_out.ptr.value + 4), None) == -1
        assert "aligned" in er.last_error(lib)
    finally:
        hip.hipDeviceSynchronize()
        hip.hipStreamDestroy(stream)
        d_poses.free()
        d_out.free()
        er.set_sample_counts(lib, None)
