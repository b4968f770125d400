"""ctypes bindings of libEyeRenderer3 (MI355X build).

``configureFunctions``/``configure_functions`` declares the same argument and
result types as the reference's helper
(python-examples/eyeRendererHelperFunctions.py:40-71), so scripts written for
the reference keep working; the other helpers here mirror that module's
public functions (packet lists, .eye files, render size, first compound eye).
``configure_extensions`` declares the additive entry points of
include/eye_renderer.h, and the ``*_np`` helpers wrap them with numpy arrays.
"""
from __future__ import annotations

import ctypes
import math
from ctypes import (POINTER, Structure, c_bool, c_char_p, c_double, c_float, c_int, c_int32, c_int64,
                    c_size_t, c_ubyte, c_uint32, c_uint64, c_void_p)

import numpy as np
from numpy.ctypeslib import ndpointer

MLP_IN = 16      # features per unit of the learned sampler
NEIGHBOURS = 6   # nearest ommatidia pooled into those features


class c_float3(Structure):
    """The 12-byte float triple returned by getGeometry{Max,Min}Bounds."""

    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]

    def toNumpy(self):
        return np.asarray([self.x, self.y, self.z])


class c_ommatidiumPacket(Structure):
    """OmmatidiumPacket (libEyeRenderer3/libEyeRenderer.h:8-14)."""

    _fields_ = [
        ("posX", c_float), ("posY", c_float), ("posZ", c_float),
        ("dirX", c_float), ("dirY", c_float), ("dirZ", c_float),
        ("acceptanceAngle", c_float), ("focalpointOffset", c_float),
    ]


class Ommatidium:
    """One ommatidium: lens position, axis, acceptance angle (rad), focal offset."""

    def __init__(self, position, direction, acceptanceAngle, focalpointOffset):
        self.position = np.asarray(position, dtype=float)
        self.direction = np.asarray(direction, dtype=float)
        self.acceptanceAngle = float(acceptanceAngle)
        self.focalpointOffset = float(focalpointOffset)

    def getSolidAngle(self):
        """Solid angle (sr) of a cone with full angle acceptanceAngle."""
        return 2.0 * math.pi * (1.0 - math.cos(self.acceptanceAngle / 2.0))

    def copy(self):
        return Ommatidium(self.position.copy(), self.direction.copy(), self.acceptanceAngle, self.focalpointOffset)




def configure_functions(lib):
    """Argument/result types of the reference ABI (libEyeRenderer.h:16-67)."""
    lib.setVerbosity.argtypes = [c_bool]
    lib.loadGlTFscene.argtypes = [c_char_p]
    lib.stop.argtypes = []
    lib.setRenderSize.argtypes = [c_int, c_int]
    lib.renderFrame.restype = c_double
    lib.saveFrameAs.argtypes = [c_char_p]
    lib.getFramePointer.restype = POINTER(c_ubyte)
    lib.getCameraCount.restype = c_size_t
    lib.getCurrentCameraIndex.restype = c_size_t
    lib.getCurrentCameraName.restype = c_char_p
    lib.gotoCamera.argtypes = [c_int]
    lib.gotoCameraByName.argtypes = [c_char_p]
    lib.gotoCameraByName.restype = c_bool
    lib.setCameraPosition.argtypes = [c_float] * 3
    lib.getCameraPosition.argtypes = [POINTER(c_float)] * 3
    lib.setCameraLocalSpace.argtypes = [c_float] * 9
    lib.rotateCameraAround.argtypes = [c_float] * 4
    lib.rotateCameraLocallyAround.argtypes = [c_float] * 4
    lib.translateCamera.argtypes = [c_float] * 3
    lib.translateCameraLocally.argtypes = [c_float] * 3
    lib.setCameraPose.argtypes = [c_float] * 6
    lib.isCompoundEyeActive.restype = c_bool
    lib.setCurrentEyeSamplesPerOmmatidium.argtypes = [c_int]
    lib.getCurrentEyeSamplesPerOmmatidium.restype = c_int
    lib.changeCurrentEyeSamplesPerOmmatidiumBy.argtypes = [c_int]
    lib.getCurrentEyeOmmatidialCount.restype = c_size_t
    lib.getCurrentEyeDataPath.restype = c_char_p
    lib.setCurrentEyeShaderName.argtypes = [c_char_p]
    lib.getGeometryMaxBounds.argtypes = [c_char_p]
    lib.getGeometryMaxBounds.restype = c_float3
    lib.getGeometryMinBounds.argtypes = [c_char_p]
    lib.getGeometryMinBounds.restype = c_float3
    lib.isInsideHitGeometry.argtypes = [c_float, c_float, c_float, c_char_p]
    lib.isInsideHitGeometry.restype = c_bool
    lib.getWindowPointer.restype = c_void_p


configureFunctions = configure_functions


def configure_extensions(lib):
    """Argument/result types of the additive entry points."""
    f32p = ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    u32p = ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    i32p = ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    i64p = ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    lib.eyeRendererGetLastError.restype = c_char_p
    lib.getOmmatidialColours.argtypes = [f32p, c_size_t]
    lib.getOmmatidialColours.restype = c_size_t
    lib.setCurrentEyeSampleCounts.argtypes = [c_void_p, c_size_t]
    lib.getCurrentEyeTotalSamples.restype = c_uint64
    lib.setRandomSeed.argtypes = [c_uint32]
    lib.getCurrentEyeFrameIndex.restype = c_uint32
    lib.setCurrentEyeFrameIndex.argtypes = [c_uint32]
    lib.getCurrentEyeIndex.restype = c_uint32
    lib.getCameraPose.argtypes = [f32p]
    lib.getCameraPose.restype = c_int
    lib.getLastRenderTimings.argtypes = [ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")]
    lib.getLastRenderTimings.restype = c_int
    lib.getCurrentEyeTraversalStats.argtypes = [ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")]
    lib.getCurrentEyeTraversalStats.restype = c_int
    lib.getCurrentEyeRays.argtypes = [f32p, c_size_t]
    lib.getCurrentEyeRays.restype = c_size_t
    lib.getCurrentEyeSamples.argtypes = [f32p, c_size_t]
    lib.getCurrentEyeSamples.restype = c_size_t
    lib.traceRays.argtypes = [f32p, c_size_t, i32p, f32p]
    lib.traceRays.restype = c_int
    lib.probePacketGuard.argtypes = [f32p, c_int, i32p]
    lib.probePacketGuard.restype = c_int
    lib.renderPoseBatch.argtypes = [f32p, c_size_t, c_void_p]
    lib.renderPoseBatch.restype = c_double
    lib.renderPoseBatchInto.argtypes = [c_void_p, c_size_t, c_void_p, c_void_p]
    lib.renderPoseBatchInto.restype = c_int
    lib.adaptiveTrainingSet.argtypes = [c_void_p, c_size_t, c_void_p, c_int, c_int, c_float, c_void_p, c_void_p,
                                        c_size_t]
    lib.adaptiveTrainingSet.restype = c_int64
    lib.nearestOmmatidia.argtypes = [c_void_p, c_size_t, c_int, c_void_p]
    lib.nearestOmmatidia.restype = c_int
    lib.renderPoseBatchAdaptive.argtypes = [c_void_p, c_size_t, c_float, c_int, c_float, c_void_p, c_void_p,
                                            c_void_p]
    lib.renderPoseBatchAdaptive.restype = c_double
    lib.getAdaptiveStages.argtypes = [c_void_p, c_void_p, c_void_p, c_size_t]
    lib.getAdaptiveStages.restype = c_int64
    lib.setAdaptiveLabelError.argtypes = [c_float]
    lib.setAdaptiveLabelError.restype = c_int
    lib.setAdaptiveSampler.argtypes = [c_int]
    lib.setAdaptiveSampler.restype = c_int
    lib.setPacketCounting.argtypes = [c_bool]
    lib.setPacketCounting.restype = None
    lib.getPacketCounts.argtypes = [ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")]
    lib.getPacketCounts.restype = c_int
    lib.getAdaptiveTimings.argtypes = [ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")]
    lib.getAdaptiveTimings.restype = c_int
    lib.setTraversalMode.argtypes = [c_int]
    lib.setCurrentEyeSampleBase.argtypes = [c_uint32]
    lib.setSampleRecording.argtypes = [c_bool]
    lib.setRenderTuning.argtypes = [c_int, c_int]
    lib.getLastRenderKernel.restype = c_char_p
    lib.setSplitLaunches.argtypes = [c_bool]
    lib.getCurrentEyeTaskSplits.restype = c_int
    lib.setTaskSplit.argtypes = [c_int]
    lib.getRenderTuning.argtypes = [i32p]
    lib.getRenderTuning.restype = c_int
    lib.setBatchAutotune.argtypes = [c_int]
    lib.setAutotuneCacheFile.argtypes = [c_char_p]
    lib.setAutotuneCacheFile.restype = c_int
    lib.autotuneRenderBatch.argtypes = [f32p, c_size_t, c_int, c_void_p, c_size_t]
    lib.autotuneRenderBatch.restype = c_int
    lib.runCopyProbe.argtypes = [c_size_t, c_int]
    lib.runCopyProbe.restype = c_double
    lib.mlpParamCount.restype = c_size_t
    lib.mlpInit.argtypes = [c_uint32]
    lib.mlpSetParams.argtypes = [c_void_p, c_size_t]
    lib.mlpGetParams.argtypes = [c_void_p, c_size_t]
    lib.mlpPredict.argtypes = [c_void_p, c_size_t, c_void_p]
    lib.mlpLossGrad.argtypes = [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]
    lib.mlpTrain.argtypes = [c_void_p, c_void_p, c_size_t, c_int, c_int, c_float, c_uint32, c_void_p]
    lib.getSceneBVHInfo.argtypes = [i64p]
    lib.getSceneBVHInfo.restype = c_int
    lib.copySceneBVHNodes.argtypes = [c_void_p, c_size_t]
    lib.copySceneBVHNodes.restype = c_size_t
    lib.copySceneSortOrder.argtypes = [c_void_p, c_void_p, c_size_t]
    lib.copySceneSortOrder.restype = c_size_t
    lib.setSceneLeafSize.argtypes = [c_int]
    lib.setSceneLeafSize.restype = c_int
    lib.parseGlTFTriangles.argtypes = [c_char_p, c_void_p, c_void_p, c_void_p, c_size_t]
    lib.parseGlTFTriangles.restype = c_int64
    lib.getCurrentCameraRays.argtypes = [c_void_p, c_size_t]
    lib.getCurrentCameraRays.restype = c_size_t
    lib.parseGlTFTextures.argtypes = [c_char_p, c_void_p, c_size_t, c_void_p, c_size_t, c_void_p]
    lib.parseGlTFTextures.restype = c_int64


# ---------------------------------------------------------------- helpers ----

def setCameraLocalSpace(eyeRenderer, npMatrix):
    """Sets the local axes from the columns of a 3x3 matrix."""
    m = np.asarray(npMatrix, dtype=float)
    eyeRenderer.setCameraLocalSpace(*m[:, 0], *m[:, 1], *m[:, 2])


def setRenderSize(eyeRenderer, width, height):
    """Resizes the frame and retypes getFramePointer as an (h, w, 4) uint8 array."""
    eyeRenderer.setRenderSize(width, height)
    eyeRenderer.getFramePointer.restype = ndpointer(dtype=c_ubyte, shape=(height, width, 4))


def setOmmatidiaFromPacketList(eyeRenderer, packetList):
    n = len(packetList)
    arr_t = c_ommatidiumPacket * n
    eyeRenderer.setOmmatidia.argtypes = [arr_t, c_size_t]
    eyeRenderer.setOmmatidia(arr_t(*packetList), c_size_t(n))


def setOmmatidiaFromOmmatidiumList(eyeRenderer, ommList):
    packets = [c_ommatidiumPacket(*[float(v) for v in o.position], *[float(v) for v in o.direction],
                                  o.acceptanceAngle, o.focalpointOffset) for o in ommList]
    setOmmatidiaFromPacketList(eyeRenderer, packets)


def set_ommatidia_np(lib, omm: np.ndarray):
    """Sets the current eye's ommatidia from an (n, 8) float32 array."""
    a = np.ascontiguousarray(omm, dtype=np.float32).reshape(-1, 8)
    lib.setOmmatidia.argtypes = [c_void_p, c_size_t]
    lib.setOmmatidia(a.ctypes.data_as(c_void_p), c_size_t(a.shape[0]))


def gotoFirstCompoundEye(eyeRenderer):
    for i in range(eyeRenderer.getCameraCount()):
        eyeRenderer.gotoCamera(int(i))
        if eyeRenderer.isCompoundEyeActive():
            return
    raise RuntimeError("Could not find a compound eye in the loaded glTF scene.")


def gotoFirstRegularCamera(eyeRenderer):
    for i in range(eyeRenderer.getCameraCount()):
        eyeRenderer.gotoCamera(int(i))
        if not eyeRenderer.isCompoundEyeActive():
            return
    raise RuntimeError("Could not find a regular camera in the loaded glTF scene.")


def _eye_line(line):
    v = [float(t) for t in line.split(" ") if t.strip()]
    return Ommatidium(v[0:3], v[3:6], v[6], v[7])


def readEyeFile(path):
    """Reads a .eye file (8 space-separated values per line) into Ommatidium objects."""
    with open(path) as f:
        return [_eye_line(line) for line in f if line.strip()]


def saveEyeFile(path, omms):
    with open(path, "w") as f:
        for o in omms:
            vals = [*o.position, *o.direction, o.acceptanceAngle, o.focalpointOffset]
            f.write(" ".join(f"{float(v):0.10f}" for v in vals) + "\n")


def decodeProjectionMapID(RGBAquadlet):
    """32-bit id packed big-endian into an RGBA pixel."""
    r, g, b, a = (int(v) for v in RGBAquadlet[:4])
    return (r << 24) | (g << 16) | (b << 8) | a


# ------------------------------------------------------- numpy wrappers ----

def last_error(lib) -> str:
    return lib.eyeRendererGetLastError().decode()


def colours_np(lib) -> np.ndarray:
    n = lib.getCurrentEyeOmmatidialCount()
    out = np.zeros((n, 3), dtype=np.float32)
    got = lib.getOmmatidialColours(out, n)
    return out[:got]


def set_sample_counts(lib, counts) -> None:
    if counts is None:
        lib.setCurrentEyeSampleCounts(None, 0)
        return
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    lib.setCurrentEyeSampleCounts(c.ctypes.data_as(c_void_p), c.size)


def pose_np(lib) -> np.ndarray:
    out = np.zeros(12, dtype=np.float32)
    if lib.getCameraPose(out) != 0:
        raise RuntimeError("no camera")
    return out


def timings_np(lib) -> np.ndarray:
    """(trace, reduce, projection, sum) GPU ms of the last renderFrame."""
    out = np.zeros(4, dtype=np.float64)
    lib.getLastRenderTimings(out)
    return out


def traversal_stats(lib) -> dict:
    out = np.zeros(3, dtype=np.uint64)
    if lib.getCurrentEyeTraversalStats(out) != 0:
        raise RuntimeError(last_error(lib))
    return {"nodes_visited": int(out[0]), "tris_tested": int(out[1]), "hits": int(out[2])}


def rays_np(lib) -> np.ndarray:
    n = int(lib.getCurrentEyeTotalSamples())
    out = np.zeros((n, 8), dtype=np.float32)
    got = lib.getCurrentEyeRays(out, n)
    return out[:got]


def samples_np(lib) -> np.ndarray:
    n = int(lib.getCurrentEyeTotalSamples())
    out = np.zeros((n, 4), dtype=np.float32)
    got = lib.getCurrentEyeSamples(out, n)
    return out[:got]


def trace_np(lib, rays: np.ndarray):
    r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
    ids = np.zeros(r.shape[0], dtype=np.int32)
    tbb = np.zeros((r.shape[0], 3), dtype=np.float32)
    if lib.traceRays(r, r.shape[0], ids, tbb) != 0:
        raise RuntimeError(last_error(lib))
    return ids, tbb


def render_batch_np(lib, poses6: np.ndarray, want_rgb: bool = True):
    p = np.ascontiguousarray(poses6, dtype=np.float32).reshape(-1, 6)
    n = lib.getCurrentEyeOmmatidialCount()
    out = np.zeros((p.shape[0], n, 3), dtype=np.float32) if want_rgb else None
    ms = lib.renderPoseBatch(p, p.shape[0], out.ctypes.data_as(c_void_p) if want_rgb else None)
    if ms < 0:
        raise RuntimeError(last_error(lib))
    return ms, out


def render_tuning(lib) -> dict:
    v = np.zeros(3, np.int32)
    lib.getRenderTuning(v)
    return {"units_per_grab": int(v[0]), "blocks": int(v[1]), "trace_mode": int(v[2])}


def autotune_np(lib, poses6: np.ndarray, repeats: int = 1) -> np.ndarray:
    """Times the candidate launch shapes on these poses and installs the
    fastest; returns rows (units_per_grab, blocks, traversal, ms) with
    traversal 0 = packet, 1 = packet with scalar loads, 2 = per-ray."""
    p = np.ascontiguousarray(poses6, dtype=np.float32).reshape(-1, 6)
    rows = np.zeros((128, 4), np.float32)
    n = lib.autotuneRenderBatch(p, p.shape[0], repeats, rows.ctypes.data_as(c_void_p), rows.shape[0])
    if n < 0:
        raise RuntimeError(last_error(lib))
    return rows[:n]


def nearest_ommatidia_np(lib, omm: np.ndarray, k: int = NEIGHBOURS) -> np.ndarray:
    """(n, k) indices of each ommatidium's k nearest others (the learned
    sampler's neighbour table; host computation, no GPU)."""
    o = np.ascontiguousarray(omm, dtype=np.float32).reshape(-1, 8)
    out = np.zeros((o.shape[0], k), np.uint32)
    if lib.nearestOmmatidia(o.ctypes.data_as(c_void_p), o.shape[0], k, out.ctypes.data_as(c_void_p)) != 0:
        raise RuntimeError(last_error(lib))
    return out


def training_set_np(lib, poses6, unit_mask=None, pilot=8, label_spp=1024, error_target=0.01):
    """(features (n, MLP_IN), labels (n,)) for the masked (pose, ommatidium) units."""
    p = np.ascontiguousarray(poses6, dtype=np.float32).reshape(-1, 6)
    n_omm = lib.getCurrentEyeOmmatidialCount()
    m = None if unit_mask is None else np.ascontiguousarray(unit_mask, dtype=np.uint8).reshape(-1)
    rows = int(m.sum()) if m is not None else p.shape[0] * n_omm
    f = np.zeros((rows, MLP_IN), np.float32)
    lab = np.zeros(rows, np.float32)
    got = lib.adaptiveTrainingSet(p.ctypes.data_as(c_void_p), p.shape[0],
                                  None if m is None else m.ctypes.data_as(c_void_p), pilot, label_spp,
                                  error_target, f.ctypes.data_as(c_void_p), lab.ctypes.data_as(c_void_p), rows)
    if got < 0:
        raise RuntimeError(last_error(lib))
    return f[:got], lab[:got]


def mlp_train_np(lib, features, labels, epochs=10, batch=4096, lr=1e-3, seed=1):
    x = np.ascontiguousarray(features, dtype=np.float32)
    t = np.ascontiguousarray(labels, dtype=np.float32)
    losses = np.zeros(epochs, np.float32)
    if lib.mlpTrain(x.ctypes.data_as(c_void_p), t.ctypes.data_as(c_void_p), x.shape[0], epochs, batch, lr, seed,
                    losses.ctypes.data_as(c_void_p)) != 0:
        raise RuntimeError(last_error(lib))
    return losses


def mlp_predict_np(lib, features):
    x = np.ascontiguousarray(features, dtype=np.float32)
    y = np.zeros(x.shape[0], np.float32)
    if lib.mlpPredict(x.ctypes.data_as(c_void_p), x.shape[0], y.ctypes.data_as(c_void_p)) != 0:
        raise RuntimeError(last_error(lib))
    return y


def render_adaptive_np(lib, poses6, error_target=0.01, max_spp=512, budget_scale=1.0, want_rgb=True):
    p = np.ascontiguousarray(poses6, dtype=np.float32).reshape(-1, 6)
    n = lib.getCurrentEyeOmmatidialCount()
    rgb = np.zeros((p.shape[0], n, 3), np.float32) if want_rgb else None
    counts = np.zeros((p.shape[0], n), np.uint32) if want_rgb else None
    total = np.zeros(1, np.uint64)
    ms = lib.renderPoseBatchAdaptive(p.ctypes.data_as(c_void_p), p.shape[0], error_target, max_spp, budget_scale,
                                     rgb.ctypes.data_as(c_void_p) if want_rgb else None,
                                     counts.ctypes.data_as(c_void_p) if want_rgb else None,
                                     total.ctypes.data_as(c_void_p))
    if ms < 0:
        raise RuntimeError(last_error(lib))
    return ms, rgb, counts, int(total[0])


def packet_counts(lib) -> dict:
    """Packet node visits and leaf triangles counted since the last read (see setPacketCounting)."""
    out = np.zeros(2, np.uint64)
    if lib.getPacketCounts(out) != 0:
        raise RuntimeError(last_error(lib))
    return {"nodes": int(out[0]), "tris": int(out[1])}


def adaptive_timings_np(lib) -> np.ndarray:
    """(pilot, allocation, top-up, combination) device ms of the last adaptive render."""
    out = np.zeros(4, np.float64)
    if lib.getAdaptiveTimings(out) != 0:
        raise RuntimeError(last_error(lib))
    return out


def adaptive_stages_np(lib, units: int):
    """(pilot (units, 4), top-up (units, 4), top-up counts) of the last adaptive render."""
    a = np.zeros((units, 4), np.float32)
    b = np.zeros((units, 4), np.float32)
    e = np.zeros(units, np.uint32)
    if lib.getAdaptiveStages(a.ctypes.data_as(c_void_p), b.ctypes.data_as(c_void_p), e.ctypes.data_as(c_void_p),
                             units) < 0:
        raise RuntimeError(last_error(lib))
    return a, b, e


def bvh_info(lib) -> dict:
    info = np.zeros(8, dtype=np.int64)
    if lib.getSceneBVHInfo(info) != 0:
        raise RuntimeError("no scene loaded")
    return {"tris": int(info[0]), "nodes": int(info[1]), "stack_bound": int(info[2]), "levels": int(info[3]),
            "leaf_size": int(info[4]), "build_ms": info[5] / 1000.0, "node_bytes": int(info[6])}


def bvh_nodes_np(lib) -> np.ndarray:
    info = bvh_info(lib)
    buf = np.zeros(info["nodes"] * info["node_bytes"], dtype=np.uint8)
    got = lib.copySceneBVHNodes(buf.ctypes.data_as(c_void_p), buf.size)
    return buf[:got]


def sort_order_np(lib):
    n = bvh_info(lib)["tris"]
    ids = np.zeros(n, dtype=np.uint32)
    codes = np.zeros(n, dtype=np.uint32)
    lib.copySceneSortOrder(ids.ctypes.data_as(c_void_p), codes.ctypes.data_as(c_void_p), n)
    return ids, codes


def parse_gltf_np(lib, path: str):
    n = lib.parseGlTFTriangles(path.encode(), None, None, None, 0)
    if n < 0:
        raise RuntimeError(last_error(lib))
    pos = np.zeros((n, 9), dtype=np.float32)
    col = np.zeros((n, 9), dtype=np.float32)
    mode = np.zeros(n, dtype=np.uint32)
    lib.parseGlTFTriangles(path.encode(), pos.ctypes.data_as(c_void_p), col.ctypes.data_as(c_void_p),
                           mode.ctypes.data_as(c_void_p), n)
    return pos, col, mode


def camera_rays_np(lib) -> np.ndarray:
    """(h * w, 8) primary rays of the current regular camera."""
    n = lib.getCurrentCameraRays(None, 0)
    out = np.zeros((n, 8), np.float32)
    if n:
        lib.getCurrentCameraRays(out.ctypes.data_as(c_void_p), n)
    return out


def parse_textures_np(lib, path: str):
    """(descriptors (n, 6) int32, texels uint32) of the glTF's base-colour textures."""
    count = np.zeros(1, np.int64)
    n = lib.parseGlTFTextures(path.encode(), None, 0, None, 0, count.ctypes.data_as(c_void_p))
    if n < 0:
        raise RuntimeError(last_error(lib))
    desc = np.zeros((n, 6), np.int32)
    texels = np.zeros(int(count[0]), np.uint32)
    lib.parseGlTFTextures(path.encode(), desc.ctypes.data_as(c_void_p), n, texels.ctypes.data_as(c_void_p),
                          texels.size, count.ctypes.data_as(c_void_p))
    return desc, texels
