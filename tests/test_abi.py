"""C-ABI library: symbol exports, ctypes signatures and the host-only glTF
parse path (no GPU calls)."""
import base64
import ctypes
import json
import pathlib
import re

import numpy as np
import pytest

import compound_ray_amd as cra
from compound_ray_amd import eyerenderer as er

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "eye_renderer.h"

# The reference's exported symbols (libEyeRenderer3/libEyeRenderer.h:16-69).
REFERENCE_SYMBOLS = [
    "setVerbosity", "loadGlTFscene", "stop", "setRenderSize", "renderFrame", "displayFrame", "saveFrameAs",
    "getFramePointer", "getCameraCount", "nextCamera", "previousCamera", "getCurrentCameraIndex",
    "getCurrentCameraName", "gotoCamera", "gotoCameraByName", "setCameraPosition", "getCameraPosition",
    "setCameraLocalSpace", "rotateCameraAround", "rotateCameraLocallyAround", "translateCamera",
    "translateCameraLocally", "resetCameraPose", "setCameraPose", "isCompoundEyeActive",
    "setCurrentEyeSamplesPerOmmatidium", "getCurrentEyeSamplesPerOmmatidium",
    "changeCurrentEyeSamplesPerOmmatidiumBy", "getCurrentEyeOmmatidialCount", "setOmmatidia",
    "getCurrentEyeDataPath", "setCurrentEyeShaderName", "isInsideHitGeometry", "getGeometryMaxBounds",
    "getGeometryMinBounds", "getWindowPointer",
]


def header_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;]*\)\s*;", text, flags=re.M)
    return [n for n in names if n not in ("if", "while")]


@pytest.fixture(scope="module")
def hostlib():
    lib = cra.load_library()
    lib.setVerbosity(False)
    return lib


def test_library_exports_every_declared_symbol(hostlib):
    declared = header_functions()
    assert len(declared) >= 55
    for name in declared:
        assert hasattr(hostlib, name), name
    for name in REFERENCE_SYMBOLS:
        assert name in declared, name


def test_struct_layouts():
    assert ctypes.sizeof(er.c_ommatidiumPacket) == 32
    assert ctypes.sizeof(er.c_float3) == 12


def test_state_before_loading(hostlib):
    assert hostlib.getCameraCount() == 0
    assert not hostlib.isCompoundEyeActive()
    assert hostlib.getCurrentEyeSamplesPerOmmatidium() == -1
    assert hostlib.getCurrentEyeOmmatidialCount() == 0
    assert hostlib.getCurrentEyeDataPath() == b""
    assert hostlib.getWindowPointer() is None
    assert hostlib.renderFrame() == -1.0
    assert "no scene loaded" in er.last_error(hostlib)
    # compound setters on no camera are no-ops, as on a regular camera in the reference
    hostlib.setCurrentEyeSamplesPerOmmatidium(5)
    hostlib.gotoCamera(3)
    assert hostlib.getCurrentCameraIndex() == 0
    b = hostlib.getGeometryMaxBounds(b"nothing")
    assert (b.x, b.y, b.z) == (0.0, 0.0, 0.0)


def test_render_tuning_clamps_grab(hostlib):
    """A wave holds one grab's sample counts in its 64 lanes: larger grabs are
    clamped to 64, negative values mean automatic."""
    try:
        for asked, kept in ((200, 64), (64, 64), (7, 7), (-3, 0)):
            hostlib.setRenderTuning(asked, 5)
            assert er.render_tuning(hostlib)["units_per_grab"] == kept
            assert er.render_tuning(hostlib)["blocks"] == 5
    finally:
        hostlib.setRenderTuning(0, 0)


# ------------------------------------------------------------- glTF parse ----

def write_doc(tmp_path, name, doc, blob=None, embed=False):
    if blob is not None:
        if embed:
            doc["buffers"] = [{"uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode(),
                               "byteLength": len(blob)}]
        else:
            (tmp_path / f"{name}.bin").write_bytes(blob)
            doc["buffers"] = [{"uri": f"{name}.bin", "byteLength": len(blob)}]
    p = tmp_path / f"{name}.gltf"
    p.write_text(json.dumps(doc))
    return str(p)


def quad_doc(pos, idx, idx_type, col=None, col_type=None, col_comps=4, stride=0, nodes=None, material=None,
             extras=None):
    blob = bytearray()
    views, accs = [], []

    def view(b, st=0):
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(b), **({"byteStride": st} if st else {})})
        blob.extend(b)
        return len(views) - 1
    if stride:
        inter = np.zeros((pos.shape[0], stride // 4), np.float32)
        inter[:, :3] = pos
        pv = view(inter.tobytes(), stride)
    else:
        pv = view(pos.astype(np.float32).tobytes())
    accs.append({"bufferView": pv, "componentType": 5126, "count": pos.shape[0], "type": "VEC3"})
    attrs = {"POSITION": 0}
    if col is not None:
        cv = view(col.tobytes())
        accs.append({"bufferView": cv, "componentType": col_type, "count": col.shape[0],
                     "type": "VEC4" if col_comps == 4 else "VEC3", "normalized": col_type != 5126})
        attrs["COLOR_0"] = len(accs) - 1
    iv = view(idx.astype({5121: np.uint8, 5123: np.uint16, 5125: np.uint32}[idx_type]).tobytes())
    accs.append({"bufferView": iv, "componentType": idx_type, "count": idx.size, "type": "SCALAR"})
    prim = {"attributes": attrs, "indices": len(accs) - 1}
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0]}], "accessors": accs, "bufferViews": views,
           "meshes": [{"name": "m", "primitives": [prim], **({"extras": extras} if extras else {})}],
           "nodes": nodes or [{"mesh": 0}]}
    if material is not None:
        doc["materials"] = [{"pbrMetallicRoughness": {"baseColorFactor": material}}]
        prim["material"] = 0
    return doc, bytes(blob)


POS = np.float32([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]])
IDX = np.uint32([[0, 1, 2], [0, 2, 3]])


@pytest.mark.parametrize("idx_type", [5121, 5123, 5125])
@pytest.mark.parametrize("embed", [False, True])
def test_parse_indices_and_buffers(hostlib, tmp_path, idx_type, embed):
    col = np.float32([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [1, 1, 1, 1]])
    doc, blob = quad_doc(POS, IDX, idx_type, col, 5126)
    pos, c, mode = er.parse_gltf_np(hostlib, write_doc(tmp_path, f"q{idx_type}{embed}", doc, blob, embed))
    assert np.array_equal(pos, POS[IDX].reshape(-1, 9))
    assert np.array_equal(c, col[IDX][:, :, :3].reshape(-1, 9))
    assert np.all(mode == 0)


@pytest.mark.parametrize("ctype,scale,dtype", [(5121, 255.0, np.uint8), (5123, 65535.0, np.uint16)])
def test_parse_normalised_colours(hostlib, tmp_path, ctype, scale, dtype):
    raw = np.array([[255, 0, 128, 255], [3, 77, 200, 255], [0, 0, 0, 255], [12, 34, 56, 255]])
    raw = (raw if dtype == np.uint8 else raw * 257).astype(dtype)
    doc, blob = quad_doc(POS, IDX, 5125, raw, ctype)
    _, c, mode = er.parse_gltf_np(hostlib, write_doc(tmp_path, f"c{ctype}", doc, blob))
    want = (raw[IDX][:, :, :3].astype(np.float32) / np.float32(scale)).reshape(-1, 9)
    assert np.array_equal(c, want)
    assert np.all(mode == 0)


def test_parse_vec3_colours_use_material(hostlib, tmp_path):
    col3 = np.float32([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]])
    doc, blob = quad_doc(POS, IDX, 5125, col3, 5126, col_comps=3, material=[0.2, 0.3, 0.4, 1.0])
    _, c, mode = er.parse_gltf_np(hostlib, write_doc(tmp_path, "v3", doc, blob))
    assert np.all(mode == 1)
    assert np.allclose(c, np.tile(np.float32([0.2, 0.3, 0.4] * 3), (2, 1)))
    doc, blob = quad_doc(POS, IDX, 5125)
    _, c, mode = er.parse_gltf_np(hostlib, write_doc(tmp_path, "nocol", doc, blob))
    assert np.all(mode == 1) and np.all(c == 1.0)   # default white material


def test_parse_strided_positions(hostlib, tmp_path):
    doc, blob = quad_doc(POS, IDX, 5125, stride=24)
    pos, _, _ = er.parse_gltf_np(hostlib, write_doc(tmp_path, "strided", doc, blob))
    assert np.array_equal(pos, POS[IDX].reshape(-1, 9))


def test_parse_node_transforms(hostlib, tmp_path):
    """parent * matrix * T * R * S, children of mesh-less nodes only."""
    q = [0.0, np.sin(np.pi / 4), 0.0, np.cos(np.pi / 4)]  # 90 deg about +y
    nodes = [{"children": [1]}, {"mesh": 0, "translation": [1, 2, 3], "rotation": q, "scale": [2, 2, 2]}]
    nodes[0]["matrix"] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 10, 0, 0, 1]  # column-major: translate x by 10
    doc, blob = quad_doc(POS, IDX, 5125, nodes=nodes)
    pos, _, _ = er.parse_gltf_np(hostlib, write_doc(tmp_path, "xf", doc, blob))
    v = POS * 2.0
    rot = np.stack([v[:, 2], v[:, 1], -v[:, 0]], axis=1)   # (x, y, z) -> (z, y, -x)
    want = rot + np.float32([11, 2, 3])
    assert np.allclose(pos, want[IDX].reshape(-1, 9), atol=1e-5)


def test_parse_skips_hitbox_and_lines(hostlib, tmp_path):
    doc, blob = quad_doc(POS, IDX, 5125, extras={"hitbox": "True"})
    pos, _, _ = er.parse_gltf_np(hostlib, write_doc(tmp_path, "hb", doc, blob))
    assert pos.shape[0] == 0
    doc, blob = quad_doc(POS, IDX, 5125)
    doc["meshes"][0]["primitives"][0]["mode"] = 1
    pos, _, _ = er.parse_gltf_np(hostlib, write_doc(tmp_path, "lines", doc, blob))
    assert pos.shape[0] == 0


def test_parse_errors(hostlib, tmp_path):
    assert hostlib.parseGlTFTriangles(str(tmp_path / "missing.gltf").encode(), None, None, None, 0) == -1
    assert "cannot open" in er.last_error(hostlib)
    bad = tmp_path / "bad.gltf"
    bad.write_text("{ \"asset\": ")
    assert hostlib.parseGlTFTriangles(str(bad).encode(), None, None, None, 0) == -1
    assert "JSON" in er.last_error(hostlib)
    doc, blob = quad_doc(POS, np.uint32([[0, 1, 9]]), 5125)
    assert hostlib.parseGlTFTriangles(write_doc(tmp_path, "oob", doc, blob).encode(), None, None, None, 0) == -1
    assert "index out of range" in er.last_error(hostlib)


def _parse_error(hostlib, path):
    assert hostlib.parseGlTFTriangles(path.encode(), None, None, None, 0) == -1
    return er.last_error(hostlib)


@pytest.mark.parametrize("hitbox", [False, True])
def test_parse_rejects_malformed_accessors(hostlib, tmp_path, hitbox):
    """Untrusted accessor fields are checked before any read, on the hitbox
    path as on the mesh path: POSITION must be float VEC3 and indices unsigned
    scalars; counts, offsets and strides must be non-negative integers; a
    count or stride that would read past the buffer is refused even where
    offset + stride * (count - 1) wraps around 2^64."""
    extras = {"hitbox": "True"} if hitbox else None
    tag = "hb" if hitbox else "mesh"

    def doc_with(edit):
        doc, blob = quad_doc(POS, IDX, 5125, extras=extras)
        edit(doc)
        return write_doc(tmp_path, f"{tag}{len(list(tmp_path.iterdir()))}", doc, blob)

    def pos_acc(**kw):
        return lambda d: d["accessors"][0].update(kw)

    assert "POSITION must be float VEC3" in _parse_error(hostlib, doc_with(pos_acc(type="VEC2")))
    assert "POSITION must be float VEC3" in _parse_error(hostlib, doc_with(pos_acc(type="SCALAR")))
    assert "POSITION must be float VEC3" in _parse_error(hostlib, doc_with(pos_acc(componentType=5123)))
    assert "indices must be unsigned" in _parse_error(hostlib, doc_with(lambda d: d["accessors"][1].update(type="VEC2", count=3)))
    assert "indices must be unsigned" in _parse_error(
        hostlib, doc_with(lambda d: d["accessors"][1].update(componentType=5126)))
    for bad in (-1, 2.5, 1e300, float(2 ** 60)):
        assert "count must be" in _parse_error(hostlib, doc_with(pos_acc(count=bad))), bad
        assert "non-negative integers" in _parse_error(hostlib, doc_with(pos_acc(byteOffset=bad))), bad
        assert "non-negative integers" in _parse_error(
            hostlib, doc_with(lambda d, b=bad: d["bufferViews"][0].update(byteStride=b))), bad
    # in range for the JSON checks, but past the buffer (and, unchecked, wrapping)
    assert "past the end" in _parse_error(hostlib, doc_with(pos_acc(count=5)))
    assert "past the end" in _parse_error(hostlib, doc_with(pos_acc(count=2 ** 40)))
    assert "past the end" in _parse_error(
        hostlib, doc_with(lambda d: d["bufferViews"][0].update(byteStride=2 ** 47)))
    assert "past the end" in _parse_error(hostlib, doc_with(pos_acc(byteOffset=2 ** 47)))
    # the well-formed document still parses
    pos, _, _ = er.parse_gltf_np(hostlib, doc_with(lambda d: None))
    assert pos.shape[0] == (0 if hitbox else 2)


def test_eye_file_helpers(tmp_path):
    omms = [er.Ommatidium([0.1, 0.2, 0.3], [0, 0, 1], 0.05, 0.0), er.Ommatidium([1, 2, 3], [1, 0, 0], 0.1, 0.01)]
    p = tmp_path / "x.eye"
    er.saveEyeFile(p, omms)
    back = er.readEyeFile(p)
    assert len(back) == 2
    assert np.allclose(back[1].position, [1, 2, 3]) and back[1].focalpointOffset == pytest.approx(0.01)
    assert er.decodeProjectionMapID([1, 2, 3, 4]) == 0x01020304
