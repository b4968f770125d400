"""Hitbox meshes and geometry bounds (MulticamScene.cpp:1756-1818,
sutil/hitscanprocessing.cpp:20-105) through the C ABI of a loaded scene."""
import json
import math

import numpy as np
import pytest

from compound_ray_amd import eyerenderer as er

pytestmark = pytest.mark.gpu

CUBE_POS = np.float32([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)])
# 12 outward-wound triangles of the cube (vertex index = 4x + 2y + z with 0/1 bits)
CUBE_IDX = np.uint32([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                      [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]])


def write_scene(tmp_path):
    blob = bytearray()
    views, accs = [], []

    def view(b):
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(b)})
        blob.extend(b)
        return len(views) - 1

    ground = np.float32([[-5, 0, -5], [5, 0, -5], [5, 0, 5], [-5, 0, 5]])
    gi = np.uint32([[0, 1, 2], [0, 2, 3]])
    accs.append({"bufferView": view(ground.tobytes()), "componentType": 5126, "count": 4, "type": "VEC3",
                 "min": [-5, 0, -5], "max": [5, 0, 5]})
    accs.append({"bufferView": view(gi.tobytes()), "componentType": 5125, "count": 6, "type": "SCALAR"})
    accs.append({"bufferView": view(CUBE_POS.tobytes()), "componentType": 5126, "count": 8, "type": "VEC3"})
    accs.append({"bufferView": view(CUBE_IDX.tobytes()), "componentType": 5125, "count": 36, "type": "SCALAR"})
    q = [0.0, math.sin(0.3), 0.0, math.cos(0.3)]   # 0.6 rad about y
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0, 1]}], "accessors": accs, "bufferViews": views,
           "buffers": [{"uri": "hb.bin", "byteLength": len(blob)}],
           "meshes": [{"name": "ground", "primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]},
                      {"name": "box", "extras": {"hitbox": "true"},
                       "primitives": [{"attributes": {"POSITION": 2}, "indices": 3}]}],
           "nodes": [{"mesh": 0, "translation": [0, -1, 0]},
                     {"mesh": 1, "translation": [2, 0.5, -1], "rotation": q, "scale": [0.5, 1.0, 2.0]}]}
    (tmp_path / "hb.bin").write_bytes(bytes(blob))
    path = tmp_path / "hb.gltf"
    path.write_text(json.dumps(doc))
    return path


def box_world(p):
    """World point -> cube object space of the hitbox node."""
    c, s = math.cos(0.6), math.sin(0.6)
    q = np.asarray(p, float) - [2, 0.5, -1]
    # inverse of R_y(0.6): rotate by -0.6
    x = c * q[0] - s * q[2]
    z = s * q[0] + c * q[2]
    return np.array([x / 0.5, q[1] / 1.0, z / 2.0])


def test_hitbox_queries_and_bounds(lib, tmp_path):
    lib.loadGlTFscene(str(write_scene(tmp_path)).encode())
    assert er.last_error(lib) == ""
    rng = np.random.default_rng(1)
    pts = rng.uniform([0.5, -1.0, -3.5], [3.5, 2.0, 1.5], size=(300, 3))
    inside = 0
    for p in pts:
        o = box_world(p)
        margin = np.min(1.0 - np.abs(o))
        if abs(margin) < 1e-3:
            continue   # too close to a face to call
        got = lib.isInsideHitGeometry(float(p[0]), float(p[1]), float(p[2]), b"box")
        assert bool(got) == (margin > 0), (p, o)
        inside += margin > 0
    assert 20 < inside < 280
    assert not lib.isInsideHitGeometry(2.0, 0.5, -1.0, b"no-such-box")
    # hitbox bounds: box around the transformed corners of the object box
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    c, s = math.cos(0.6), math.sin(0.6)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    world = (R @ (corners * [0.5, 1.0, 2.0]).T).T + [2, 0.5, -1]
    hi, lo = lib.getGeometryMaxBounds(b"box"), lib.getGeometryMinBounds(b"box")
    assert np.allclose([hi.x, hi.y, hi.z], world.max(axis=0), atol=1e-5)
    assert np.allclose([lo.x, lo.y, lo.z], world.min(axis=0), atol=1e-5)
    # render mesh bounds from the accessor's min/max, translated
    hi, lo = lib.getGeometryMaxBounds(b"ground"), lib.getGeometryMinBounds(b"ground")
    assert [hi.x, hi.y, hi.z] == [5, -1, 5] and [lo.x, lo.y, lo.z] == [-5, -1, -5]
    hi = lib.getGeometryMaxBounds(b"unknown")
    assert [hi.x, hi.y, hi.z] == [0, 0, 0]
