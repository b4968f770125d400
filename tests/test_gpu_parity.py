"""GPU parity of the HIP renderer against the CPU oracle, through the C ABI.

Bar: rays, hit ids, hit distances, per-sample colours and per-ommatidium
colours (canonical wave order) are bit-identical to the oracle; colours are
within 1e-5 relative of the reference's sequential summation order; the BVH
node array is byte-identical to the oracle's CPU build of the same codes.
"""
import numpy as np
import pytest

import oracle
from compound_ray_amd import eyerenderer as er
from compound_ray_amd import scenes

pytestmark = pytest.mark.gpu


def load(lib, path):
    lib.loadGlTFscene(str(path).encode())
    err = er.last_error(lib)
    assert err == "", err
    er.gotoFirstCompoundEye(lib)


def setup_eye(lib, cfg, omm=None, spp=32, pose6=None, record=True):
    lib.setSampleRecording(record)
    load(lib, cfg["gltf"])
    if omm is not None:
        er.set_ommatidia_np(lib, omm)
    lib.setCurrentEyeSamplesPerOmmatidium(spp)
    p = cfg["pose"] if pose6 is None else pose6
    lib.setCameraPose(*[float(v) for v in p])


def rel_close(a, b, rtol):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-12))


def test_pose_matches_oracle(lib, config1):
    setup_eye(lib, config1)
    got = er.pose_np(lib)
    want = oracle.pose12(config1["pose"])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_bvh_bit_exact(lib, config1, oracle_c1):
    load(lib, config1["gltf"])
    info = er.bvh_info(lib)
    oinfo = oracle_c1.info()
    assert info["tris"] == oinfo["tris"] == config1["scene"].tri_count()
    assert info["nodes"] == oinfo["nodes"]
    assert info["stack_bound"] == oinfo["stack_bound"]
    ids, codes = er.sort_order_np(lib)
    oids, ocodes = oracle_c1.order()
    assert np.array_equal(codes, ocodes)
    assert np.array_equal(ids, oids)
    assert np.array_equal(er.bvh_nodes_np(lib), oracle_c1.nodes())


@pytest.mark.parametrize("leaf", [1, 2, 7, 16])
def test_bvh_leaf_sizes(lib, config1, leaf):
    load(lib, config1["gltf"])
    assert lib.setSceneLeafSize(leaf) == 0
    o = oracle.OracleScene.from_scene(config1["scene"], leaf_size=leaf)
    assert np.array_equal(er.bvh_nodes_np(lib), o.nodes())
    setup_eye(lib, config1, omm=config1["eyes"][0][:128], spp=8)
    assert lib.setSceneLeafSize(leaf) == 0
    assert lib.renderFrame() >= 0
    want = o.render(config1["eyes"][0][:128], er.pose_np(lib), spp=8)["colour_gpu_order"][0]
    assert np.array_equal(er.colours_np(lib), want)
    assert lib.setSceneLeafSize(4) == 0


def test_rays_bit_exact(lib, config1):
    setup_eye(lib, config1, spp=32)
    rays = er.rays_np(lib)
    want = oracle.generate_rays(config1["eyes"][0], er.pose_np(lib), spp=32, seed=42,
                                eye=lib.getCurrentEyeIndex(), frame=lib.getCurrentEyeFrameIndex())
    assert rays.shape == (1000 * 32, 8)
    assert np.array_equal(rays.view(np.uint32), want.view(np.uint32))


def test_hits_match_oracle(lib, config1, oracle_c1):
    setup_eye(lib, config1, spp=32)
    rays = er.rays_np(lib)
    ids, tbb = er.trace_np(lib, rays)
    oids, otbb = oracle_c1.trace(rays, bvh=True)
    agree = np.mean(ids == oids)
    assert agree >= 0.9999, agree
    assert np.array_equal(ids, oids)
    hit = ids >= 0
    assert np.array_equal(tbb[hit].view(np.uint32), otbb[hit].view(np.uint32))
    # and against brute force over every triangle on a subset
    bids, btbb = oracle_c1.trace(rays[:3000], bvh=False)
    assert np.array_equal(ids[:3000], bids)


def test_render_bit_exact_c1(lib, config1, oracle_c1):
    setup_eye(lib, config1, spp=32)
    ms = lib.renderFrame()
    assert ms >= 0, er.last_error(lib)
    samples = er.samples_np(lib)
    col = er.colours_np(lib)
    ref = oracle_c1.render(config1["eyes"][0], er.pose_np(lib), spp=32, want_samples=True)
    assert np.array_equal(samples.view(np.uint32), ref["samples"].view(np.uint32))
    assert np.array_equal(col, ref["colour_gpu_order"][0])
    # the reference's own (sequential) summation order: within 1e-5 relative
    assert rel_close(col, ref["colour_seq"][0], 1e-5)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("spp", [7, 32, 64, 100])
def test_traversal_modes_bit_exact(lib, config1, oracle_c1, mode, spp):
    """Per-ray and wave-packet traversal (vector or scalar node loads) give the
    oracle's samples exactly."""
    lib.setTraversalMode(mode)
    try:
        omm = config1["eyes"][0][::3]
        setup_eye(lib, config1, omm=omm, spp=spp)
        assert lib.renderFrame() >= 0, er.last_error(lib)
        ref = oracle_c1.render(omm, er.pose_np(lib), spp=spp, want_samples=True)
        assert np.array_equal(er.samples_np(lib).view(np.uint32), ref["samples"].view(np.uint32))
        assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
    finally:
        lib.setTraversalMode(2)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("spp", [1, 32, 33, 64, 65])
def test_split_launch_routing(lib, config1, oracle_c1, mode, spp):
    """One pose with uniform counts: each unit's 64-sample chunks are tasks of
    their own (k_render_tasks for packets at any count, k_render_chunks per
    ray above 32 samples); per ray at 32 or fewer the unit kernel packs units
    into lanes.  Either way, and with split launches switched off, the
    colours are the oracle's."""
    lib.setTraversalMode(mode)
    try:
        omm = config1["eyes"][0][::2]
        setup_eye(lib, config1, omm=omm, spp=spp, record=False)
        for split in (True, False):
            lib.setSplitLaunches(split)
            assert lib.renderFrame() >= 0, er.last_error(lib)
            want = b"k_render_units"
            if split and mode == 1:
                want = b"k_render_tasks"
            elif split and spp > 32:
                want = b"k_render_chunks"
            assert lib.getLastRenderKernel() == want
            ref = oracle_c1.render(omm, er.pose_np(lib), spp=spp, eye=lib.getCurrentEyeIndex(),
                                   frame=lib.getCurrentEyeFrameIndex() - 1)
            assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
    finally:
        lib.setSplitLaunches(True)
        lib.setTraversalMode(2)
        lib.setSampleRecording(True)


def test_eye_changes_between_renders(lib, config1, oracle_c1):
    """One eye re-rendered after each kind of change: the sample rate (up and
    down), uneven counts, back to a uniform rate, new ommatidia of another
    count, and the same ommatidia again.  The library uploads only what
    changed (the ommatidia and their processing order, or the count and
    sample-record offsets); every render's samples and colours equal the
    oracle's."""
    rng = np.random.default_rng(31)
    omm_a = config1["eyes"][0][:240]
    omm_b = config1["eyes"][0][300:557]
    setup_eye(lib, config1, omm=omm_a, spp=32)

    def check(omm, spp=None, counts=None):
        assert lib.renderFrame() >= 0, er.last_error(lib)
        frame = lib.getCurrentEyeFrameIndex() - 1
        ref = oracle_c1.render(omm, er.pose_np(lib), spp=spp or 1, counts=counts, eye=lib.getCurrentEyeIndex(),
                               frame=frame, want_samples=True)
        assert np.array_equal(er.samples_np(lib).view(np.uint32), ref["samples"].view(np.uint32))
        assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])

    check(omm_a, spp=32)
    lib.setCurrentEyeSamplesPerOmmatidium(7)
    check(omm_a, spp=7)
    lib.setCurrentEyeSamplesPerOmmatidium(100)
    check(omm_a, spp=100)
    lib.setCurrentEyeSamplesPerOmmatidium(100)   # unchanged rate
    check(omm_a, spp=100)
    counts = rng.integers(1, 90, size=len(omm_a)).astype(np.uint32)
    er.set_sample_counts(lib, counts)
    check(omm_a, counts=counts)
    er.set_sample_counts(lib, None)
    lib.setCurrentEyeSamplesPerOmmatidium(16)
    check(omm_a, spp=16)
    er.set_ommatidia_np(lib, omm_b)
    check(omm_b, spp=16)
    er.set_ommatidia_np(lib, omm_a)
    lib.setCurrentEyeSamplesPerOmmatidium(65)
    check(omm_a, spp=65)


def test_unrecorded_render_matches(lib, config1, oracle_c1):
    """Default renderFrame path (samples reduced in registers, nothing recorded)."""
    setup_eye(lib, config1, spp=48, record=False)
    assert lib.renderFrame() >= 0, er.last_error(lib)
    ref = oracle_c1.render(config1["eyes"][0], er.pose_np(lib), spp=48)
    assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
    assert er.samples_np(lib).shape[0] == 0
    lib.setSampleRecording(True)


@pytest.mark.parametrize("grab,blocks", [(1, 1), (3, 7), (16, 0), (200, 0)])
def test_launch_shapes_same_result(lib, config1, oracle_c1, grab, blocks):
    """Any launch shape gives the oracle's colours; grabs above 64 units are
    clamped to 64 (a wave keeps one grab's sample counts in its lanes)."""
    omm = config1["eyes"][0][:400]
    setup_eye(lib, config1, omm=omm, spp=24)
    lib.setRenderTuning(grab, blocks)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        ref = oracle_c1.render(omm, er.pose_np(lib), spp=24)
        assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
    finally:
        lib.setRenderTuning(0, 0)


def test_traversal_stats_match_oracle(lib, config1, oracle_c1):
    omm = config1["eyes"][0][:300]
    setup_eye(lib, config1, omm=omm, spp=16)
    st = er.traversal_stats(lib)
    ref = oracle_c1.render(omm, er.pose_np(lib), spp=16)
    assert st["nodes_visited"] == ref["nodes_visited"]
    assert st["tris_tested"] == ref["tris_tested"]


def test_frame_buffer_make_color(lib, config1, oracle_c1):
    setup_eye(lib, config1, spp=16)
    er.setRenderSize(lib, 1000, 1)
    lib.renderFrame()
    frame = np.array(lib.getFramePointer(), copy=True)
    col = er.colours_np(lib)
    assert np.array_equal(frame[0], oracle.make_color(col))
    er.setRenderSize(lib, 400, 400)


def test_frame_counter_and_restart(lib, config1, oracle_c1):
    omm = config1["eyes"][0][:256]
    setup_eye(lib, config1, omm=omm, spp=8)
    pose = er.pose_np(lib)
    lib.renderFrame()
    lib.renderFrame()
    assert lib.getCurrentEyeFrameIndex() == 2
    want = oracle_c1.render(omm, pose, spp=8, frame=1)["colour_gpu_order"][0]
    assert np.array_equal(er.colours_np(lib), want)
    lib.setCurrentEyeSamplesPerOmmatidium(8)   # restarts the streams, as the reference re-seeds
    assert lib.getCurrentEyeFrameIndex() == 0
    lib.renderFrame()
    want0 = oracle_c1.render(omm, pose, spp=8, frame=0)["colour_gpu_order"][0]
    assert np.array_equal(er.colours_np(lib), want0)


def test_uneven_sample_counts(lib, config1, oracle_c1):
    omm = config1["eyes"][0][:500]
    counts = scenes.log_uniform_counts(500, 1, 300, seed=9)
    counts[7] = 0
    counts[8] = 64
    counts[9] = 65
    setup_eye(lib, config1, omm=omm, spp=4)
    er.set_sample_counts(lib, counts)
    assert lib.getCurrentEyeTotalSamples() == int(counts.sum())
    assert lib.renderFrame() >= 0, er.last_error(lib)
    ref = oracle_c1.render(omm, er.pose_np(lib), counts=counts, want_samples=True)
    assert np.array_equal(er.samples_np(lib).view(np.uint32), ref["samples"].view(np.uint32))
    col = er.colours_np(lib)
    assert np.array_equal(col, ref["colour_gpu_order"][0])
    assert np.all(col[7] == 0.0)
    er.set_sample_counts(lib, None)


@pytest.mark.parametrize("grab", [1, 3, 16, 64])
def test_packed_small_units(lib, config1, oracle_c1, grab):
    """Units of at most 32 samples share a wave in power-of-two lane segments;
    every packing (grab sizes, mixed segment sizes, zero counts, a 33-sample
    unit breaking a pack) gives the oracle's colours and samples exactly."""
    omm = config1["eyes"][0][:400]
    rng = np.random.default_rng(grab)
    counts = rng.integers(0, 34, size=400).astype(np.uint32)
    counts[:16] = [1, 1, 2, 3, 4, 5, 8, 16, 17, 32, 33, 0, 7, 31, 9, 1]
    # many zero and one-sample units: packed rounds then often leave a gap of
    # idle lanes before an aligned segment, whose unit must still get its count
    counts[16:24] = [0, 0, 0, 20, 1, 0, 0, 30]
    counts[24:200:3] = 0
    setup_eye(lib, config1, omm=omm, spp=4)
    er.set_sample_counts(lib, counts)
    lib.setRenderTuning(grab, 0)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        ref = oracle_c1.render(omm, er.pose_np(lib), counts=counts, want_samples=True)
        assert np.array_equal(er.samples_np(lib).view(np.uint32), ref["samples"].view(np.uint32))
        assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
        poses6 = scenes.batch_poses(config1["scene"], 5, seed=grab, extent=10.0)
        _, rgb = er.render_batch_np(lib, poses6)
        poses12 = np.stack([oracle.pose12(p) for p in poses6])
        want = oracle_c1.render(omm, poses12, counts=counts, frame=lib.getCurrentEyeFrameIndex())
        assert np.array_equal(rgb, want["colour_gpu_order"])
    finally:
        lib.setRenderTuning(0, 0)
        er.set_sample_counts(lib, None)


def test_merged_eyes_bit_exact(lib, config1, oracle_c1):
    """Copies of an eye moved to world positions metres apart and merged into
    one ommatidium list (units then sort by position cell before direction),
    with counts from 1 to 79 (packed small units and whole-wave units): the
    samples and colours are the oracle's."""
    base = config1["eyes"][0][:120]
    poses = scenes.batch_poses(config1["scene"], 4, seed=31, extent=10.0)
    omm = np.concatenate([scenes.world_ommatidia(base, oracle.pose12(p)) for p in poses]).astype(np.float32)
    assert np.ptp(omm[:, 0]) > 2.0 or np.ptp(omm[:, 2]) > 2.0
    counts = np.random.default_rng(31).integers(1, 80, size=len(omm)).astype(np.uint32)
    setup_eye(lib, config1, omm=omm, spp=4, pose6=np.zeros(6))
    er.set_sample_counts(lib, counts)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        ref = oracle_c1.render(omm, er.pose_np(lib), counts=counts, want_samples=True)
        assert np.array_equal(er.samples_np(lib).view(np.uint32), ref["samples"].view(np.uint32))
        assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
    finally:
        er.set_sample_counts(lib, None)


def test_pose_batch(lib, config1, oracle_c1):
    omm = config1["eyes"][0][:300]
    setup_eye(lib, config1, omm=omm, spp=16)
    poses6 = scenes.batch_poses(config1["scene"], 12, seed=4, extent=10.0)
    ms, rgb = er.render_batch_np(lib, poses6)
    assert ms >= 0
    poses12 = np.stack([oracle.pose12(p) for p in poses6])
    ref = oracle_c1.render(omm, poses12, spp=16, frame=lib.getCurrentEyeFrameIndex())
    assert np.array_equal(rgb, ref["colour_gpu_order"])


def test_autotuned_batch_matches_oracle(lib, config1, oracle_c1, tmp_path):
    """The autotuner times every candidate launch shape; the batch rendered with
    the winning shape (and through the tuning cache) equals the oracle."""
    omm = config1["eyes"][0][:300]
    setup_eye(lib, config1, omm=omm, spp=16)
    counts = scenes.log_uniform_counts(300, 8, 64, seed=3)
    er.set_sample_counts(lib, counts)
    poses6 = scenes.batch_poses(config1["scene"], 10, seed=8, extent=10.0)
    poses12 = np.stack([oracle.pose12(p) for p in poses6])
    ref = oracle_c1.render(omm, poses12, counts=counts, frame=lib.getCurrentEyeFrameIndex())["colour_gpu_order"]
    try:
        rows = er.autotune_np(lib, poses6, repeats=1)
        assert rows.shape == (45, 4) and np.all(rows[:, 3] > 0)
        best = rows[np.argmin(rows[:, 3])]
        tun = er.render_tuning(lib)
        assert tun["units_per_grab"] == int(best[0]) and tun["blocks"] == int(best[1])
        # traversal 0 packet, 1 packet with scalar loads, 2 per-ray -> trace modes 1, 2, 0
        assert tun["trace_mode"] == {0: 1, 1: 2, 2: 0}[int(best[2])]
        _, rgb = er.render_batch_np(lib, poses6)
        assert np.array_equal(rgb, ref)
        # automatic tuning inside renderPoseBatch, saved to and reloaded from the cache file
        cache = tmp_path / "tune.txt"
        assert lib.setAutotuneCacheFile(str(cache).encode()) >= 0
        lib.setBatchAutotune(1)
        _, rgb2 = er.render_batch_np(lib, poses6)
        assert np.array_equal(rgb2, ref)
        lines = [ln for ln in cache.read_text().splitlines() if not ln.startswith("#")]
        assert any(ln.split()[0].startswith("tris19602_") for ln in lines)
        assert lib.setAutotuneCacheFile(str(cache).encode()) >= 1
    finally:
        lib.setBatchAutotune(0)
        lib.setAutotuneCacheFile(None)
        lib.setRenderTuning(0, 0)
        lib.setTraversalMode(2)


def test_spp_edge_values(lib, config1, oracle_c1):
    omm = config1["eyes"][0][:64]
    for spp in (1, 63, 64, 65, 129):
        setup_eye(lib, config1, omm=omm, spp=spp)
        lib.renderFrame()
        want = oracle_c1.render(omm, er.pose_np(lib), spp=spp)["colour_gpu_order"][0]
        assert np.array_equal(er.colours_np(lib), want), spp
    setup_eye(lib, config1, omm=omm, spp=0)
    assert lib.getCurrentEyeSamplesPerOmmatidium() == 1


def test_degenerate_axes(lib, config1, oracle_c1):
    # axis straight up (perp falls back to +z) and axes with x == z (the
    # reference's component-sum test also falls back there)
    omm = np.zeros((4, 8), np.float32)
    omm[0, 3:6] = (0, 1, 0)
    omm[1, 3:6] = (0, -1, 0)
    omm[2, 3:6] = (0.6, 0.52915025, 0.6)
    omm[3, 3:6] = (-0.5, -0.70710677, -0.5)
    omm[:, 6] = 0.05
    setup_eye(lib, config1, omm=omm, spp=64)
    rays = er.rays_np(lib)
    want = oracle.generate_rays(omm, er.pose_np(lib), spp=64, eye=lib.getCurrentEyeIndex())
    assert np.array_equal(rays.view(np.uint32), want.view(np.uint32))
    lib.renderFrame()
    ref = oracle_c1.render(omm, er.pose_np(lib), spp=64)["colour_gpu_order"][0]
    assert np.array_equal(er.colours_np(lib), ref)


def test_empty_and_tiny_scenes(lib, tmp_path):
    rng = np.random.default_rng(3)
    eye = scenes.fibonacci_eye(200, 0.001, 0.05)
    # a scene with no triangles: every sample is sky
    empty = scenes.Scene([], eyes=[scenes.Eye("eye", eye)])
    path = scenes.write_gltf(tmp_path, "empty", empty)
    load(lib, path)
    lib.setCurrentEyeSamplesPerOmmatidium(8)
    lib.setCameraPose(0, 0, 0, 0, 0, 0)
    assert lib.renderFrame() >= 0, er.last_error(lib)
    o = oracle.OracleScene(np.zeros((0, 9), np.float32), np.zeros((0, 9), np.float32), np.zeros(0, np.uint32), 1)
    want = o.render(eye, er.pose_np(lib), spp=8)["colour_gpu_order"][0]
    assert np.array_equal(er.colours_np(lib), want)
    # a single triangle and a two-triangle scene (below the leaf size)
    for ntri in (1, 2, 5):
        pos = rng.uniform(-1, 1, size=(3 * ntri, 3)).astype(np.float32) + np.float32([0, 0, 2])
        idx = np.arange(3 * ntri, dtype=np.uint32).reshape(-1, 3)
        col = np.concatenate([rng.uniform(0, 1, (3 * ntri, 3)), np.ones((3 * ntri, 1))], 1).astype(np.float32)
        sc = scenes.Scene([scenes.Mesh("t", pos, idx, col)], eyes=[scenes.Eye("eye", eye)])
        path = scenes.write_gltf(tmp_path, f"tri{ntri}", sc)
        load(lib, path)
        lib.setCurrentEyeSamplesPerOmmatidium(16)
        lib.setCameraPose(0, 0, 0, 0, 0, 0)
        lib.renderFrame()
        o = oracle.OracleScene.from_scene(sc)
        assert np.array_equal(er.bvh_nodes_np(lib), o.nodes())
        want = o.render(eye, er.pose_np(lib), spp=16)["colour_gpu_order"][0]
        assert np.array_equal(er.colours_np(lib), want), ntri


@pytest.mark.parametrize("shader,variant,ids", [("spherical_positionwise", 0, False),
                                                ("spherical_orientationwise", 1, False),
                                                ("spherical_split_orientationwise", 2, False),
                                                ("spherical_orientationwise_ids", 1, True),
                                                ("spherical_positionwise_ids", 0, True)])
def test_spherical_projections(lib, config1, oracle_c1, shader, variant, ids):
    """Window pixels show the ommatidium closest in angle (oracle ids), as its
    make_color colour or as its index in the RGBA bytes."""
    omm = config1["eyes"][0][:500]
    setup_eye(lib, config1, omm=omm, spp=8, record=False)
    lib.setCurrentEyeShaderName(shader.encode())
    w, h = 72, 36
    er.setRenderSize(lib, w, h)
    try:
        for _ in range(2):   # second frame reuses the cached pixel map
            assert lib.renderFrame() >= 0, er.last_error(lib)
            frame = np.array(lib.getFramePointer(), copy=True).reshape(h, w, 4)
            want_ids = oracle.spherical_ids(omm, w, h, variant)
            if ids:
                enc = np.stack([want_ids >> 24, (want_ids >> 16) & 255, (want_ids >> 8) & 255, want_ids & 255], -1)
                assert np.array_equal(frame, enc.astype(np.uint8))
            else:
                col = oracle.make_color(er.colours_np(lib))
                assert np.array_equal(frame, col[want_ids])
    finally:
        lib.setCurrentEyeShaderName(b"single_dimension_fast")
        er.setRenderSize(lib, 400, 400)


def test_batch_graph_rebuilt_with_scene(lib, config1, oracle_c1, tmp_path):
    """The captured batch graph belongs to one scene build.  Reloading the same
    triangles under the default background (same sizes, so the allocator can
    hand back the same addresses) and rebuilding with another leaf size
    between batches give each new scene's oracle colours: the graph is
    captured again, with the new sky mode and the new BVH's stack bound."""
    omm = config1["eyes"][0][:300]
    poses6 = scenes.batch_poses(config1["scene"], 6, seed=12, extent=10.0)
    poses12 = np.stack([oracle.pose12(p) for p in poses6])
    setup_eye(lib, config1, omm=omm, spp=16, record=False)
    _, rgb = er.render_batch_np(lib, poses6)
    frame = lib.getCurrentEyeFrameIndex()
    assert np.array_equal(rgb, oracle_c1.render(omm, poses12, spp=16, frame=frame)["colour_gpu_order"])
    sc = config1["scene"]
    dflt = scenes.Scene(sc.meshes, eyes=sc.eyes, background="default_background", heightfield=sc.heightfield)
    path = scenes.write_gltf(tmp_path, "config1_default_bg", dflt)
    setup_eye(lib, {"gltf": path, "pose": config1["pose"]}, omm=omm, spp=16, record=False)
    _, rgb0 = er.render_batch_np(lib, poses6)
    o0 = oracle.OracleScene(sc.tri_pos(), sc.tri_col(), sc.tri_mode(), 0, 4)
    want0 = o0.render(omm, poses12, spp=16, frame=lib.getCurrentEyeFrameIndex())["colour_gpu_order"]
    assert not np.array_equal(want0, rgb)
    assert np.array_equal(rgb0, want0)
    try:
        assert lib.setSceneLeafSize(1) == 0
        _, rgb1 = er.render_batch_np(lib, poses6)
        o1 = oracle.OracleScene(sc.tri_pos(), sc.tri_col(), sc.tri_mode(), 0, 1)
        assert er.bvh_info(lib)["stack_bound"] == o1.info()["stack_bound"]
        assert np.array_equal(rgb1, want0)
    finally:
        assert lib.setSceneLeafSize(4) == 0


def test_packet_guard_flags_partial_waves(lib, config1, oracle_c1):
    """The wave-packet traversal keeps its stack in the lanes of the wave, so
    it needs every lane.  Entered by the whole wave it returns the oracle's
    hits and flags nothing; entered by a partial wave (lanes skipping the
    call, forced from the host) it flags the launch and returns misses
    without walking the tree; the flag is cleared and later renders are
    unaffected."""
    setup_eye(lib, config1, spp=32)
    rays = np.ascontiguousarray(er.rays_np(lib)[::500][:64])
    assert rays.shape == (64, 8)
    oids, _ = oracle_c1.trace(rays)
    assert (oids >= 0).any() and (oids < 0).any()
    ids = np.zeros(64, np.int32)
    assert lib.probePacketGuard(rays, 64, ids) == 0, er.last_error(lib)
    assert np.array_equal(ids, oids)
    for active in (1, 32, 63):
        ids[:] = 0
        assert lib.probePacketGuard(rays, active, ids) == 1
        assert np.all(ids[:active] == -1) and np.all(ids[active:] == -2)
    assert lib.probePacketGuard(rays, 64, ids) == 0
    assert np.array_equal(ids, oids)
    assert lib.renderFrame() >= 0, er.last_error(lib)
    want = oracle_c1.render(config1["eyes"][0], er.pose_np(lib), spp=32, frame=lib.getCurrentEyeFrameIndex() - 1)
    assert np.array_equal(er.colours_np(lib), want["colour_gpu_order"][0])


@pytest.mark.parametrize("width,height", [(333, 3), (1500, 2)])
def test_single_dimension_projection(lib, config1, oracle_c1, width, height):
    """single_dimension (shaders.cu:375-389): pixel (x, y) shows the summed
    colour of ommatidium x * n / width on every row."""
    omm = config1["eyes"][0]
    setup_eye(lib, config1, omm=omm, spp=8, record=False)
    lib.setCurrentEyeShaderName(b"single_dimension")
    er.setRenderSize(lib, width, height)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        want = oracle_c1.render(omm, er.pose_np(lib), spp=8, frame=lib.getCurrentEyeFrameIndex() - 1)
        col = want["colour_gpu_order"][0]
        assert np.array_equal(er.colours_np(lib), col)
        frame = np.array(lib.getFramePointer(), copy=True).reshape(height, width, 4)
        idx = (np.arange(width, dtype=np.int64) * len(omm)) // width
        assert np.array_equal(frame, np.broadcast_to(oracle.make_color(col)[idx], (height, width, 4)))
    finally:
        lib.setCurrentEyeShaderName(b"single_dimension_fast")
        er.setRenderSize(lib, 400, 400)


@pytest.mark.parametrize("uneven", [False, True])
def test_raw_ommatidial_samples_projection(lib, config1, oracle_c1, uneven):
    """raw_ommatidial_samples (shaders.cu:354-368): column x is ommatidium x,
    row y its sample y pre-scaled by 1/n (n: that ommatidium's sample count,
    spp when uniform), through make_color.  Rows at or beyond n, and
    columns beyond the ommatidia, are not written."""
    n_omm, spp = 150, 12
    omm = config1["eyes"][0][:n_omm]
    setup_eye(lib, config1, omm=omm, spp=spp, record=False)
    counts = None
    if uneven:
        counts = np.random.default_rng(8).integers(0, 20, size=n_omm).astype(np.uint32)
        counts[:3] = [0, 1, 19]
        er.set_sample_counts(lib, counts)
    lib.setCurrentEyeShaderName(b"raw_ommatidial_samples")
    w, h = 160, 24
    er.setRenderSize(lib, w, h)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        frame_index = lib.getCurrentEyeFrameIndex() - 1
        ref = oracle_c1.render(omm, er.pose_np(lib), spp=spp, counts=counts, frame=frame_index, want_samples=True)
        n = np.full(n_omm, spp, np.int64) if counts is None else counts.astype(np.int64)
        first = np.concatenate([[0], np.cumsum(n)])
        frame = np.array(lib.getFramePointer(), copy=True).reshape(h, w, 4)
        for x in range(n_omm):
            s = ref["samples"][first[x]:first[x + 1], :3]
            inv = np.float32(1.0) / np.float32(n[x])
            assert np.array_equal(frame[:n[x], x], oracle.make_color((s * inv).astype(np.float32))), x
    finally:
        lib.setCurrentEyeShaderName(b"single_dimension_fast")
        er.set_sample_counts(lib, None)
        er.setRenderSize(lib, 400, 400)


def test_packet_counts(lib, config1, oracle_c1):
    """With counting on, renderFrame's packet traversal reports the nodes
    its waves visit and the triangles of the leaves they enter.  Every lane
    walks a superset of its own ray's nodes, so 64 x the packet visits bound
    the rays' own visits (the oracle's per-ray count) from above; with
    counting off nothing is added, and the colours are the same."""
    omm = config1["eyes"][0][:640]
    setup_eye(lib, config1, omm=omm, spp=64, record=False)
    lib.setPacketCounting(False)
    er.packet_counts(lib)
    assert lib.renderFrame() >= 0
    assert er.packet_counts(lib) == {"nodes": 0, "tris": 0}
    lib.setCurrentEyeFrameIndex(0)
    lib.setPacketCounting(True)
    try:
        assert lib.renderFrame() >= 0, er.last_error(lib)
        pk = er.packet_counts(lib)
    finally:
        lib.setPacketCounting(False)
    ref = oracle_c1.render(omm, er.pose_np(lib), spp=64, frame=0)
    assert np.array_equal(er.colours_np(lib), ref["colour_gpu_order"][0])
    assert 0 < pk["nodes"] and ref["nodes_visited"] <= 64 * pk["nodes"]
    assert 0 < pk["tris"] and ref["tris_tested"] <= 64 * pk["tris"]


def test_convergence_against_fp64_oracle(lib, config1, oracle_c1):
    """Converged GPU colours against the FP64-accumulated 4,096-sample oracle:
    the GPU's own 4,096-sample render (FP32 accumulation, the same samples)
    stays within 1e-4 of it, and renders of independent samples approach it
    as N^-1/2, within 1 % RMS from 1,024 samples on."""
    omm = config1["eyes"][0][:120]
    setup_eye(lib, config1, omm=omm, spp=4096, record=False)
    try:
        ref = oracle_c1.render(omm, er.pose_np(lib), spp=4096, want_f64=True, eye=lib.getCurrentEyeIndex(),
                               frame=7)["colour_f64"][0]
        lib.setCurrentEyeFrameIndex(7)
        assert lib.renderFrame() >= 0, er.last_error(lib)
        same = er.colours_np(lib).astype(np.float64)
        assert float(np.max(np.abs(same - ref))) < 1e-4
        errs = {}
        for spp in (64, 256, 1024):
            lib.setCurrentEyeSamplesPerOmmatidium(spp)
            lib.setCurrentEyeFrameIndex(100 + spp)   # samples independent of the reference's
            assert lib.renderFrame() >= 0, er.last_error(lib)
            errs[spp] = float(np.sqrt(np.mean((er.colours_np(lib).astype(np.float64) - ref) ** 2)))
        assert errs[1024] < 0.01, errs
        assert 1.4 < errs[64] / errs[256] < 2.9 and 1.4 < errs[256] / errs[1024] < 2.9, errs
    finally:
        lib.setSampleRecording(True)
