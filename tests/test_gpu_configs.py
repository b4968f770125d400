"""GPU parity at the benchmark workloads, through the C ABI.

The other parity tests run on the 19,602-triangle heightfield.  These run the
exact kernels and shapes the benchmark times, at the benchmark's scenes, and
compare with the oracle at pose and eye counts it finishes in seconds:

- bee scale (2 eyes x 5,500 ommatidia, 128 samples, 2,189,058 triangles):
  BVH bytes, sort order and Morton codes; both eyes through renderFrame with
  sample recording off, in all three traversals (the split-launch
  k_render_tasks for the packet traversals); the uint8 frame row; the
  traversal counters the roofline is computed from;
- the 4,096-pose batch (5,000 ommatidia, counts log-uniform in [8, 256]):
  16 of its poses through renderPoseBatch with the autotuned launch shape,
  captured in a graph and replayed;
- learned adaptive sampling on the bee-scale scene and eyes: pilot + top-up
  colours of both eyes at two poses;
- the large heterogeneous batch (16,316,418 triangles, 256 eyes): BVH bytes,
  sort order and Morton codes, and three eyes of the merged ommatidium list
  rendered uniformly, with uneven counts and adaptively.
"""
import numpy as np
import pytest

import oracle
from compound_ray_amd import eyerenderer as er
from compound_ray_amd import scenes

pytestmark = pytest.mark.gpu


def load(lib, path):
    lib.loadGlTFscene(str(path).encode())
    assert er.last_error(lib) == "", er.last_error(lib)


def compound_eyes(lib):
    out = []
    for i in range(lib.getCameraCount()):
        lib.gotoCamera(i)
        if lib.isCompoundEyeActive():
            out.append(i)
    return out


def check_bvh(lib, osc):
    info, oinfo = er.bvh_info(lib), osc.info()
    for k in ("tris", "nodes", "stack_bound", "levels"):
        assert info[k] == oinfo[k], k
    ids, codes = er.sort_order_np(lib)
    oids, ocodes = osc.order()
    assert np.array_equal(codes, ocodes)
    assert np.array_equal(ids, oids)
    assert np.array_equal(er.bvh_nodes_np(lib), osc.nodes())


def adaptive_want(osc, omm, poses12, counts, eye, frame):
    """Pilot (8 samples) and top-up (samples 8 .. n-1) combined by sample count."""
    pilot = osc.render(omm, poses12, spp=8, eye=eye, frame=frame)["colour_gpu_order"]
    extra_n = (counts - 8).astype(np.uint32)
    extra = osc.render(omm, poses12, counts=extra_n, eye=eye, frame=frame, sample_base=8)["colour_gpu_order"]
    wa = np.float32(8)
    wb = extra_n.astype(np.float32)[..., None]
    inv = np.float32(1) / (wa + wb)
    return np.where(extra_n[..., None] > 0, (wa * pilot + wb * extra) * inv, pilot).astype(np.float32)


# ------------------------------------------------------------- bee scale ----

@pytest.fixture(scope="module")
def config2(tmp_path_factory):
    cfg = scenes.build_config(2, tmp_path_factory.mktemp("config2"))
    cfg["oracle"] = oracle.OracleScene.from_scene(cfg["scene"])
    return cfg


def test_c2_bvh_bit_exact(lib, config2):
    load(lib, config2["gltf"])
    assert er.bvh_info(lib)["tris"] == 2_189_058
    check_bvh(lib, config2["oracle"])


@pytest.mark.parametrize("mode", [1, 2])
def test_c2_long_task_split_bit_exact(lib, config2, mode):
    """renderFrame of one bee-scale eye frame after frame at 128 spp: each
    frame lists its longest tasks and the next frame traces them first, the
    longest as four or two packets of neighbouring-elevation rays, the next
    whole.  Every frame's colours equal the oracle's, with sample recording
    off and on."""
    lib.setTraversalMode(mode)
    lib.setTaskSplit(1)
    load(lib, config2["gltf"])
    eye = compound_eyes(lib)[0]
    er.setRenderSize(lib, 5500, 1)

    def frame(spp, record):
        lib.setSampleRecording(record)
        assert lib.renderFrame() >= 0, er.last_error(lib)
        assert lib.getLastRenderKernel() == b"k_render_tasks"
        want = config2["oracle"].render(config2["eyes"][0], er.pose_np(lib), spp=spp,
                                        eye=lib.getCurrentEyeIndex(),
                                        frame=lib.getCurrentEyeFrameIndex() - 1, want_samples=record)
        assert np.array_equal(er.colours_np(lib), want["colour_gpu_order"][0])
        if record:
            assert np.array_equal(er.samples_np(lib).view(np.uint32), want["samples"].view(np.uint32))
        return lib.getCurrentEyeTaskSplits()

    try:
        lib.gotoCamera(eye)
        lib.setCurrentEyeSamplesPerOmmatidium(128)
        lib.setCameraPose(*[float(v) for v in config2["pose"]])
        splits = [frame(128, record) for record in (False, False, False, True)]
        assert splits[0] == 0 and min(splits[1:]) > 0, splits
        # a change of rate starts over; at 96 spp the second chunk of each
        # ommatidium has 32 rays, and its split packets trace what falls to
        # them of those
        lib.setCurrentEyeSamplesPerOmmatidium(96)
        splits = [frame(96, False) for _ in range(3)]
        assert splits[0] == 0 and min(splits[1:]) > 0, splits
    finally:
        lib.setTaskSplit(2)
        lib.setSampleRecording(False)
        lib.setTraversalMode(2)
        er.setRenderSize(lib, 400, 400)


def test_c2_task_split_trial(lib, config2):
    """The default choice of the long-task split: renders 0-2 of an eye trace
    without it, render 3 records the tasks' steps, renders 4 and 5 trace with
    it, and later renders keep the faster way; every render's colours equal
    the oracle's."""
    lib.setTraversalMode(1)
    lib.setTaskSplit(2)
    load(lib, config2["gltf"])
    eye = compound_eyes(lib)[0]
    er.setRenderSize(lib, 5500, 1)
    try:
        lib.gotoCamera(eye)
        lib.setCurrentEyeSamplesPerOmmatidium(128)
        lib.setCameraPose(*[float(v) for v in config2["pose"]])
        splits = []
        for k in range(8):
            assert lib.renderFrame() >= 0, er.last_error(lib)
            splits.append(lib.getCurrentEyeTaskSplits())
            want = config2["oracle"].render(config2["eyes"][0], er.pose_np(lib), spp=128,
                                            eye=lib.getCurrentEyeIndex(),
                                            frame=lib.getCurrentEyeFrameIndex() - 1)
            assert np.array_equal(er.colours_np(lib), want["colour_gpu_order"][0])
        assert splits[:4] == [0, 0, 0, 0] and min(splits[4:6]) > 0, splits
        # renders 6 and 7 keep one way
        assert (splits[6] > 0) == (splits[7] > 0), splits
        lib.setTaskSplit(3)
        assert er.last_error(lib)
    finally:
        lib.setTaskSplit(2)
        lib.setTraversalMode(2)
        er.setRenderSize(lib, 400, 400)


def test_c2_task_split_many_tasks(lib, config2):
    """More tasks than the split plan holds in one round (a thread keeps 16
    step counts, 16,384 per round): 300 of the bee eye's ommatidia at 4,096
    samples are 19,200 tasks.  With the split on, every frame equals the
    oracle and the later frames list long tasks."""
    lib.setTraversalMode(2)
    lib.setTaskSplit(1)
    load(lib, config2["gltf"])
    eye = compound_eyes(lib)[0]
    omm = np.ascontiguousarray(config2["eyes"][0][::18][:300])
    assert omm.shape[0] == 300
    er.setRenderSize(lib, 300, 1)
    try:
        lib.gotoCamera(eye)
        er.set_ommatidia_np(lib, omm)
        lib.setCurrentEyeSamplesPerOmmatidium(4096)
        lib.setCameraPose(*[float(v) for v in config2["pose"]])
        splits = []
        for k in range(3):
            assert lib.renderFrame() >= 0, er.last_error(lib)
            assert lib.getLastRenderKernel() == b"k_render_tasks"
            splits.append(lib.getCurrentEyeTaskSplits())
            want = config2["oracle"].render(omm, er.pose_np(lib), spp=4096, eye=lib.getCurrentEyeIndex(),
                                            frame=lib.getCurrentEyeFrameIndex() - 1)
            assert np.array_equal(er.colours_np(lib), want["colour_gpu_order"][0]), k
        assert splits[0] == 0 and min(splits[1:]) > 0, splits
    finally:
        lib.setTaskSplit(2)
        er.set_ommatidia_np(lib, config2["eyes"][0])
        er.setRenderSize(lib, 400, 400)


def test_c2_task_split_two_eyes_alternating(lib, config2):
    """Both bee-scale eyes rendered in turn, as the benchmark does, with the
    split always on: each eye keeps its own step counts and lists (the other
    eye's render in between changes neither), and every render of both eyes
    equals the oracle."""
    lib.setTraversalMode(2)
    lib.setTaskSplit(1)
    load(lib, config2["gltf"])
    eyes = compound_eyes(lib)
    er.setRenderSize(lib, 5500, 1)
    try:
        for i in eyes:
            lib.gotoCamera(i)
            lib.setCurrentEyeSamplesPerOmmatidium(128)
            lib.setCameraPose(*[float(v) for v in config2["pose"]])
        splits = {i: [] for i in eyes}
        for k in range(3):
            for n, i in enumerate(eyes):
                lib.gotoCamera(i)
                assert lib.renderFrame() >= 0, er.last_error(lib)
                splits[i].append(lib.getCurrentEyeTaskSplits())
                want = config2["oracle"].render(config2["eyes"][n], er.pose_np(lib), spp=128,
                                                eye=lib.getCurrentEyeIndex(),
                                                frame=lib.getCurrentEyeFrameIndex() - 1)
                assert np.array_equal(er.colours_np(lib), want["colour_gpu_order"][0]), (k, i)
        for i in eyes:
            assert splits[i][0] == 0 and min(splits[i][1:]) > 0, splits
    finally:
        lib.setTaskSplit(2)
        er.setRenderSize(lib, 400, 400)


@pytest.mark.parametrize("mode,kernel", [(1, b"k_render_tasks"), (2, b"k_render_tasks"), (0, b"k_render_chunks")])
def test_c2_render_frame_bit_exact(lib, config2, mode, kernel):
    """Both eyes at 128 samples, recording off: the benchmark's renderFrame."""
    lib.setSampleRecording(False)
    lib.setTraversalMode(mode)
    load(lib, config2["gltf"])
    eyes = compound_eyes(lib)
    assert len(eyes) == 2
    er.setRenderSize(lib, 5500, 1)
    try:
        for k, i in enumerate(eyes):
            lib.gotoCamera(i)
            lib.setCurrentEyeSamplesPerOmmatidium(128)
            lib.setCameraPose(*[float(v) for v in config2["pose"]])
            st = er.traversal_stats(lib) if mode == 1 else None   # counts of the frame about to render
            assert lib.renderFrame() >= 0, er.last_error(lib)
            assert lib.getLastRenderKernel() == kernel
            omm = config2["eyes"][k]
            want = config2["oracle"].render(omm, er.pose_np(lib), spp=128, eye=lib.getCurrentEyeIndex(),
                                            frame=lib.getCurrentEyeFrameIndex() - 1)
            colours = er.colours_np(lib)
            assert np.array_equal(colours, want["colour_gpu_order"][0])
            frame = np.array(lib.getFramePointer(), copy=True)
            assert np.array_equal(frame[0], oracle.make_color(colours))
            if st is not None:
                assert st["nodes_visited"] == want["nodes_visited"]
                assert st["tris_tested"] == want["tris_tested"]
    finally:
        lib.setTraversalMode(2)
        er.setRenderSize(lib, 400, 400)


def test_c4_adaptive_bee_scale_bit_exact(lib, config2):
    """Learned adaptive render of both bee-scale eyes at two test poses (pilot
    packing eight ommatidia per wave, neighbour features, allocation, packed
    top-up) equals the oracle rendering the same counts."""
    lib.setSampleRecording(False)
    load(lib, config2["gltf"])
    eyes = compound_eyes(lib)
    assert lib.mlpInit(1) == 0, er.last_error(lib)
    poses6 = scenes.batch_poses(config2["scene"], 2, seed=4242, extent=40.0)
    poses12 = np.stack([oracle.pose12(p) for p in poses6])
    for k, i in enumerate(eyes):
        lib.gotoCamera(i)
        lib.setCurrentEyeFrameIndex(0)
        _, rgb, counts, total = er.render_adaptive_np(lib, poses6, error_target=0.01, max_spp=256)
        assert counts.min() >= 8 and counts.max() <= 256 and total == int(counts.sum())
        want = adaptive_want(config2["oracle"], config2["eyes"][k], poses12, counts, lib.getCurrentEyeIndex(),
                             lib.getCurrentEyeFrameIndex())
        assert np.array_equal(rgb, want)


# ------------------------------------------------------ 4,096-pose batch ----

def test_c3_batch_autotuned_graph_bit_exact(lib, tmp_path):
    cfg = scenes.build_config(3, tmp_path)
    load(lib, cfg["gltf"])
    lib.gotoCamera(compound_eyes(lib)[0])
    assert lib.getCurrentEyeOmmatidialCount() == 5000
    er.set_sample_counts(lib, cfg["counts"])
    poses6 = cfg["poses"][::256]
    assert len(poses6) == 16
    osc = oracle.OracleScene.from_scene(cfg["scene"])
    try:
        er.autotune_np(lib, poses6, repeats=1)
        want = osc.render(cfg["eyes"][0], np.stack([oracle.pose12(p) for p in poses6]), counts=cfg["counts"],
                          frame=lib.getCurrentEyeFrameIndex())["colour_gpu_order"]
        for _ in range(2):   # capture, then replay
            _, rgb = er.render_batch_np(lib, poses6)
            assert np.array_equal(rgb, want)
    finally:
        lib.setRenderTuning(0, 0)
        lib.setTraversalMode(2)
        er.set_sample_counts(lib, None)


# --------------------------------------------- large heterogeneous batch ----

@pytest.fixture(scope="module")
def config5(tmp_path_factory):
    scene = scenes.config5_scene()
    eyes = scenes.config5_eyes()
    scene.eyes = [scenes.Eye("batch", eyes[0])]
    path = scenes.write_gltf(tmp_path_factory.mktemp("config5"), "config5", scene)
    return {"gltf": path, "scene": scene, "eyes": eyes, "oracle": oracle.OracleScene.from_scene(scene)}


def test_c5_bvh_bit_exact(lib, config5):
    load(lib, config5["gltf"])
    assert er.bvh_info(lib)["tris"] == 16_316_418
    check_bvh(lib, config5["oracle"])


def merged_slice(config5, picks):
    """World-space ommatidia of eyes `picks`, each at its pose of the
    benchmark's test batch, concatenated as the benchmark merges all 256."""
    scene, eyes = config5["scene"], config5["eyes"]
    poses = scenes.batch_poses(scene, len(eyes), seed=902, extent=80.0)
    return np.concatenate([scenes.world_ommatidia(eyes[k], scenes.pose_matrix12(poses[k]))
                           for k in picks]).astype(np.float32)


def test_c5_merged_slice_bit_exact(lib, config5):
    """Three eyes of the merged batch (both layouts, 1k-30k ommatidia each)
    rendered as one ommatidium list: uniform, uneven counts and adaptive."""
    sizes = np.array([len(e) for e in config5["eyes"]])
    even = [k for k in range(0, 32, 2)]
    odd = [k for k in range(1, 32, 2)]
    picks = [min(even, key=lambda k: sizes[k]), min(odd, key=lambda k: sizes[k]), int(np.argsort(sizes)[128])]
    omm = merged_slice(config5, picks)
    osc = config5["oracle"]
    ident6 = np.zeros((1, 6), np.float32)
    lib.setSampleRecording(False)
    load(lib, config5["gltf"])
    lib.gotoCamera(compound_eyes(lib)[0])
    er.set_ommatidia_np(lib, omm)
    lib.setCameraPose(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    pose = er.pose_np(lib)
    eye = lib.getCurrentEyeIndex()
    # uniform, one batch
    lib.setCurrentEyeSamplesPerOmmatidium(24)
    _, rgb = er.render_batch_np(lib, ident6)
    want = osc.render(omm, pose, spp=24, eye=eye, frame=lib.getCurrentEyeFrameIndex())["colour_gpu_order"]
    assert np.array_equal(rgb, want)
    # uneven counts in [4, 512]
    counts = scenes.log_uniform_counts(len(omm), 4, 512, seed=55)
    er.set_sample_counts(lib, counts)
    try:
        _, rgb = er.render_batch_np(lib, ident6)
        want = osc.render(omm, pose, counts=counts, eye=eye, frame=lib.getCurrentEyeFrameIndex())["colour_gpu_order"]
        assert np.array_equal(rgb, want)
    finally:
        er.set_sample_counts(lib, None)
    # adaptive
    assert lib.mlpInit(5) == 0, er.last_error(lib)
    lib.setCurrentEyeFrameIndex(0)
    _, rgb, counts, total = er.render_adaptive_np(lib, ident6, error_target=0.01, max_spp=512)
    assert counts.min() >= 8 and counts.max() <= 512 and total == int(counts.sum())
    want = adaptive_want(osc, omm, pose, counts, eye, lib.getCurrentEyeFrameIndex())
    assert np.array_equal(rgb, want)
