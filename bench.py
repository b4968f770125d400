"""Benchmark: compound-eye rendering on one MI355X.

Headline workload (the large heterogeneous batch): 256 eyes with seeded
ommatidium counts log-uniform in [1k, 30k] (2,149,386 ommatidia), alternating
Fibonacci-sphere and ellipsoidal-patch layouts, acceptance 1-5 deg, each at
its own seeded pose over a 16,316,418-triangle scene (2048^2 heightfield +
20,000 64-segment cylinders + 200 icosphere rocks), generated in-run from
fixed seeds and loaded through loadGlTFscene.  The eyes' ommatidia are moved
to world space and rendered as one batch.  One step is one learned adaptive
render of the whole batch (8-sample pilot, sampler network, Neyman
allocation, top-up: 256 eye-frames) at <= 1 % RMS colour error against a
4,096-sample render; inputs are resident in HBM before the timed region.
The metric is eye-frames/s; Grays/s, the gain over uniform sampling (and over
the same allocation without the network) at the same error, the render
kernel's roofline fraction and a CPU-oracle baseline are reported beside it.

Extra keys: the bee-scale plain method through renderFrame (2 eyes x 5,500
ommatidia x 128 spp, 2,189,058 triangles), learned adaptive sampling on the
bee-scale scene (1,024 poses), and the 4,096-pose batch on the
19,602-triangle heightfield.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import platform
import sys
import tempfile
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

METRIC = "eye-frames/s at ≤1% RMS colour err; Grays/s; learned-adaptive vs uniform"
HBM_PEAK_GBS = 8000.0


PROFILES = ROOT / "profiles"


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` (c2, c5), from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary named for both
    (already scaled by the copy-kernel factors recorded there); None when no
    such summary is present."""
    base = kernel.split("<")[0]
    try:
        for line in (PROFILES / f"{base}_{workload}_traffic.txt").read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(cfg: dict, spp: int, pose12: np.ndarray, target_s: float = 10.0) -> dict:
    """Times the CPU oracle on a bounded sample of the same workload (eye 0)."""
    import oracle  # the CPU oracle, timed here as the CPU baseline

    t0 = time.time()
    osc = oracle.OracleScene.from_scene(cfg["scene"])
    build_s = time.time() - t0
    omm = cfg["eyes"][0]
    threads = oracle.default_threads()
    m = 64
    t0 = time.perf_counter()
    osc.render(omm[:m], pose12, spp=spp, threads=threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    m = int(min(len(omm), max(64, m * target_s / dt)))
    # whole frames of the first m ommatidia (consecutive frame indices) until
    # about target_s seconds of CPU work
    frames = 0
    t0 = time.perf_counter()
    while True:
        osc.render(omm[:m], pose12, spp=spp, frame=frames, threads=threads)
        frames += 1
        dt = time.perf_counter() - t0
        if dt >= target_s or frames >= 1000:
            break
    rays = m * spp * frames
    rays_per_s = rays / dt
    eye_frame_rays = len(omm) * spp
    cpu_model = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": rays_per_s / eye_frame_rays,
        "unit": "eye-frames/s",
        "cores": threads,
        "kind": "port",
        "sample": f"eye 0, first {m} of {len(omm)} ommatidia x {spp} spp x {frames} frames ({rays} rays) in "
                  f"{dt:.2f} s; "
                  f"{rays_per_s / 1e6:.3f} Mrays/s; oracle BVH build {build_s:.1f} s excluded; CPU {cpu_model}",
        "grays_per_s": rays_per_s / 1e9,
    }


def bench_batched(lib, er, scenes, tmp: str, repeats: int = 3) -> dict:
    """Batched configuration: 4,096 seeded poses of one 5,000-ommatidium eye
    over the 19,602-triangle heightfield, per-ommatidium samples log-uniform in
    [8, 256], one persistent graph-launched render per batch."""
    cfg = scenes.build_config(3, os.path.join(tmp, "c3"))
    lib.loadGlTFscene(str(cfg["gltf"]).encode())
    if er.last_error(lib):
        raise RuntimeError(er.last_error(lib))
    er.gotoFirstCompoundEye(lib)
    er.set_sample_counts(lib, cfg["counts"])
    poses = cfg["poses"]
    rays = int(cfg["counts"].sum()) * len(poses)
    # launch shape (units per queue grab, blocks, node-load variant) from the autotuner
    t0 = time.perf_counter()
    trials = er.autotune_np(lib, poses, repeats=1)
    tune_s = time.perf_counter() - t0
    tuning = er.render_tuning(lib)
    default_ms = float(trials[(trials[:, 0] == 4) & (trials[:, 2] == 0)][:, 3].min())
    log(f"autotune: {len(trials)} shapes in {tune_s:.1f} s -> {tuning}; sample ms best "
        f"{trials[:, 3].min():.2f}, worst {trials[:, 3].max():.2f}")
    er.render_batch_np(lib, poses, want_rgb=False)  # warm-up (captures the graph)
    times = []
    t0 = time.perf_counter()
    for _ in range(repeats):
        ms, _ = er.render_batch_np(lib, poses, want_rgb=False)
        times.append(ms)
    wall = (time.perf_counter() - t0) / repeats
    gpu_ms = float(np.mean(times))
    return {
        "workload": "4,096 poses x 5,000 ommatidia, spp log-uniform [8, 256] "
                    f"({rays} rays/batch), 19,602-triangle heightfield",
        "eye_frames_per_s": len(poses) / (wall),
        "grays_per_s": rays / wall / 1e9,
        "gpu_ms_per_batch": gpu_ms,
        "wall_ms_per_batch": wall * 1e3,
        "eye_frames_per_s_gpu": len(poses) / (gpu_ms * 1e-3),
        "autotune": {"shape": tuning, "candidates": int(len(trials)), "seconds": tune_s,
                     "sample_ms_best": float(trials[:, 3].min()), "sample_ms_worst": float(trials[:, 3].max()),
                     "sample_ms_grab4_packet": default_ms},
    }


def rms_abs(a, b) -> float:
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))


def calibrate_budget(err_at, target: float, iters: int = 6):
    """Budget scale of the adaptive render whose RMS error is at most `target`
    and at least 97 % of it (error ~ 1/sqrt(budget)), the way the uniform rate
    is set: returns (scale, error, extra) of the smallest scale tried that met
    the target, or of the last one tried when none did."""
    scale, best, last = 1.0, None, None
    for _ in range(iters):
        e, extra = err_at(scale)
        last = (scale, e, extra)
        if e <= target and (best is None or scale < best[0]):
            best = last
        if 0.97 * target <= e <= target:
            break
        scale *= (e / target) ** 2 * (1.03 if e > target else 1.0)
    return best if best is not None else last


def neyman_mean_spp(nreq, floor: float = 1.0, cap: float = 4096.0) -> float:
    """Smallest mean sample count of any allocation n_i in [floor, cap] that
    meets the RMS target, given each unit's required samples nreq_i (its
    variance over the squared target): the mean squared error is
    mean(nreq_i / n_i) <= 1, minimised in total by n_i proportional to
    sqrt(nreq_i) (Neyman) and clipped to the bounds; the multiplier is found
    by bisection."""
    n = np.asarray(nreq, np.float64)
    lo, hi = 1e-9, 1e9
    for _ in range(200):
        lam = math.sqrt(lo * hi)
        if np.mean(n / np.clip(lam * np.sqrt(n), floor, cap)) <= 1.0:
            hi = lam
        else:
            lo = lam
    return float(np.clip(hi * np.sqrt(n), floor, cap).mean())


def bench_adaptive(lib, er, scenes, cfg, eyes, target: float = 0.01, n_test: int = 1024, n_train: int = 8192,
                   train_units: int = 2_000_000, label_spp: int = 1024, n_ref: int = 16) -> dict:
    """Learned adaptive sampling on the bee-scale scene and eyes: the sampler
    network is trained on renders of held-out poses, then 1,024 test poses of
    both eyes are rendered (a) adaptively and (b) uniformly at the smallest
    rate reaching the same RMS colour error against a 4,096-spp render, in
    one batched launch each; (c) uniformly through renderFrame, pose by pose
    (the reference's way of rendering)."""
    scene = cfg["scene"]
    test = scenes.batch_poses(scene, n_test, seed=4242, extent=40.0)
    train = scenes.batch_poses(scene, n_train, seed=777, extent=40.0)
    ref_poses = test[:n_ref]
    # 1. training set: pilot features + labels from label_spp further samples
    t0 = time.perf_counter()
    rng = np.random.default_rng(9)
    feats, labels = [], []
    chunk = 512
    for i in eyes:
        lib.gotoCamera(i)
        n_omm = lib.getCurrentEyeOmmatidialCount()
        frac = train_units / len(eyes) / (n_train * n_omm)
        for c0 in range(0, n_train, chunk):
            P = train[c0:c0 + chunk]
            mask = (rng.random((len(P), n_omm)) < frac).astype(np.uint8)
            lib.setCurrentEyeFrameIndex(100000 + c0)
            f, lab = er.training_set_np(lib, P, unit_mask=mask, pilot=8, label_spp=label_spp, error_target=target)
            feats.append(f)
            labels.append(lab)
    F = np.concatenate(feats)
    L = np.concatenate(labels)
    t_set = time.perf_counter() - t0
    log(f"adaptive: training set {len(L)} rows in {t_set:.1f} s")
    # 2. training on the GPU (BF16 MFMA forward/backward, Adam)
    if lib.mlpInit(1) != 0:
        raise RuntimeError(er.last_error(lib))
    t0 = time.perf_counter()
    losses = er.mlp_train_np(lib, F, L, epochs=6, batch=4096, lr=2e-3, seed=2)
    t_train = time.perf_counter() - t0
    log(f"adaptive: trained in {t_train:.1f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}")

    def per_eye(fn):
        out = []
        for i in eyes:
            lib.gotoCamera(i)
            out.append(fn())
        return out

    # 3. 4,096-spp reference at the first n_ref test poses (independent frame)
    def ref_render():
        lib.setCurrentEyeSamplesPerOmmatidium(4096)
        lib.setCurrentEyeFrameIndex(1 << 20)
        return er.render_batch_np(lib, ref_poses)[1]
    ref = per_eye(ref_render)

    def uniform_err(spp):
        def f():
            lib.setCurrentEyeSamplesPerOmmatidium(spp)
            lib.setCurrentEyeFrameIndex(0)
            return er.render_batch_np(lib, ref_poses)[1]
        return rms_abs(np.stack(per_eye(f)), np.stack(ref))

    # uniform rate for the target: RMS error falls as 1/sqrt(spp)
    e64 = uniform_err(64)
    spp_u = max(1, int(math.ceil(64 * (e64 / target) ** 2)))
    e_u = uniform_err(spp_u)
    while e_u > target and spp_u < 4096:
        spp_u = int(math.ceil(spp_u * 1.1))
        e_u = uniform_err(spp_u)

    # launch shape of both timed renders, autotuned on the uniform batch: the
    # traversal and blocks apply to both; the adaptive passes pick their grab size
    lib.gotoCamera(eyes[0])
    lib.setCurrentEyeSamplesPerOmmatidium(spp_u)
    er.autotune_np(lib, test[:256], repeats=1)
    shape = er.render_tuning(lib)

    # adaptive error on the reference poses, budget calibrated to the target
    def adaptive_err(scale):
        got, spp_sum, units = [], 0, 0
        for i in eyes:
            lib.gotoCamera(i)
            lib.setCurrentEyeFrameIndex(0)
            _, rgb, _, total = er.render_adaptive_np(lib, ref_poses, error_target=target, max_spp=1024,
                                                     budget_scale=scale)
            got.append(rgb)
            spp_sum += total
            units += rgb.shape[0] * rgb.shape[1]
        return rms_abs(np.stack(got), np.stack(ref)), spp_sum / units

    # the same allocation without the network (pooled pilot deviation)
    lib.setAdaptiveSampler(1)
    scale_nf, e_nf, spp_nf_ref = calibrate_budget(adaptive_err, target)
    lib.setAdaptiveSampler(0)
    scale, e_a, spp_a_ref = calibrate_budget(adaptive_err, target)

    # 4. timed: 1,024 poses x both eyes, adaptive vs uniform batched vs renderFrame
    def timed(fn, repeats=2):
        fn()  # warm-up
        t = time.perf_counter()
        for _ in range(repeats):
            fn()
        return (time.perf_counter() - t) / repeats

    total_a = [0]

    def run_adaptive():
        total_a[0] = 0
        for i in eyes:
            lib.gotoCamera(i)
            lib.setCurrentEyeFrameIndex(0)
            total_a[0] += er.render_adaptive_np(lib, test, error_target=target, max_spp=1024, budget_scale=scale,
                                                want_rgb=False)[3]

    def run_network_free():
        lib.setAdaptiveSampler(1)
        try:
            for i in eyes:
                lib.gotoCamera(i)
                lib.setCurrentEyeFrameIndex(0)
                er.render_adaptive_np(lib, test, error_target=target, max_spp=1024, budget_scale=scale_nf,
                                      want_rgb=False)
        finally:
            lib.setAdaptiveSampler(0)

    def run_uniform():
        for i in eyes:
            lib.gotoCamera(i)
            lib.setCurrentEyeSamplesPerOmmatidium(spp_u)
            er.render_batch_np(lib, test, want_rgb=False)

    n_plain = 128

    def run_plain():
        for i in eyes:
            lib.gotoCamera(i)
            lib.setCurrentEyeSamplesPerOmmatidium(spp_u)
            for p in test[:n_plain]:
                lib.setCameraPose(*[float(v) for v in p])
                if lib.renderFrame() < 0:
                    raise RuntimeError(er.last_error(lib))

    frames = n_test * len(eyes)
    t_a = timed(run_adaptive)
    t_u = timed(run_uniform)
    t_nf = timed(run_network_free)
    t_p = timed(run_plain, repeats=1)
    n_omm_total = sum(per_eye(lambda: lib.getCurrentEyeOmmatidialCount()))
    efs_a = frames / t_a
    efs_u = frames / t_u
    efs_p = n_plain * len(eyes) / t_p
    return {
        "workload": f"learned adaptive vs uniform at <= {target} RMS colour error: 2 eyes x 5,500 ommatidia x "
                    f"{n_test} poses over the bee-scale scene",
        "eye_frames_per_s_adaptive": efs_a,
        "eye_frames_per_s_uniform_batched": efs_u,
        "eye_frames_per_s_uniform_renderframe": efs_p,
        "gain_vs_uniform_batched": efs_a / efs_u,
        "gain_vs_uniform_renderframe": efs_a / efs_p,
        "eye_frames_per_s_network_free": frames / t_nf,
        "gain_vs_network_free": t_nf / t_a,
        "rms_abs_network_free": e_nf,
        "network_free_mean_spp_ref_poses": spp_nf_ref,
        "adaptive_mean_spp_ref_poses": spp_a_ref,
        "uniform_spp": spp_u,
        "adaptive_mean_spp": total_a[0] / (n_test * n_omm_total),
        "rms_abs_uniform": e_u,
        "rms_abs_adaptive": e_a,
        "rms_reference": f"{n_ref} test poses x both eyes, uniform 4,096-spp render on the GPU",
        "budget_scale": scale,
        "launch_shape": shape,
        "training": {"rows": int(len(L)), "label_spp": label_spp, "poses": n_train, "set_s": t_set,
                     "train_s": t_train, "epochs": len(losses), "loss_first": float(losses[0]),
                     "loss_last": float(losses[-1])},
    }


def cpu_model() -> str:
    name = platform.processor() or "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return name


def node_bytes(lib, er) -> float:
    """Bytes of one BVH node (read whole at every visit) of the loaded scene."""
    return float(er.bvh_info(lib)["node_bytes"])


def traversal_bytes(st: dict, units: int, per_unit: int, nbytes: float) -> float:
    """Algorithmic bytes of traced rays: BVH nodes visited x `nbytes` + triangles
    tested x 48 B + shading fetch (3 x 16 B) per hit, plus `per_unit` bytes of
    per-unit records (ommatidium read, colour and statistics written)."""
    return st["nodes_visited"] * nbytes + st["tris_tested"] * 48.0 + st["hits"] * 48.0 + units * float(per_unit)


def bench_large(lib, er, scenes, tmp: str, steps: int, warmup: int, target: float = 0.01,
                ref_per_eye: int = 400, cpu_per_eye: int = 20, cpu: bool = True, profile: bool = False) -> dict:
    """The headline: 256 eyes (log-uniform 1k-30k ommatidia, Fibonacci-sphere
    and ellipsoidal layouts, acceptance 1-5 deg), each at its own pose over a
    16.3M-triangle scene, rendered as one batch (all eyes' ommatidia moved to
    world space and traced in one launch sequence).  The sampler network is
    trained on a held-out batch; the error is measured on ref_per_eye seeded
    random ommatidia of every eye against a 4,096-sample render, and on the
    first cpu_per_eye of them against the CPU oracle's FP64-accumulated
    4,096-sample render, which is also the timed CPU baseline."""
    t0 = time.perf_counter()
    scene = scenes.config5_scene()
    eyes = scenes.config5_eyes()
    scene.eyes = [scenes.Eye("batch", eyes[0])]
    path = scenes.write_gltf(os.path.join(tmp, "c5"), "config5", scene)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    lib.loadGlTFscene(str(path).encode())
    if er.last_error(lib):
        raise RuntimeError(er.last_error(lib))
    bvh = er.bvh_info(lib)
    load_s = time.perf_counter() - t0
    log(f"C5: {scene.tri_count()} triangles generated in {gen_s:.1f} s, loaded + BVH in {load_s:.1f} s "
        f"(build {bvh['build_ms']:.1f} ms)")
    er.gotoFirstCompoundEye(lib)
    lib.setSampleRecording(False)
    sizes = np.array([len(e) for e in eyes])

    def world_batch(seed):
        poses = scenes.batch_poses(scene, len(eyes), seed=seed, extent=80.0)
        return np.concatenate([scenes.world_ommatidia(e, scenes.pose_matrix12(p)) for e, p in zip(eyes, poses)])

    identity = np.zeros((1, 6), np.float32)
    # 1. training on a held-out batch (other poses), a quarter of its units
    t0 = time.perf_counter()
    train = world_batch(seed=901)
    er.set_ommatidia_np(lib, train)
    lib.setCameraPose(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    rng = np.random.default_rng(17)
    mask = (rng.random((1, len(train))) < 0.25).astype(np.uint8)
    lib.setCurrentEyeFrameIndex(77)
    F, L = er.training_set_np(lib, identity, unit_mask=mask, pilot=8, label_spp=1024, error_target=target)
    if lib.mlpInit(5) != 0:
        raise RuntimeError(er.last_error(lib))
    losses = er.mlp_train_np(lib, F, L, epochs=6, batch=4096, lr=2e-3, seed=6)
    train_s = time.perf_counter() - t0
    log(f"C5: trained on {len(L)} rows in {train_s:.1f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    # 2. the test batch and its error sample: ref_per_eye random ommatidia of
    # every eye, the first cpu_per_eye of each eye placed first
    batch = world_batch(seed=902)
    n_omm = len(batch)
    first = np.concatenate([[0], np.cumsum(sizes)])
    rng = np.random.default_rng(903)
    picks = [first[k] + rng.choice(sizes[k], size=min(ref_per_eye, sizes[k]), replace=False)
             for k in range(len(eyes))]
    ref_idx = np.concatenate([pk[:cpu_per_eye] for pk in picks] + [pk[cpu_per_eye:] for pk in picks])
    n_cpu = cpu_per_eye * len(eyes)
    er.set_ommatidia_np(lib, batch[ref_idx])
    lib.setCameraPose(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    ref_eye = lib.getCurrentEyeIndex()
    ref_frame = 1 << 20
    lib.setCurrentEyeSamplesPerOmmatidium(4096)
    lib.setCurrentEyeFrameIndex(ref_frame)
    ref = er.render_batch_np(lib, identity)[1][0]
    er.set_ommatidia_np(lib, batch)
    lib.setCameraPose(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    pose12 = er.pose_np(lib)

    def uniform(spp, want_rgb=True):
        lib.setCurrentEyeSamplesPerOmmatidium(spp)
        lib.setCurrentEyeFrameIndex(0)
        return er.render_batch_np(lib, identity, want_rgb=want_rgb)

    def err_of(rgb):
        return rms_abs(rgb[0, ref_idx], ref)

    # 3. the uniform rate reaching the target (error ~ 1/sqrt(spp))
    e64 = err_of(uniform(64)[1])
    spp_u = max(1, int(math.ceil(64 * (e64 / target) ** 2)))
    rgb_u = uniform(spp_u)[1]
    e_u = err_of(rgb_u)
    while e_u > target and spp_u < 4096:
        spp_u = int(math.ceil(spp_u * 1.1))
        rgb_u = uniform(spp_u)[1]
        e_u = err_of(rgb_u)
    # launch shape of every timed render, autotuned on the uniform batch (the
    # adaptive passes keep its traversal and blocks and choose their grab size)
    lib.setCurrentEyeSamplesPerOmmatidium(spp_u)
    er.autotune_np(lib, identity, repeats=1)
    shape = er.render_tuning(lib)

    def adaptive(scale, want_rgb=True):
        lib.setCurrentEyeFrameIndex(0)
        return er.render_adaptive_np(lib, identity, error_target=target, max_spp=512, budget_scale=scale,
                                     want_rgb=want_rgb)

    def calibrated(sampler):
        if lib.setAdaptiveSampler(sampler) != 0:
            raise RuntimeError(er.last_error(lib))

        def err_at(scale):
            _, rgb, _, total = adaptive(scale)
            return err_of(rgb), total
        scale, e, total = calibrate_budget(err_at, target)
        _, rgb, counts, _ = adaptive(scale)
        return scale, e, total, rgb, counts

    scale_nf, e_nf, total_nf, rgb_nf, _ = calibrated(1)
    scale, e_a, total, rgb_a, counts_a = calibrated(0)
    # 4. the timed headline: warmup + steps learned adaptive renders of the batch
    for _ in range(warmup):
        adaptive(scale, want_rgb=False)
    stage = []
    t0 = time.perf_counter()
    for _ in range(steps):
        adaptive(scale, want_rgb=False)
        stage.append(er.adaptive_timings_np(lib))
    elapsed = time.perf_counter() - t0
    stage = np.array(stage)
    if profile:
        # profiler runs: the timed adaptive renders are the last launches
        return {"_timed": {"elapsed": elapsed, "steps": steps, "frames_per_step": len(eyes),
                           "samples_per_step": total},
                "workload": "profile run", "bvh": bvh, "grays_per_s_adaptive": total / (elapsed / steps) / 1e9,
                "gain_vs_uniform_batched": None, "gain_vs_network_free": None, "rms_abs_adaptive": e_a,
                "stage_ms": {"pilot": float(stage[:, 0].mean()), "allocation": float(stage[:, 1].mean()),
                             "top_up": float(stage[:, 2].mean()), "combine": float(stage[:, 3].mean())},
                "roofline": None}

    def timed(fn, repeats=3):
        fn()
        t = time.perf_counter()
        for _ in range(repeats):
            fn()
        return (time.perf_counter() - t) / repeats

    t_u = timed(lambda: uniform(spp_u, want_rgb=False))
    lib.setAdaptiveSampler(1)
    t_nf = timed(lambda: adaptive(scale_nf, want_rgb=False))
    lib.setAdaptiveSampler(0)
    frames = len(eyes)
    t_a = elapsed / steps
    # 5. roofline of the render kernel: the exact traversal counts of the rays
    # the two k_render_units launches trace (pilot: samples 0-7 of every unit;
    # top-up: samples 8 .. n-1), counted by the per-ray statistics pass
    er.set_sample_counts(lib, counts_a[0])
    lib.setCurrentEyeFrameIndex(0)
    st_all = er.traversal_stats(lib)
    er.set_sample_counts(lib, None)
    lib.setCurrentEyeSamplesPerOmmatidium(8)
    lib.setCurrentEyeFrameIndex(0)
    st_pilot = er.traversal_stats(lib)
    st_top = {k: st_all[k] - st_pilot[k] for k in st_all}
    # pilot: ommatidium 32 B read, colour 16 B + statistics 32 B written per
    # unit; top-up: ommatidium 32 B + count 4 B read, colour 16 B written
    unit_io = n_omm * ((32 + 16 + 32) + (32 + 4 + 16))
    nb = node_bytes(lib, er)
    b_per_ray = traversal_bytes(st_pilot, 0, 0, nb) + traversal_bytes(st_top, 0, 0, nb) + unit_io
    # what the waves fetch: a packet reads each node it visits and the
    # triangles of each leaf it enters once for its 64 rays (counted on the
    # GPU in an untimed run of the same render); shading gathers stay per ray
    lib.setPacketCounting(True)
    er.packet_counts(lib)
    adaptive(scale, want_rgb=False)
    pk = er.packet_counts(lib)
    lib.setPacketCounting(False)
    b_packet = pk["nodes"] * nb + pk["tris"] * 48.0 + st_all["hits"] * 48.0 + unit_io
    ms_pilot = float(stage[:, 0].mean())
    ms_top = float(stage[:, 2].mean())
    launch_ms = (ms_pilot + ms_top) / 2
    bytes_launch = b_packet / 2
    achieved = bytes_launch / (launch_ms * 1e-3) / 1e9
    traffic = measured_traffic("k_render_units", "c5")
    # the most an allocation could save with exact knowledge of every unit's
    # variance: required samples of a seeded tenth of the test batch's units
    # from 1,024 samples each, against the uniform rate they imply
    lib.setCurrentEyeFrameIndex(99)
    bmask = (np.random.default_rng(905).random((1, n_omm)) < 0.1).astype(np.uint8)
    _, blab = er.training_set_np(lib, identity, unit_mask=bmask, pilot=8, label_spp=1024, error_target=target)
    nreq = np.exp2(blab.astype(np.float64))
    bound = {"units": int(len(nreq)), "uniform_spp": float(nreq.mean()),
             "ideal_mean_spp": neyman_mean_spp(nreq), "ideal_mean_spp_pilot_floor": neyman_mean_spp(nreq, floor=8.0)}
    bound["ideal_sample_ratio"] = bound["uniform_spp"] / bound["ideal_mean_spp"]
    bound["ideal_sample_ratio_pilot_floor"] = bound["uniform_spp"] / bound["ideal_mean_spp_pilot_floor"]
    out = {
        "workload": f"{len(eyes)} eyes ({n_omm} ommatidia, log-uniform 1k-30k per eye), each at its own pose, "
                    f"{scene.tri_count()} triangles (2048^2 heightfield + 20,000 cylinders + 200 rocks), "
                    "one batch per render",
        "eye_frames_per_s_adaptive": frames / t_a,
        "eye_frames_per_s_uniform_batched": frames / t_u,
        "eye_frames_per_s_network_free": frames / t_nf,
        "gain_vs_uniform_batched": t_u / t_a,
        "gain_vs_network_free": t_nf / t_a,
        "grays_per_s_adaptive": total / t_a / 1e9,
        "grays_per_s_uniform": n_omm * spp_u / t_u / 1e9,
        "uniform_spp": spp_u,
        "adaptive_mean_spp": total / n_omm,
        "network_free_mean_spp": total_nf / n_omm,
        "rms_abs_uniform": e_u,
        "rms_abs_adaptive": e_a,
        "rms_abs_network_free": e_nf,
        "rms_rel_adaptive": e_a / float(np.sqrt(np.mean(ref.astype(np.float64) ** 2))),
        "rms_reference": f"{len(ref_idx)} ommatidia ({ref_per_eye} seeded random per eye, all 256 eyes), "
                         "4,096-spp render on the GPU",
        "budget_scale": scale,
        "budget_scale_network_free": scale_nf,
        "stage_ms": {"pilot": ms_pilot, "allocation": float(stage[:, 1].mean()), "top_up": ms_top,
                     "combine": float(stage[:, 3].mean())},
        "traversal_per_ray": {"nodes": st_all["nodes_visited"] / total, "tris": st_all["tris_tested"] / total,
                              "hit_fraction": st_all["hits"] / total},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_render_units",
                     "bytes_per_launch": bytes_launch, "launch_ms": launch_ms,
                     "launches": "pilot and top-up of one adaptive render (HIP events on the library stream)",
                     "bytes": f"per wave: packet node visits x {nb:.0f} B (one node) + leaf triangles x 48 B; per hit "
                              "48 B of shading data; per unit 132 B of ommatidium, count, colour and statistics",
                     "packet_nodes_per_64_rays": pk["nodes"] * 64.0 / total,
                     "packet_tris_per_64_rays": pk["tris"] * 64.0 / total,
                     "per_ray_equivalent_GBps": b_per_ray / 2 / (launch_ms * 1e-3) / 1e9,
                     "hbm_frac_measured": None if traffic is None else traffic / (launch_ms * 1e-3) / 1e9 /
                     HBM_PEAK_GBS},
        "sampling_bound": bound,
        "launch_shape": shape,
        "bvh": bvh,
        "scene_load_s": load_s,
        "training": {"rows": int(len(L)), "label_spp": 1024, "seconds": train_s, "loss_first": float(losses[0]),
                     "loss_last": float(losses[-1])},
        "_timed": {"elapsed": elapsed, "steps": steps, "frames_per_step": frames, "samples_per_step": total},
    }
    if cpu:
        import oracle  # the CPU oracle: timed here as the CPU baseline, and the FP64 reference
        t0 = time.perf_counter()
        osc = oracle.OracleScene.from_scene(scene)
        build_s = time.perf_counter() - t0
        sample = batch[ref_idx[:n_cpu]]
        threads = oracle.default_threads()
        t0 = time.perf_counter()
        r = osc.render(sample, pose12, spp=4096, eye=ref_eye, frame=ref_frame, want_f64=True, threads=threads)
        dt = time.perf_counter() - t0
        rays = n_cpu * 4096
        f64 = r["colour_f64"][0]
        sel = ref_idx[:n_cpu]
        # the GPU's 4,096-sample reference renders the same rays in FP32
        out["rms_vs_fp64_oracle"] = {
            "ommatidia": n_cpu,
            "adaptive": rms_abs(rgb_a[0, sel], f64),
            "uniform": rms_abs(rgb_u[0, sel], f64),
            "network_free": rms_abs(rgb_nf[0, sel], f64),
            "gpu_4096_reference_max_abs_diff": float(np.max(np.abs(ref[:n_cpu].astype(np.float64) - f64))),
            "gpu_4096_reference_bit_identical_to_oracle_fp32": bool(np.array_equal(ref[:n_cpu],
                                                                                  r["colour_gpu_order"][0])),
        }
        rays_per_s = rays / dt
        out["cpu_baseline"] = {
            "value": rays_per_s / (n_omm / len(eyes) * spp_u),
            "unit": "eye-frames/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{n_cpu} ommatidia ({cpu_per_eye} per eye, all 256 eyes) x 4096 spp ({rays} rays, "
                      f"FP64 accumulation) in {dt:.2f} s = {rays_per_s / 1e6:.3f} Mrays/s, scaled to eye-frames "
                      f"of the plain method at the same error ({n_omm / len(eyes):.0f} ommatidia per eye x "
                      f"{spp_u} spp); oracle BVH build {build_s:.1f} s excluded; CPU {cpu_model()}",
            "grays_per_s": rays_per_s / 1e9,
        }
    return out


def bench_speed_test(lib, er, tmp: str, frames: int = 200, spp_list=(1, 2, 4, 8, 16, 32, 64, 128, 256, 512,
                                                                        1000, 2000, 3200)) -> dict:
    """The reference's speed test (python-examples/speed-test/speedTest.py:
    81-124) on the scene it ships (data/natural-standin-sky.gltf; its lab and
    natural speed-test scenes are not shipped): the 1,000-ommatidium
    equidistant eye through the first compound camera, single_dimension_fast
    into a 1000 x 1 frame, a warm-up, then for each sample count two untimed
    frames and the mean of renderFrame()'s host wall milliseconds over
    `frames` frames (the reference averages 500)."""
    import gzip
    golden = ROOT / "tests" / "golden"
    d = pathlib.Path(tmp) / "speedtest"
    (d / "eyes").mkdir(parents=True, exist_ok=True)
    gltf = d / "natural-standin-sky.gltf"
    gltf.write_bytes(gzip.decompress((golden / "natural_standin_sky.gltf.gz").read_bytes()))
    eye = gzip.decompress((golden / "standin_eye.eye.gz").read_bytes())
    for where in (d, d / "eyes"):
        (where / "1000-horizontallyAcute-variableDegree.eye").write_bytes(eye)
    rows = [[float(t) for t in ln.split()] for ln in
            gzip.decompress((golden / "equidistant_eye.eye.gz").read_bytes()).decode().splitlines() if ln.strip()]
    lib.loadGlTFscene(str(gltf).encode())
    if er.last_error(lib):
        raise RuntimeError(er.last_error(lib))
    er.gotoFirstCompoundEye(lib)
    er.set_ommatidia_np(lib, np.asarray(rows, np.float32))
    lib.setCurrentEyeShaderName(b"single_dimension_fast")
    er.setRenderSize(lib, lib.getCurrentEyeOmmatidialCount(), 1)
    lib.setCurrentEyeSamplesPerOmmatidium(1)
    t0 = time.time()
    while time.time() - t0 < 1.0:
        lib.renderFrame()
    out = {}
    for spp in spp_list:
        lib.setCurrentEyeSamplesPerOmmatidium(spp)
        lib.renderFrame()
        lib.renderFrame()
        total = 0.0
        for _ in range(frames):
            ms = lib.renderFrame()
            if ms < 0:
                raise RuntimeError(er.last_error(lib))
            total += ms
        out[str(spp)] = total / frames
    n = lib.getCurrentEyeOmmatidialCount()
    return {
        "workload": "reference speed test: natural-standin-sky.gltf (24,200 triangles, textured), 1,000-ommatidium "
                    f"equidistant eye, single_dimension_fast 1000x1, mean renderFrame() host ms over {frames} frames",
        "ms_per_frame": out,
        "grays_per_s": {k: n * int(k) / v / 1e6 for k, v in out.items()},
        "reference_rtx_2080_ti_natural_ms": {"32": 0.2096, "128": 0.2576, "1000": 0.6804, "3200": 1.7482},
        "reference_gtx_1080_ti_natural_ms": {"1000": 4.3611, "3200": 13.5052},
        "note": "the reference's figures are on its own natural scene (not shipped); context, not a target",
    }


def bench_plain(lib, er, cfg: dict, eyes: list, steps: int, warmup: int, trace_mode: int, copy_probe: float,
                cpu: bool, rms: bool) -> dict:
    """The plain method at bee scale: both eyes through renderFrame at 128
    spp, the reference's way of rendering (libEyeRenderer.cpp:229-242)."""
    spp = cfg["spp"]
    rays_per_step = 0
    stats = {"nodes_visited": 0, "tris_tested": 0, "hits": 0}
    omm_total = 0
    for i in eyes:
        lib.gotoCamera(i)
        lib.setCurrentEyeSamplesPerOmmatidium(spp)
        lib.setCameraPose(*[float(v) for v in cfg["pose"]])
        rays_per_step += int(lib.getCurrentEyeTotalSamples())
        omm_total += int(lib.getCurrentEyeOmmatidialCount())
        s = er.traversal_stats(lib)
        for k in stats:
            stats[k] += s[k]

    def step():
        for i in eyes:
            lib.gotoCamera(i)
            if lib.renderFrame() < 0:
                raise RuntimeError(er.last_error(lib))

    lib.setPacketCounting(True)
    er.packet_counts(lib)
    step()
    pk = er.packet_counts(lib)
    lib.setPacketCounting(False)
    for _ in range(warmup):
        step()
    trace_ms = []
    kernel_name = ""
    t0 = time.perf_counter()
    for _ in range(steps):
        for i in eyes:
            lib.gotoCamera(i)
            if lib.renderFrame() < 0:
                raise RuntimeError(er.last_error(lib))
            trace_ms.append(er.timings_np(lib)[0])
            kernel_name = lib.getLastRenderKernel().decode()
    elapsed = time.perf_counter() - t0
    trace_avg_ms = float(np.mean(trace_ms))
    # algorithmic bytes of one render launch (one eye-frame), per ray as the
    # rays' own traversals count them (a packet fetches a node once per wave)
    nb = node_bytes(lib, er)
    bytes_launch = traversal_bytes(stats, omm_total, 32 + 16, nb) / len(eyes)
    achieved = bytes_launch / (trace_avg_ms * 1e-3) / 1e9
    # what the waves fetch (each node and leaf once per 64-ray packet)
    packet_launch = (pk["nodes"] * nb + pk["tris"] * 48.0 + stats["hits"] * 48.0 + omm_total * 48.0) / len(eyes)
    traffic = measured_traffic(kernel_name, "c2")
    if copy_probe > 0:
        nbytes = int(copy_probe * 2 ** 20)
        cms = lib.runCopyProbe(nbytes, 10)
        log(f"copy probe: {nbytes} B in {cms:.4f} ms = {2 * nbytes / cms / 1e6:.1f} GB/s (read + write)")
    out = {
        "workload": "plain method, bee-scale: 2 eyes x 5,500 ommatidia x 128 spp through renderFrame, "
                    f"{cfg['scene'].tri_count()} triangles (1024^2 fBm heightfield + 2,000 cylinders)",
        "eye_frames_per_s": steps * len(eyes) / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "grays_per_s": steps * rays_per_step / elapsed / 1e9,
        "rays_per_step": rays_per_step,
        "trace_kernel_ms": trace_avg_ms,
        "traversal_per_ray": {"nodes": stats["nodes_visited"] / rays_per_step,
                              "tris": stats["tris_tested"] / rays_per_step,
                              "hit_fraction": stats["hits"] / rays_per_step},
        "roofline": {"bound": "latency of dependent node fetches (L2/MALL-resident scene)",
                     "achieved_packet": packet_launch / (trace_avg_ms * 1e-3) / 1e9,
                     "frac_packet": packet_launch / (trace_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "per_ray_equivalent_GBps": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "packet_nodes_per_64_rays": pk["nodes"] * 64.0 / rays_per_step,
                     "traffic": traffic,
                     "hbm_frac_measured": None if traffic is None else traffic / (trace_avg_ms * 1e-3) / 1e9 /
                     HBM_PEAK_GBS,
                     "kernel": kernel_name, "bytes_per_launch": bytes_launch, "trace_mode": trace_mode},
    }
    if rms:
        # RMS colour error of this render against a 4096-spp render of the same eye
        lib.gotoCamera(eyes[0])
        col128 = er.colours_np(lib)
        lib.setCurrentEyeSamplesPerOmmatidium(4096)
        lib.setCurrentEyeFrameIndex(1 << 20)
        lib.renderFrame()
        col4096 = er.colours_np(lib)
        lib.setCurrentEyeSamplesPerOmmatidium(spp)
        e = float(np.sqrt(np.mean((col128.astype(np.float64) - col4096) ** 2)))
        out["rms_colour_err_vs_4096spp"] = {"abs": e, "rel": e / float(np.sqrt(np.mean(col4096.astype(np.float64) ** 2))),
                                            "note": "eye 0, GPU 4096-spp render as reference"}
    if cpu:
        lib.gotoCamera(eyes[0])
        out["cpu_baseline"] = cpu_baseline(cfg, spp, er.pose_np(lib), target_s=5.0)
        out["gpu_vs_cpu"] = out["eye_frames_per_s"] / out["cpu_baseline"]["value"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="headline renders only (no CPU baseline, no other workloads): for profiler runs")
    ap.add_argument("--skip-batched", action="store_true", help="skip the 4,096-pose batched configuration")
    ap.add_argument("--skip-adaptive", action="store_true", help="skip learned adaptive sampling at bee scale")
    ap.add_argument("--skip-plain", action="store_true", help="skip the bee-scale renderFrame workload")
    ap.add_argument("--skip-speed-test", action="store_true", help="skip the reference speed-test sweep")
    ap.add_argument("--skip-large", action="store_true",
                    help="skip the headline large batch (the line then carries no headline value)")
    ap.add_argument("--adaptive-only", action="store_true",
                    help="run only learned adaptive sampling at bee scale (profiler runs)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the 4,096-pose batched configuration (profiler runs of the persistent renderer)")
    ap.add_argument("--plain-only", action="store_true",
                    help="run only the bee-scale renderFrame workload (profiler runs of k_render_tasks)")
    ap.add_argument("--no-split", action="store_true", help="disable split launches (chunks of a unit on several waves)")
    ap.add_argument("--trace-mode", type=int, default=2, choices=(0, 1, 2),
                    help="renderFrame traversal: 0 per-ray, 1 wave packets, 2 wave packets with scalar node loads "
                         "(default)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="also run a plain copy kernel of this many MiB (profiler bandwidth reference)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this renderer runs on one GPU; --gpus above 1 is not supported "
              "(the workload would shard by pose/eye, see the design notes)", file=sys.stderr)
        return 2
    cra = __graft_entry__.import_package()
    from compound_ray_amd import eyerenderer as er
    from compound_ray_amd import scenes

    lib = cra.load_library()
    lib.setVerbosity(False)
    lib.setTraversalMode(args.trace_mode)
    lib.setSplitLaunches(not args.no_split)
    tmp = tempfile.mkdtemp(prefix="eyebench-")
    if args.batched_only:
        out = {"batched": bench_batched(lib, er, scenes, tmp)}
        if args.copy_probe > 0:
            lib.runCopyProbe(int(args.copy_probe * 2 ** 20), 10)
        lib.stop()
        print(json.dumps(out))
        return 0
    only = args.adaptive_only or args.plain_only
    out = {
        "metric": METRIC,
        "value": None,
        "unit": "eye-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded scenes and eyes generated in-run, loaded via loadGlTFscene; "
                "sampler network trained in-run on held-out poses",
    }
    if not (only or args.skip_large):
        log("large heterogeneous batch (headline) ...")
        large = bench_large(lib, er, scenes, tmp, args.steps, args.warmup,
                            cpu=not (args.no_cpu_baseline or args.profile), profile=args.profile)
        timed = large.pop("_timed")
        out["value"] = timed["frames_per_step"] * timed["steps"] / timed["elapsed"]
        out["ms_per_step"] = timed["elapsed"] / timed["steps"] * 1e3
        out["config"] = {
            "workload": "learned adaptive + batched, large heterogeneous batch: " + large["workload"]
                        + f"; <= 0.01 RMS colour error; one step = one adaptive render of the batch "
                          f"({timed['frames_per_step']} eye-frames)",
            "eye_frames_per_step": timed["frames_per_step"],
            "samples_per_step": timed["samples_per_step"],
            "bvh": large["bvh"],
        }
        out["grays_per_s"] = large["grays_per_s_adaptive"]
        out["gain_vs_uniform_batched"] = large["gain_vs_uniform_batched"]
        out["gain_vs_network_free"] = large["gain_vs_network_free"]
        out["rms_abs_adaptive"] = large["rms_abs_adaptive"]
        out["roofline"] = large.pop("roofline")
        if "cpu_baseline" in large:
            out["cpu_baseline"] = large.pop("cpu_baseline")
        out["large_batch"] = large
        if not args.profile:
            log(f"C5: {out['value']:.0f} eye-frames/s adaptive, gain {large['gain_vs_uniform_batched']:.2f}x over "
                f"uniform, {large['gain_vs_network_free']:.2f}x over network-free")
    if args.profile and not only:
        if args.copy_probe > 0:
            lib.runCopyProbe(int(args.copy_probe * 2 ** 20), 10)
        lib.stop()
        print(json.dumps(out))
        return 0
    want_plain = not args.skip_plain and not args.adaptive_only
    want_adaptive = not args.skip_adaptive and not args.plain_only
    if want_plain or want_adaptive:
        t0 = time.time()
        cfg = scenes.build_config(2, tmp)
        log(f"generated C2 scene ({cfg['scene'].tri_count()} triangles) in {time.time() - t0:.1f} s")
        t0 = time.time()
        lib.loadGlTFscene(str(cfg["gltf"]).encode())
        err = er.last_error(lib)
        if err:
            raise RuntimeError(err)
        bvh = er.bvh_info(lib)
        log(f"loaded + GPU LBVH in {time.time() - t0:.1f} s (build {bvh['build_ms']:.1f} ms, {bvh['nodes']} nodes)")
        eyes = []
        for i in range(lib.getCameraCount()):
            lib.gotoCamera(i)
            if lib.isCompoundEyeActive():
                lib.setCurrentEyeSamplesPerOmmatidium(cfg["spp"])
                lib.setCameraPose(*[float(v) for v in cfg["pose"]])
                eyes.append(i)
        assert len(eyes) == 2
        if want_plain:
            log("plain method at bee scale (renderFrame) ...")
            plain = bench_plain(lib, er, cfg, eyes, args.steps, args.warmup, args.trace_mode, args.copy_probe,
                                cpu=not (args.no_cpu_baseline or args.profile), rms=not args.profile)
            plain["bvh"] = bvh
            out["plain_bee_scale"] = plain
        if want_adaptive:
            log("learned adaptive sampling at bee scale ...")
            out["learned_adaptive_bee_scale"] = bench_adaptive(lib, er, scenes, cfg, eyes)
    if not only and not args.skip_batched:
        log("batched configuration ...")
        out["batched"] = bench_batched(lib, er, scenes, tmp)
    if not only and not args.skip_speed_test:
        log("reference speed-test sweep ...")
        out["speed_test"] = bench_speed_test(lib, er, tmp)
    lib.stop()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
