"""Benchmark: compound-eye rendering on one MI355X.

Workload (the bee-scale plain-method configuration): two 5,500-ommatidium
ellipsoidal-patch eyes (160 x 120 deg, acceptance U[1.5, 3] deg) at 128
samples per ommatidium over a 2,189,058-triangle scene (1024^2 fBm heightfield
+ 2,000 16-segment cylinders), generated in-run from fixed seeds and loaded
through loadGlTFscene.  One step renders both eyes once (2 eye-frames,
1,408,000 rays) through renderFrame; inputs are resident in HBM before the
timed region.  The metric is eye-frames/s; Grays/s, the trace kernel's
roofline fraction and a CPU-oracle baseline are reported beside it.

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


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` on the bee-scale workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary named for that kernel
    (already scaled by the copy-kernel factors recorded there); None when no
    summary of this kernel is present."""
    base = kernel.split("<")[0]
    try:
        for line in (PROFILES / f"{base}_c2_traffic.txt").read_text().splitlines():
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


def bench_large(lib, er, scenes, tmp: str, target: float = 0.01, n_ref: int = 100_000) -> dict:
    """Large heterogeneous batch: 256 eyes (log-uniform 1k-30k ommatidia,
    Fibonacci-sphere and ellipsoidal layouts, acceptance 1-5 deg) at 256 poses
    over a 13.4M-triangle scene, rendered as one batch (all eyes' ommatidia
    moved to world space and traced in one persistent launch), learned
    adaptive sampling against uniform sampling at the same RMS error."""
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
    log(f"C5: {scene.tri_count()} triangles generated in {gen_s:.1f} s, loaded + BVH in {load_s:.1f} s")
    er.gotoFirstCompoundEye(lib)

    def world_batch(seed):
        poses = scenes.batch_poses(scene, len(eyes), seed=seed, extent=80.0)
        return np.concatenate([scenes.world_ommatidia(e, scenes.pose_matrix12(p)) for e, p in zip(eyes, poses)])

    identity = np.zeros((1, 6), np.float32)
    # training on held-out poses
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
    log(f"C5: trained on {len(L)} rows in {train_s:.1f} s")
    # the test batch; reference colours of its first n_ref ommatidia at 4,096 spp
    batch = world_batch(seed=902)
    n_omm = len(batch)
    er.set_ommatidia_np(lib, batch[:n_ref])
    lib.setCurrentEyeSamplesPerOmmatidium(4096)
    lib.setCurrentEyeFrameIndex(1 << 20)
    ref = er.render_batch_np(lib, identity)[1][0]
    er.set_ommatidia_np(lib, batch)
    lib.setCameraPose(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)

    def uniform(spp, want_rgb=True):
        lib.setCurrentEyeSamplesPerOmmatidium(spp)
        lib.setCurrentEyeFrameIndex(0)
        return er.render_batch_np(lib, identity, want_rgb=want_rgb)

    def err_of(rgb):
        return rms_abs(rgb[0, :n_ref], ref)

    e64 = err_of(uniform(64)[1])
    spp_u = max(1, int(math.ceil(64 * (e64 / target) ** 2)))
    e_u = err_of(uniform(spp_u)[1])
    while e_u > target and spp_u < 4096:
        spp_u = int(math.ceil(spp_u * 1.1))
        e_u = err_of(uniform(spp_u)[1])

    # launch shape of both timed renders, autotuned on the uniform batch
    lib.setCurrentEyeSamplesPerOmmatidium(spp_u)
    er.autotune_np(lib, identity, repeats=1)
    shape = er.render_tuning(lib)

    def adaptive(scale, want_rgb=True):
        lib.setCurrentEyeFrameIndex(0)
        return er.render_adaptive_np(lib, identity, error_target=target, max_spp=512, budget_scale=scale,
                                     want_rgb=want_rgb)

    def adaptive_err(scale):
        _, rgb, _, total = adaptive(scale)
        return err_of(rgb), total

    scale, e_a, total = calibrate_budget(adaptive_err, target)

    def timed(fn, repeats=3):
        fn()
        t = time.perf_counter()
        for _ in range(repeats):
            fn()
        return (time.perf_counter() - t) / repeats

    t_a = timed(lambda: adaptive(scale, want_rgb=False))
    t_u = timed(lambda: uniform(spp_u, want_rgb=False))
    frames = len(eyes)
    return {
        "workload": f"{len(eyes)} eyes ({n_omm} ommatidia, log-uniform 1k-30k per eye) at {len(eyes)} poses, "
                    f"{scene.tri_count()} triangles (2048^2 heightfield + 20,000 cylinders + 200 rocks), "
                    "one batch per render",
        "eye_frames_per_s_adaptive": frames / t_a,
        "eye_frames_per_s_uniform_batched": frames / t_u,
        "gain_vs_uniform_batched": t_u / t_a,
        "grays_per_s_adaptive": total / t_a / 1e9,
        "grays_per_s_uniform": n_omm * spp_u / t_u / 1e9,
        "uniform_spp": spp_u,
        "adaptive_mean_spp": total / n_omm,
        "rms_abs_uniform": e_u,
        "rms_abs_adaptive": e_a,
        "rms_reference": f"first {n_ref} ommatidia of the batch, uniform 4,096-spp render on the GPU",
        "budget_scale": scale,
        "launch_shape": shape,
        "bvh": bvh,
        "scene_load_s": load_s,
        "training": {"rows": int(len(L)), "label_spp": 1024, "seconds": train_s, "loss_first": float(losses[0]),
                     "loss_last": float(losses[-1])},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="render loop only (no 4096-spp error check, no CPU baseline): for profiler runs")
    ap.add_argument("--skip-batched", action="store_true", help="skip the 4,096-pose batched configuration")
    ap.add_argument("--skip-adaptive", action="store_true", help="skip the learned adaptive sampling comparison")
    ap.add_argument("--adaptive-only", action="store_true",
                    help="run only the learned adaptive comparison on the bee-scale scene (profiler runs)")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the 4,096-pose batched configuration (profiler runs of the persistent renderer)")
    ap.add_argument("--large", action="store_true",
                    help="also run the 256-eye, 13.4M-triangle heterogeneous batch (several minutes)")
    ap.add_argument("--no-split", action="store_true", help="disable split launches (chunks of a unit on several waves)")
    ap.add_argument("--trace-mode", type=int, default=1, choices=(0, 1, 2),
                    help="traversal: 0 per-ray, 1 wave packets (default), 2 wave packets with scalar node loads")
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
    t0 = time.time()
    cfg = scenes.build_config(2, tmp)
    log(f"generated C2 scene ({cfg['scene'].tri_count()} triangles) in {time.time() - t0:.1f} s")
    spp = cfg["spp"]
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
            lib.setCurrentEyeSamplesPerOmmatidium(spp)
            lib.setCameraPose(*[float(v) for v in cfg["pose"]])
            eyes.append(i)
    assert len(eyes) == 2
    if args.adaptive_only:
        out = {"learned_adaptive": bench_adaptive(lib, er, scenes, cfg, eyes)}
        lib.stop()
        print(json.dumps(out))
        return 0
    rays_per_step = 0
    stats = {"nodes_visited": 0, "tris_tested": 0, "hits": 0}
    omm_counts = []
    for i in eyes:
        lib.gotoCamera(i)
        rays_per_step += int(lib.getCurrentEyeTotalSamples())
        omm_counts.append(int(lib.getCurrentEyeOmmatidialCount()))
        s = er.traversal_stats(lib)
        for k in stats:
            stats[k] += s[k]

    def step():
        for i in eyes:
            lib.gotoCamera(i)
            if lib.renderFrame() < 0:
                raise RuntimeError(er.last_error(lib))

    for _ in range(args.warmup):
        step()
    trace_ms = []
    kernel_name = ""
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for i in eyes:
            lib.gotoCamera(i)
            if lib.renderFrame() < 0:
                raise RuntimeError(er.last_error(lib))
            trace_ms.append(er.timings_np(lib)[0])
            kernel_name = lib.getLastRenderKernel().decode()
    elapsed = time.perf_counter() - t0
    eye_frames = args.steps * len(eyes)
    value = eye_frames / elapsed
    grays = args.steps * rays_per_step / elapsed / 1e9
    trace_avg_ms = float(np.mean(trace_ms))
    # algorithmic bytes of one render launch (one eye-frame): BVH nodes visited
    # x 128 B + triangles tested x 48 B + shading fetch (3 x 16 B per hit) +
    # per ommatidium its 32 B record read and its 16 B colour written
    omm_total = sum(omm_counts)
    bytes_launch = (stats["nodes_visited"] * 128 + stats["tris_tested"] * 48 + stats["hits"] * 48
                    + omm_total * (32 + 16)) / len(eyes)
    achieved = bytes_launch / (trace_avg_ms * 1e-3) / 1e9
    if args.copy_probe > 0:
        nbytes = int(args.copy_probe * 2 ** 20)
        cms = lib.runCopyProbe(nbytes, 10)
        log(f"copy probe: {nbytes} B in {cms:.4f} ms = {2 * nbytes / cms / 1e6:.1f} GB/s (read + write)")
    rms = rms_rel = None
    if not args.profile:
        # RMS colour error of this render against a 4096-spp render of the same eye
        lib.gotoCamera(eyes[0])
        col128 = er.colours_np(lib)
        lib.setCurrentEyeSamplesPerOmmatidium(4096)
        lib.setCurrentEyeFrameIndex(1 << 20)
        lib.renderFrame()
        col4096 = er.colours_np(lib)
        lib.setCurrentEyeSamplesPerOmmatidium(spp)
        rms = float(np.sqrt(np.mean((col128.astype(np.float64) - col4096) ** 2)))
        rms_rel = rms / float(np.sqrt(np.mean(col4096.astype(np.float64) ** 2)))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "eye-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded C2 scene and eyes generated in-run, loaded via loadGlTFscene",
        "config": {
            "workload": "plain method, bee-scale: 2 eyes x 5,500 ommatidia x 128 spp, "
                        f"{cfg['scene'].tri_count()} triangles (1024^2 fBm heightfield + 2,000 cylinders)",
            "rays_per_step": rays_per_step,
            "eye_frames_per_step": len(eyes),
            "bvh": bvh,
        },
        "grays_per_s": grays,
        "rms_colour_err_vs_4096spp": {"abs": rms, "rel": rms_rel, "note": "eye 0, GPU 4096-spp render as reference"},
        "trace_kernel_ms": trace_avg_ms,
        "traversal_per_ray": {"nodes": stats["nodes_visited"] / rays_per_step,
                              "tris": stats["tris_tested"] / rays_per_step,
                              "hit_fraction": stats["hits"] / rays_per_step},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kernel_name),
                     "kernel": kernel_name, "bytes_per_launch": bytes_launch,
                     "trace_mode": args.trace_mode},
    }
    if not (args.no_cpu_baseline or args.profile):
        lib.gotoCamera(eyes[0])
        pose12 = er.pose_np(lib)
        out["cpu_baseline"] = cpu_baseline(cfg, spp, pose12)
    if not args.profile and not args.skip_adaptive:
        log("learned adaptive sampling ...")
        out["learned_adaptive"] = bench_adaptive(lib, er, scenes, cfg, eyes)
    if args.large:
        log("large heterogeneous batch ...")
        out["large_batch"] = bench_large(lib, er, scenes, tmp)
    if not args.profile and not args.skip_batched:
        log("batched configuration ...")
        out["batched"] = bench_batched(lib, er, scenes, tmp)
    lib.stop()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
