#!/usr/bin/env python3
"""Benchmark of the photon-mapper hot path on one MI355X.

One step = one frame of the 1080p workload (1M surface + 2M volume photon
records, 64^3 value-noise medium, k=64 kNN radiance gather, 1920x1080 x 4
eye samples): photon emission/trace, kd-tree build of both maps, render
(eye rays, ray-march with a volume kNN query per step, surface kNN query at
the hit).  Inputs are synthetic (the scene is generated from Philox seed 1).

Prints one JSON line.  `value` = radiance queries (surface + volume kNN
density estimates) per second of whole-frame time.  `roofline` is for the
dominant kernel (the render/gather kernel), with its duration measured by HIP
events on the library's own stream inside libpm.so.  `cpu_baseline` times the
CPU oracle (same per-thread logic, std::thread over the host cores) on a
stratified sample of the same frame.  `cache` reports the learned-cache form of
the same frame (queries recorded, MLP trained online on exact kNN targets,
MLP inference with exact kNN fallback): ms/frame, speed-up over the plain
frame, its error against the plain frame on the same eye samples, and the
MFMA rate of the MLP inference.  `batch` reports the batched workload (64
random scenes per frame through one hipGraph launch) against the same scenes
rendered one at a time.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "cuda-photon-mapper_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import pmamd  # noqa: E402

METRIC = "radiance queries/s + ms/frame at 1M photons, k=64; batched/cache vs plain FP32"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_cfg():
    return pmamd.preset(2)


def workload_desc(c):
    return ("unit room, %dk surface + %dk volume photon records, %d^3 value-noise medium "
            "(sigma_s %.2g, sigma_a %.2g, HG g %.2g), k=%d kNN, r<=%.3g, %dx%d x %d spp, march step %.3g, FP32"
            % (c.cap_surface // 1000, c.cap_volume // 1000, c.noise_res, c.sigma_s, c.sigma_a, c.hg_g, c.k_volume,
               c.rcap_volume, c.width, c.height, c.spp, c.march_step))


def gather_bytes_per_query(k, bf16=False):
    # k neighbours x (12 B position + 12 B power + 12 B direction) + 36 B of
    # query I/O (position, normal or view direction, RGB out).  BF16 storage:
    # power and direction 6 B each.
    per = 12 + (6 + 6 if bf16 else 12 + 12)
    return k * per + 36


def run_gpu(cfg, steps, warmup, gather_mode, tile_regions=0):
    lib = pmamd.load()
    g = pmamd.PhotonMapper(cfg)
    g.set_option(pmamd.PM_OPT_GATHER_MODE, gather_mode)
    g.set_option(pmamd.PM_OPT_TILE_REGIONS, tile_regions)
    npx = cfg.width * cfg.height
    rgb = pmamd.DeviceBuffer(npx * 12)
    u8 = pmamd.DeviceBuffer(npx * 4)

    def frame():
        g.emit()
        g.build()
        pmamd.check(lib.pm_render(g.h, rgb.ptr, u8.ptr), "pm_render")
        return g.stats()

    for i in range(warmup):
        s = frame()
        log("warmup %d: emit %.1f ms, build %.1f ms, render %.1f ms" % (i, s.ms_emit, s.ms_build, s.ms_render))
    per = []
    t0 = time.perf_counter()
    for i in range(steps):
        s = frame()
        per.append((s.ms_emit, s.ms_build, s.ms_render, s.ms_gather_kernel, s.surface_queries, s.volume_queries))
        log("step %d: emit %.1f ms, build %.1f ms, render %.1f ms" % (i, s.ms_emit, s.ms_build, s.ms_render))
    t1 = time.perf_counter()
    img = rgb.download(np.float32, npx * 3)
    return (t1 - t0) / steps, per, g.stats(), img, g


def run_surface_gather(g, cfg, steps):
    """The surface-gather measurement: the same context and photon maps as the
    plain frame, rendered without the march, so every radiance query is a
    surface estimate (k_surface neighbours of the 1M surface records) at the
    diffuse hit of an eye path of the 1920x1080 x spp frame."""
    lib = pmamd.load()
    npx = cfg.width * cfg.height
    rgb = pmamd.DeviceBuffer(npx * 12)
    g.set_option(pmamd.PM_OPT_RENDER_MEDIA, 0)
    ms = []
    try:
        for i in range(steps + 1):
            pmamd.check(lib.pm_render(g.h, rgb.ptr, None), "pm_render")
            s = g.stats()
            if i > 0:
                ms.append(s.ms_render)
    finally:
        g.set_option(pmamd.PM_OPT_RENDER_MEDIA, 1)
    t = float(np.median(ms))
    q = int(s.surface_queries)
    bpq = gather_bytes_per_query(cfg.k_surface, cfg.storage == pmamd.PM_STORE_BF16)
    log("surface gather: %.1f ms for %d queries" % (t, q))
    ctl = run_wall_control(cfg, steps, bpq)
    return {"workload": "surface radiance estimates only (no march): the plain frame's %d surface records, k=%d, "
                        "r<=%g, at the diffuse hits of the %dx%d x %d spp eye paths"
                        % (int(s.n_surface), cfg.k_surface, cfg.rcap_surface, cfg.width, cfg.height, cfg.spp),
            "ms_per_frame": round(t, 3), "surface_queries": q, "radiance_queries_per_s": round(q / (t * 1e-3), 1),
            "bytes_per_query": bpq, "achieved_GBps": round(bpq * q / (t * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(bpq * q / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "control_uniform_walls": ctl}


def run_wall_control(cfg, steps, bpq, n=1000000):
    """The surface gather on a control map: n records uniform on the room's five
    walls (numpy default_rng(1)), each arriving along its wall's outward
    direction, rendered with the same frame and estimator."""
    lib = pmamd.load()
    rng = np.random.default_rng(1)
    w = rng.integers(0, 5, n)
    u, v = rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32)
    pos = np.zeros((n, 3), np.float32)
    dirs = np.zeros((n, 3), np.float32)
    axis = np.array([0, 0, 1, 1, 2])[w]
    side = np.array([0.0, 1.0, 0.0, 1.0, 1.0], np.float32)[w]
    others = np.array([[1, 2], [1, 2], [0, 2], [0, 2], [0, 1]])[w]
    rows = np.arange(n)
    pos[rows, axis] = side
    pos[rows, others[:, 0]] = u
    pos[rows, others[:, 1]] = v
    dirs[rows, axis] = np.where(side > 0.5, 1.0, -1.0)
    power = np.ones((n, 3), np.float32)
    g = pmamd.PhotonMapper(cfg)
    g.set_records(pmamd.PM_MAP_SURFACE, pos, power, dirs, n)
    g.build()
    g.set_option(pmamd.PM_OPT_RENDER_MEDIA, 0)
    npx = cfg.width * cfg.height
    rgb = pmamd.DeviceBuffer(npx * 12)
    ms = []
    for i in range(steps + 1):
        pmamd.check(lib.pm_render(g.h, rgb.ptr, None), "pm_render")
        s = g.stats()
        if i > 0:
            ms.append(s.ms_render)
    g.close()
    t = float(np.median(ms))
    q = int(s.surface_queries)
    log("surface gather, uniform-wall control: %.1f ms for %d queries" % (t, q))
    return {"records": n, "generator": "numpy default_rng(1), uniform on the 5 walls", "ms_per_frame": round(t, 3),
            "surface_queries": q, "radiance_queries_per_s": round(q / (t * 1e-3), 1),
            "achieved_GBps": round(bpq * q / (t * 1e-3) / 1e9, 1)}


def rel_rmse(a, b, exposure=None):
    if exposure is not None:  # displayed values: exposure, clamp to [0, 1]
        a, b = np.clip(a * exposure, 0, 1), np.clip(b * exposure, 0, 1)
    return float(np.sqrt(((a.astype(np.float64) - b) ** 2).mean()) / np.abs(b).mean())


def run_cache(g, cfg, steps, train_frames, stride, threshold, batch, plain_img):
    """Cache frames of the same workload: emit + build + cache render (record
    queries, train on 2^18 exact kNN targets, MLP inference, exact fallback),
    after `train_frames` frames of online training from a fresh network."""
    lib = pmamd.load()
    npx = cfg.width * cfg.height
    rgb = pmamd.DeviceBuffer(npx * 12)
    g.cache_reset(1)
    g.cache_configure(threshold=threshold, batch=batch, march_stride=stride)

    def frame():
        g.emit()
        g.build()
        pmamd.check(lib.pm_render_cache_crop(g.h, 0, 0, cfg.width, cfg.height, rgb.ptr, None), "pm_render_cache_crop")
        return g.cache_stats()

    for i in range(train_frames):
        cs = frame()
        if i % 4 == 0:
            log("cache training frame %d: %.1f ms, loss %.4g" % (i, cs.ms_total, cs.train_loss))
    per = []
    t0 = time.perf_counter()
    for i in range(steps):
        cs = frame()
        per.append((cs.ms_record, cs.ms_train, cs.ms_infer, cs.ms_fallback, cs.ms_accumulate, cs.ms_total))
        log("cache step %d: %.1f ms (infer %.1f ms), fallback %d of %d" % (i, cs.ms_total, cs.ms_infer, cs.fallbacks,
                                                                          cs.records))
    t1 = time.perf_counter()
    img = rgb.download(np.float32, npx * 3).reshape(-1, 3)
    m = np.mean(np.array(per), axis=0)
    infer_flop = 154368.0 * cs.records
    return {
        "ms_per_frame": round((t1 - t0) / steps * 1e3, 3),
        "breakdown_ms": {"record": round(m[0], 3), "train": round(m[1], 3), "infer": round(m[2], 3),
                         "fallback": round(m[3], 3), "accumulate": round(m[4], 3), "render_total": round(m[5], 3)},
        "march_stride": stride, "threshold": threshold, "adam_batch": batch, "train_samples_per_frame": 1 << 18,
        "train_frames_before_timing": train_frames,
        "cache_queries": int(cs.records), "fallback_queries": int(cs.fallbacks), "train_loss": round(cs.train_loss, 5),
        "rmse_vs_plain_same_samples": {"display_rel": round(rel_rmse(img, plain_img, cfg.exposure), 5),
                                       "linear_rel": round(rel_rmse(img, plain_img), 5)},
        "mlp_inference_mfma": {"bound": "mfma", "achieved": round(infer_flop / (m[2] * 1e-3) / 1e12, 2),
                               "peak": 2500.0, "unit": "TFLOP/s",
                               "frac": round(infer_flop / (m[2] * 1e-3) / 1e12 / 2500.0, 4),
                               "flop_per_query": 154368},
    }


def run_batch(steps, tune_stride, n_scenes=64, seed=1):
    """The batched workload: n_scenes random scenes (generator seed 1), one
    hipGraph launch per frame, tuning picked by the library's autotuner; then
    the same scenes one at a time through pm_create/emit/build/render."""
    lib = pmamd.load()
    cfgs = [pmamd.batch_scene(seed, i) for i in range(n_scenes)]
    t = time.perf_counter()
    b = pmamd.Batch(cfgs)
    t_setup = time.perf_counter() - t
    t = time.perf_counter()
    best, tune_ms = b.autotune(tune_stride)
    t_tune = time.perf_counter() - t
    log("batch: %d scenes, setup %.1f s, autotune %.1f s -> leaf %d, block %d, heap slots %d, key bytes %d, "
        "gather mode %d" % (n_scenes, t_setup, t_tune, best[0], best[1], best[2], best[3], best[4]))
    b.submit()
    per = []
    for i in range(steps):
        b.submit()
        per.append(b.stats().ms_frame)
        log("batch step %d: %.1f ms" % (i, per[-1]))
    st = b.stats()
    q = st.surface_queries + st.volume_queries
    ms = float(np.median(per))
    b.close()
    alone = 0.0
    # the one-scene gather with the same heap keys (or mode 7)
    mode = 7 if best[4] == 7 else {8: 1, 4: 2, 2: 4}[best[3]]
    for c in cfgs:
        c.leaf_size = best[0]
        g = pmamd.PhotonMapper(c)
        g.set_option(pmamd.PM_OPT_GATHER_MODE, mode)
        buf = pmamd.DeviceBuffer(c.width * c.height * 12)
        g.emit()
        g.build()
        pmamd.check(lib.pm_render(g.h, buf.ptr, None), "pm_render")
        s = g.stats()
        alone += s.ms_emit + s.ms_build + s.ms_render
        g.close()
    log("batch: %.1f ms per frame; the scenes one at a time: %.1f ms" % (ms, alone))
    return {
        "workload": "%d random scenes (2-256 spheres: 60%% Lambertian / 25%% mirror / 15%% dielectric IOR 1.5; "
                    "photons log-uniform 1e4-2e6; images 128^2-1024^2; homogeneous sigma_s U[0,0.3]; k=32, r<=0.05, "
                    "1 spp)" % n_scenes,
        "photons": int(st.photons), "pixels": int(st.pixels), "surface_records": int(st.surface_records),
        "volume_records": int(st.volume_records), "radiance_queries": int(q),
        "ms_per_frame": round(ms, 3), "radiance_queries_per_s": round(q / (ms * 1e-3), 1),
        "tuning": {"leaf_size": best[0], "block_threads": best[1], "heap_slots": best[2], "heap_key_bytes": best[3],
                   "gather_mode": best[4], "autotune_sample_stride": tune_stride, "autotune_s": round(t_tune, 2)},
        "graph_nodes": int(st.graph_nodes), "work_items": int(st.work_items),
        "scenes_one_at_a_time_ms": round(alone, 3), "speedup_vs_one_at_a_time": round(alone / ms, 3),
    }


def run_cpu(cfg, budget_s):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    cores = min(16, len(os.sched_getaffinity(0)))
    o = O.Oracle(cfg, cores)
    t = time.perf_counter()
    o.emit()
    t_emit = time.perf_counter() - t
    t = time.perf_counter()
    o.build()
    t_build = time.perf_counter() - t
    # Stratified rows of the same frame until the time budget is spent.
    H = cfg.height
    order = []
    stride = 64
    while stride >= 1 and len(order) < H:
        for y in range(stride // 2 if stride > 1 else 0, H, stride):
            if y not in order:
                order.append(y)
        stride //= 2
    rows = queries = 0
    t_render = 0.0
    for y in order:
        t = time.perf_counter()
        o.render(0, y, cfg.width, 1)
        t_render += time.perf_counter() - t
        rows += 1
        queries += sum(o.last_queries)
        if t_render > budget_s:
            break
    qps = queries / t_render
    return dict(qps=qps, cores=cores, rows=rows, queries=queries, t_render=t_render, t_emit=t_emit, t_build=t_build,
                cpu=platform.processor() or platform.machine())


def read_traffic():
    p = os.path.join(REPO, "profiles", "traffic_render_cfg2.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--gather-mode", type=int, default=5,
                    help="0: wave per query; lane per query with 1: 8-byte, 2: 4-byte, 3/4/6/7: 2-byte heap keys; 5: auto")
    ap.add_argument("--tile-regions", type=int, default=0,
                    help="0: raster tile order; 1, 2, 4, 8: per-XCD regions, that many across")
    ap.add_argument("--frame", type=str, default="", help="WxHxSPP override of the frame (profiling runs only)")
    ap.add_argument("--surface-steps", type=int, default=3,
                    help="timed surface-only frames (0: skip the surface-gather leg)")
    ap.add_argument("--cache-steps", type=int, default=3, help="timed cache frames (0: skip the cache leg)")
    ap.add_argument("--cache-train-frames", type=int, default=48)
    ap.add_argument("--cache-batch", type=int, default=8192, help="Adam batch (2^18 training samples per frame)")
    ap.add_argument("--cache-stride", type=int, default=4)
    ap.add_argument("--cache-threshold", type=float, default=0.12)
    ap.add_argument("--batch-steps", type=int, default=3, help="timed batch frames (0: skip the batch leg)")
    ap.add_argument("--batch-tune-stride", type=int, default=4)
    ap.add_argument("--copy-calibrate", type=int, default=0,
                    help="run only the library's copy kernel over this many bytes (memory-counter calibration)")
    args = ap.parse_args()
    if args.copy_calibrate:
        n = args.copy_calibrate // 16 * 16
        src, dst = pmamd.DeviceBuffer(n), pmamd.DeviceBuffer(n)
        for _ in range(3):
            pmamd.check(pmamd.load().pm_copy_kernel(dst.ptr, src.ptr, n), "pm_copy_kernel")
        print(json.dumps({"copy_bytes": n, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X only: this build runs on a single GPU (got --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    cfg = workload_cfg()
    if args.frame:
        w, h, spp = (int(v) for v in args.frame.lower().split("x"))
        cfg.width, cfg.height, cfg.spp = w, h, spp
    sec_per_frame, per, stats, img, g = run_gpu(cfg, args.steps, args.warmup, args.gather_mode, args.tile_regions)
    q_frame = stats.surface_queries + stats.volume_queries
    value = q_frame / sec_per_frame
    kern_ms = float(np.mean([p[3] for p in per]))
    bpq = gather_bytes_per_query(cfg.k_volume, cfg.storage == pmamd.PM_STORE_BF16)
    achieved = bpq * q_frame / (kern_ms * 1e-3) / 1e9
    tr = read_traffic()
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (tr.get("bytes_per_launch") if tr else None),
            "kernel": "k_render_lane7 (eye ray + march + kNN gather)", "bytes_per_query": bpq,
            "queries_per_launch": int(q_frame), "kernel_ms": round(kern_ms, 3)}
    photons_traced = stats.n_emit_total
    emit_ms = float(np.mean([p[0] for p in per]))
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "radiance queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_per_frame * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cfg2 plain FP32 frame: " + workload_desc(cfg),
                   "photons_emitted": int(photons_traced), "surface_records": int(stats.n_surface),
                   "volume_records": int(stats.n_volume), "surface_queries": int(stats.surface_queries),
                   "volume_queries": int(stats.volume_queries), "sigma_a_used": cfg.sigma_a,
                   "march_step_used": cfg.march_step},
        "breakdown_ms": {"emit": round(emit_ms, 3), "build": round(float(np.mean([p[1] for p in per])), 3),
                         "render": round(float(np.mean([p[2] for p in per])), 3)},
        "photons_traced_per_s": round(photons_traced / (emit_ms * 1e-3), 1),
        "roofline": roof,
        "image_mean_rgb": [round(float(x), 5) for x in img.reshape(-1, 3).mean(0)],
    }
    if args.surface_steps > 0:
        out["surface_gather"] = run_surface_gather(g, cfg, args.surface_steps)
    if args.cache_steps > 0:
        ca = run_cache(g, cfg, args.cache_steps, args.cache_train_frames, args.cache_stride, args.cache_threshold,
                       args.cache_batch, img.reshape(-1, 3))
        ca["speedup_vs_plain_frame"] = round(out["ms_per_step"] / ca["ms_per_frame"], 3)
        out["cache"] = ca
    if args.batch_steps > 0:
        out["batch"] = run_batch(args.batch_steps, args.batch_tune_stride)
    if not args.no_cpu:
        cb = run_cpu(cfg, args.cpu_seconds)
        out["cpu_baseline"] = {
            "value": round(cb["qps"], 1), "unit": "radiance queries/s", "cores": cb["cores"], "kind": "port",
            "sample": "CPU oracle (g++ -O2, std::thread x %d): full cfg2 emit (%.1f s) and kd-build (%.1f s), then "
                      "render of %d stratified rows of the 1920x1080x4 frame (%d radiance queries in %.1f s)"
                      % (cb["cores"], cb["t_emit"], cb["t_build"], cb["rows"], cb["queries"], cb["t_render"]),
            "est_ms_per_frame": round(q_frame / cb["qps"] * 1e3 + (cb["t_emit"] + cb["t_build"]) * 1e3, 1),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
