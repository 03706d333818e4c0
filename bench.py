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
events on the library's own stream inside libpm.so; it also carries the
kernel's VALU issue rate from the committed counter profile, the roof it
actually sits on.  `cpu_baseline` times the CPU oracle (same per-thread logic,
std::thread over the host-core share of one GPU) on a stratified sample of the
same frame, and the 256x256 unit-room workload in full.  `cache_bf16` reports
the frame in its BF16-power + learned-cache form (queries recorded, MLP
trained online on exact kNN targets, MLP inference with exact kNN fallback):
ms/frame, speed-up over the plain FP32 frame, and its error against the FP64
oracle render of stratified rows (also given for the plain FP32 frame as
`plain_fp32_vs_fp64_oracle`).  `batch` reports the batched workload (64 random
scenes per frame through one hipGraph launch) against the same scenes
rendered one at a time.  `voxel_grid` reports the reference's own 32^3 voxel-grid
estimate (the compatibility mode) on the same photon maps and frame, with the
CPU oracle's rate on rows of it.

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


def run_gpu(cfg, steps, warmup, gather_mode, tile_regions=0, top_levels=0):
    lib = pmamd.load()
    g = pmamd.PhotonMapper(cfg)
    g.set_option(pmamd.PM_OPT_GATHER_MODE, gather_mode)
    g.set_option(pmamd.PM_OPT_TILE_REGIONS, tile_regions)
    g.set_option(pmamd.PM_OPT_TOP_LEVELS, top_levels)
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
    # the measured leg's gather (mode 7): the automatic choice would pick the
    # one-pass gather here, since this map alone holds fewer than 2^20 records
    g.set_option(pmamd.PM_OPT_GATHER_MODE, 7)
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
    return {"records": n, "generator": "numpy default_rng(1), uniform on the 5 walls", "gather_mode": 7,
            "ms_per_frame": round(t, 3),
            "surface_queries": q, "radiance_queries_per_s": round(q / (t * 1e-3), 1),
            "achieved_GBps": round(bpq * q / (t * 1e-3) / 1e9, 1)}


def rel_rmse(a, b, exposure=None):
    if exposure is not None:  # displayed values: exposure, clamp to [0, 1]
        a, b = np.clip(a * exposure, 0, 1), np.clip(b * exposure, 0, 1)
    return float(np.sqrt(((a.astype(np.float64) - b) ** 2).mean()) / np.abs(b).mean())


FLOP_FULL = 2 * (216 * 128 + 3 * 128 * 128 + 128 * 3)    # 154,368
FLOP_VOLUME = 2 * (144 * 128 + 3 * 128 * 128 + 128 * 3)  # 135,936


def bf16_cache_cfg():
    """The 1080p workload in its BF16 precision form: photon power and
    direction stored BF16 (positions and split planes FP32), FP64
    transmittance on long segments through the density field."""
    c = workload_cfg()
    c.storage = pmamd.PM_STORE_BF16
    c.fp64_transmittance = 1
    return c


def run_cache(steps, train_frames, stride, thresholds, batch, rows, near_groups=2, extra_error_frames=0,
                  ema=950):
    """Cache frames of the 1080p workload in the BF16 form: emit + build + cache
    render (record queries, train on 2^18 exact kNN targets, MLP inference,
    exact fallback), after `train_frames` frames of online training from a
    fresh network, at each fallback threshold in turn (training continues).
    Returns the timings and the frame rows `rows` of each setting (their error
    against the FP64 oracle is computed by the CPU leg)."""
    lib = pmamd.load()
    cfg = bf16_cache_cfg()
    g = pmamd.PhotonMapper(cfg)
    npx = cfg.width * cfg.height
    rgb = pmamd.DeviceBuffer(npx * 12)
    g.cache_reset(1)
    g.cache_configure(threshold=thresholds[0], batch=batch, march_stride=stride)
    g.set_option(pmamd.PM_OPT_CACHE_NEAR_GROUPS, near_groups)
    g.set_option(pmamd.PM_OPT_CACHE_EMA, ema)

    def frame():
        g.emit()
        g.build()
        pmamd.check(lib.pm_render_cache_crop(g.h, 0, 0, cfg.width, cfg.height, rgb.ptr, None), "pm_render_cache_crop")
        return g.cache_stats()

    for i in range(train_frames):
        cs = frame()
        if i % 8 == 0:
            log("cache training frame %d: %.1f ms, loss %.4g" % (i, cs.ms_total, cs.train_loss))
    out = []
    for th in thresholds:
        g.cache_configure(threshold=th, batch=batch, march_stride=stride)
        frame()
        per = []
        t0 = time.perf_counter()
        for i in range(steps):
            cs = frame()
            per.append((cs.ms_record, cs.ms_train, cs.ms_infer, cs.ms_fallback, cs.ms_accumulate, cs.ms_total))
            log("cache step %d (threshold %g): %.1f ms (infer %.1f ms), fallback %d of %d"
                % (i, th, cs.ms_total, cs.ms_infer, cs.fallbacks, cs.records))
        t1 = time.perf_counter()
        img = rgb.download(np.float32, npx * 3).reshape(cfg.height, cfg.width, 3)
        # the error rows of more frames (untimed; training continues), for the
        # spread of the error from frame to frame
        more_rows = []
        for i in range(extra_error_frames):
            frame()
            more_rows.append(rgb.download(np.float32, npx * 3).reshape(cfg.height, cfg.width, 3)[rows].copy())
        # the same frames without the fallback's seeds from the previous frame
        # (as when the photon positions change every frame): heap passes
        g.set_option(pmamd.PM_OPT_CACHE_SEEDS, 0)
        t2 = time.perf_counter()
        fb_noseed = []
        for i in range(steps):
            fb_noseed.append(frame().ms_fallback)
        t3 = time.perf_counter()
        g.set_option(pmamd.PM_OPT_CACHE_SEEDS, 1)
        m = np.mean(np.array(per), axis=0)
        # algorithmic FLOPs of the network per row: 2 x (216 x 128 + 3 x 128^2 + 128 x 3) on the full
        # layer 1; 2 x (144 x 128 + ...) for volume rows, whose 72 normal features are constants
        infer_flop = FLOP_FULL * (cs.inferred - cs.inferred_volume) + FLOP_VOLUME * cs.inferred_volume
        out.append({
            "threshold": th,
            "ms_per_frame": round((t1 - t0) / steps * 1e3, 3),
            "breakdown_ms": {"record": round(m[0], 3), "train": round(m[1], 3), "infer": round(m[2], 3),
                             "fallback": round(m[3], 3), "accumulate": round(m[4], 3),
                             "render_total": round(m[5], 3)},
            "without_fallback_seeds": {"ms_per_frame": round((t3 - t2) / steps * 1e3, 3),
                                       "fallback_ms": round(float(np.mean(fb_noseed)), 3)},
            "cache_queries": int(cs.records), "fallback_queries": int(cs.fallbacks),
            "train_loss": round(cs.train_loss, 5),
            "mlp_inference_mfma": {"bound": "mfma", "achieved": round(infer_flop / (m[2] * 1e-3) / 1e12, 2),
                                   "peak": 2500.0, "unit": "TFLOP/s",
                                   "frac": round(infer_flop / (m[2] * 1e-3) / 1e12 / 2500.0, 4),
                                   "rows": int(cs.inferred), "rows_volume_layer1": int(cs.inferred_volume),
                                   "flop_per_row": {"full": FLOP_FULL, "volume": FLOP_VOLUME}},
            "_rows": img[rows].copy(),
            "_more_rows": more_rows,
        })
    g.close()
    return {"workload": "the same frame with photon power and direction stored BF16 (positions and splits FP32), "
                        "FP64 transmittance on segments of >= 32 march steps through the density field, and the "
                        "learned radiance cache (MLP 216 -> 4 x 128 -> 3, BF16 MFMA, Adam lr 1e-3, 2^18 training "
                        "samples per frame, exact k=64 fallback above the threshold)",
            "march_stride": stride, "near_light_groups": near_groups, "weight_average_decay": ema / 1000.0,
            "adam_batch": batch,
            "train_samples_per_frame": 1 << 18,
            "train_frames_before_timing": train_frames, "settings": out}


def row_errors(img_rows, ref_rows, exposure):
    """Errors of FP32 frame rows against the FP64 oracle's rows: RMSE relative
    to the mean of the reference, on linear radiance and on the displayed
    values (exposure, clamp to [0, 1]); relative error quantiles."""
    a = img_rows.astype(np.float64)
    b = ref_rows
    da, db = np.clip(a * exposure, 0, 1), np.clip(b * exposure, 0, 1)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6 * np.abs(b).mean())
    return {"rmse_linear_rel": round(float(np.sqrt(((a - b) ** 2).mean()) / np.abs(b).mean()), 6),
            "rmse_display_rel": round(float(np.sqrt(((da - db) ** 2).mean()) / db.mean()), 6),
            "rel_err_median": float("%.4g" % np.median(rel)), "rel_err_p99": float("%.4g" % np.quantile(rel, 0.99)),
            "rel_err_max": float("%.4g" % rel.max())}


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
        "gather mode %d, top levels %d" % (n_scenes, t_setup, t_tune, best[0], best[1], best[2], best[3], best[4],
                                           best[5]))
    b.submit()
    per = []
    for i in range(steps):
        b.submit()
        per.append(b.stats().ms_frame)
        log("batch step %d: %.1f ms" % (i, per[-1]))
    st = b.stats()
    q = st.surface_queries + st.volume_queries
    ms = float(np.median(per))
    # the dominant kernel's roofline: the same frame launched kernel by
    # kernel, with events around its render (one launch, every scene's tiles)
    ms_emit, ms_build, ms_render = b.time_phases()
    log("batch phases, kernel by kernel: emit %.1f ms, build %.1f ms, render %.1f ms" % (ms_emit, ms_build, ms_render))
    bf16 = cfgs[0].storage == pmamd.PM_STORE_BF16
    ks = {c.k_surface for c in cfgs} | {c.k_volume for c in cfgs}
    bpq = gather_bytes_per_query(max(ks), bf16)
    achieved = bpq * q / (ms_render * 1e-3) / 1e9
    tb = read_profile("traffic_batch_cfg4.json")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": (tb.get("bytes_per_launch") if tb else None),
            "kernel": "k_render_batch (persistent work queue over every scene's tiles)",
            "bytes_per_query": bpq, "queries_per_launch": int(q), "kernel_ms": round(ms_render, 3),
            "phases_ms": {"emit": round(ms_emit, 3), "build": round(ms_build, 3), "render": round(ms_render, 3)}}
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
                   "gather_mode": best[4], "top_levels": best[5], "autotune_sample_stride": tune_stride,
                   "autotune_s": round(t_tune, 2)},
        "graph_nodes": int(st.graph_nodes), "work_items": int(st.work_items),
        "scenes_one_at_a_time_ms": round(alone, 3), "speedup_vs_one_at_a_time": round(alone / ms, 3),
        "roofline": roof,
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


# Host cores given to one GPU on the benchmark box: the oracle's threads are
# sized to that share, whatever the machine's total.
CPU_SHARE_PER_GPU = 16


def run_cpu(cfg, budget_s, ref_rows=None):
    """The cpu_baseline leg: the CPU oracle (the same per-thread logic, g++ -O2,
    std::thread) timed on this host.  (1) The 1080p workload: full emit and
    kd-build, then stratified rows of the frame until `budget_s` is spent
    (extrapolated).  (2) The 256x256 unit-room workload (100k photons, k=32)
    in full.  (3) As the checker, the FP64 oracle render of the frame rows
    `ref_rows` (the reference for the BF16 + cache frame's error)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O

    aff = len(os.sched_getaffinity(0))
    cores = min(CPU_SHARE_PER_GPU, aff)
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
    ref = None
    t_ref = 0.0
    if ref_rows is not None:
        t = time.perf_counter()
        ref = np.stack([o.render(0, int(y), cfg.width, 1, fp64=True)[0] for y in ref_rows])
        t_ref = time.perf_counter() - t
    del o
    c1 = pmamd.preset(1)
    o1 = O.Oracle(c1, cores)
    t = time.perf_counter()
    o1.emit()
    t1_emit = time.perf_counter() - t
    t = time.perf_counter()
    o1.build()
    t1_build = time.perf_counter() - t
    t = time.perf_counter()
    o1.render()
    t1_render = time.perf_counter() - t
    q1 = sum(o1.last_queries)
    cfg1 = {"workload": "unit room, 100k photons, homogeneous medium, k=32, r<=0.05, 256x256 x 1 spp, in full",
            "emit_s": round(t1_emit, 3), "build_s": round(t1_build, 3), "render_s": round(t1_render, 3),
            "radiance_queries": q1, "radiance_queries_per_s": round(q1 / t1_render, 1),
            "photons_traced_per_s": round(c1.n_emit / t1_emit, 1),
            "ms_per_frame": round((t1_emit + t1_build + t1_render) * 1e3, 1)}
    return dict(qps=qps, cores=cores, cores_affinity=aff, cores_total=os.cpu_count(), rows=rows, queries=queries,
                t_render=t_render, t_emit=t_emit, t_build=t_build, cpu=cpu_model(), cfg1=cfg1, ref=ref, t_ref=t_ref)


def run_voxel(cfg, steps, cpu_budget_s):
    """The voxel-grid compatibility mode (PM_OPT_VOXEL_GRID, the reference's
    own 32^3-grid estimate) on the 1080p workload's photon maps: grids and
    summed-volume tables built once per emission, then the frame with the
    window estimates in place of the kNN gather, without and with the bilinear
    blend.  With cpu_budget_s > 0 the CPU oracle renders stratified rows of
    the same frame from the same records (the checker and the CPU figure)."""
    lib = pmamd.load()
    g = pmamd.PhotonMapper(cfg)
    g.emit()
    npx = cfg.width * cfg.height
    rgb = pmamd.DeviceBuffer(npx * 12)
    res = {"workload": "voxel-grid estimate on the 1080p workload's photon maps: " + workload_desc(cfg),
           "kernel": "k_render_voxel (eye ray + march, window sums from summed-volume tables)"}
    imgs = {}
    for mode in (1, 2):
        g.set_option(pmamd.PM_OPT_VOXEL_GRID, mode)
        t = time.perf_counter()
        pmamd.check(lib.pm_render(g.h, rgb.ptr, None), "pm_render")
        first = (time.perf_counter() - t) * 1e3
        ms = []
        for _ in range(steps):
            pmamd.check(lib.pm_render(g.h, rgb.ptr, None), "pm_render")
            ms.append(g.stats().ms_render)
        s = g.stats()
        q = int(s.surface_queries + s.volume_queries)
        tk = float(np.median(ms))
        key = "window" if mode == 1 else "bilinear"
        res[key] = {"ms_per_frame_render": round(tk, 3), "radiance_queries": q,
                    "radiance_queries_per_s": round(q / (tk * 1e-3), 1)}
        if mode == 1:
            res["first_render_ms_with_grid_build"] = round(first, 3)
        imgs[mode] = rgb.download(np.float32, npx * 3).reshape(cfg.height, cfg.width, 3)
        log("voxel mode %d: %.2f ms render, %d queries" % (mode, tk, q))
    if cpu_budget_s > 0:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O

        cores = min(CPU_SHARE_PER_GPU, len(os.sched_getaffinity(0)))
        o = O.Oracle(cfg, cores)
        for which, n_emit in ((pmamd.PM_MAP_SURFACE, s.n_emit_surface), (pmamd.PM_MAP_VOLUME, s.n_emit_volume)):
            pos, pw, d = g.records(which)
            o.set_records(which, pos, pw, d, n_emit)
        rows = queries = exact = 0
        t_render = 0.0
        for y in range(cfg.height // 32, cfg.height, cfg.height // 16):
            t = time.perf_counter()
            row = o.render_voxel(0, y, cfg.width, 1)
            t_render += time.perf_counter() - t
            queries += sum(o.last_queries)
            exact += int(np.array_equal(row[0], imgs[1][y].astype(np.float64)))
            rows += 1
            if t_render > cpu_budget_s:
                break
        res["cpu_oracle"] = {"radiance_queries_per_s": round(queries / t_render, 1), "cores": cores,
                             "kind": "port", "rows": rows, "radiance_queries": queries,
                             "seconds": round(t_render, 2), "rows_bit_exact_vs_gpu": exact}
        log("voxel CPU oracle: %d rows, %.1f s, %d of them bit-exact against the GPU" % (rows, t_render, exact))
    g.close()
    return res


def read_profile(name):
    """A committed counter summary under profiles/ (None if absent)."""
    p = os.path.join(REPO, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:
        return None


# VALU issue peak of MI355X: 256 CUs x 4 SIMDs, one wave64 VALU instruction per
# SIMD every 4 cycles at the 2.4 GHz peak engine clock.
VALU_ISSUE_PEAK_PER_S = 256 * 4 * 2.4e9 / 4


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
    ap.add_argument("--cache-error-frames", type=int, default=4,
                    help="further untimed cache frames whose rows are also compared with the FP64 oracle")
    ap.add_argument("--cache-ema", type=int, default=950,
                    help="inference on the moving average of the weights, decay per mille per Adam step (0: off)")
    ap.add_argument("--cache-near", type=int, default=2,
                    help="near-light radius of the cache frame in merged march groups (PM_OPT_CACHE_NEAR_GROUPS)")
    ap.add_argument("--cache-thresholds", type=str, default="0.12,0.05",
                    help="fallback thresholds timed in turn (comma-separated)")
    ap.add_argument("--error-rows", type=int, default=16,
                    help="stratified frame rows rendered by the FP64 oracle for the error of the BF16 + cache frame")
    ap.add_argument("--batch-steps", type=int, default=3, help="timed batch frames (0: skip the batch leg)")
    ap.add_argument("--batch-tune-stride", type=int, default=4)
    ap.add_argument("--top-levels", type=int, default=0,
                    help="kd-tree levels the mode-7 render stages in LDS per wave (PM_OPT_TOP_LEVELS)")
    ap.add_argument("--voxel-steps", type=int, default=3,
                    help="timed frames of the voxel-grid compatibility mode (0: skip the leg)")
    ap.add_argument("--voxel-cpu-seconds", type=float, default=8.0)
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
    sec_per_frame, per, stats, img, g = run_gpu(cfg, args.steps, args.warmup, args.gather_mode, args.tile_regions,
                                                args.top_levels)
    q_frame = stats.surface_queries + stats.volume_queries
    value = q_frame / sec_per_frame
    kern_ms = float(np.mean([p[3] for p in per]))
    bpq = gather_bytes_per_query(cfg.k_volume, cfg.storage == pmamd.PM_STORE_BF16)
    achieved = bpq * q_frame / (kern_ms * 1e-3) / 1e9
    tr = read_profile("traffic_render_cfg2.json")
    vp = read_profile("valu_render_cfg2.json")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (tr.get("bytes_per_launch") if tr else None),
            "kernel": "k_render_lane7s (eye ray + march + kNN gather, a pixel's samples in adjacent lanes)",
            "bytes_per_query": bpq,
            "queries_per_launch": int(q_frame), "kernel_ms": round(kern_ms, 3)}
    if vp:
        # the roof the kernel actually sits on: VALU instruction issue
        roof["valu_issue"] = {"insts_per_launch": vp["sq_insts_valu"], "peak_insts_per_s": VALU_ISSUE_PEAK_PER_S,
                              "frac": round(vp["sq_insts_valu"] / (kern_ms * 1e-3) / VALU_ISSUE_PEAK_PER_S, 4),
                              "salu_insts_per_launch": vp.get("sq_insts_salu"),
                              "source": "profiles/valu_render_cfg2.json (rocprofv3 SQ_INSTS_VALU of the same kernel)"}
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
        "config": {"workload": "1080p workload, plain FP32 frame: " + workload_desc(cfg),
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
    g.close()
    H = cfg.height
    err_rows = [int((i + 0.5) * H / args.error_rows) for i in range(args.error_rows)]
    plain_rows = img.reshape(H, cfg.width, 3)[err_rows].copy()
    ns = None
    if args.cache_steps > 0:
        ths = [float(v) for v in args.cache_thresholds.split(",") if v]
        ns = run_cache(args.cache_steps, args.cache_train_frames, args.cache_stride, ths, args.cache_batch,
                           err_rows, args.cache_near, args.cache_error_frames, args.cache_ema)
        for st in ns["settings"]:
            st["speedup_vs_plain_fp32_frame"] = round(out["ms_per_step"] / st["ms_per_frame"], 3)
            ws = st["without_fallback_seeds"]
            ws["speedup_vs_plain_fp32_frame"] = round(out["ms_per_step"] / ws["ms_per_frame"], 3)
            st["rmse_vs_plain_fp32_same_samples"] = {
                k: v for k, v in row_errors(st["_rows"], plain_rows.astype(np.float64), cfg.exposure).items()
                if k.startswith("rmse")}
    if args.batch_steps > 0:
        out["batch"] = run_batch(args.batch_steps, args.batch_tune_stride)
    if args.voxel_steps > 0:
        out["voxel_grid"] = run_voxel(cfg, args.voxel_steps, 0.0 if args.no_cpu else args.voxel_cpu_seconds)
        out["voxel_grid"]["knn_render_ms_same_frame"] = out["breakdown_ms"]["render"]
    if not args.no_cpu:
        cb = run_cpu(cfg, args.cpu_seconds, err_rows if ns is not None else None)
        out["cpu_baseline"] = {
            "value": round(cb["qps"], 1), "unit": "radiance queries/s", "cores": cb["cores"], "kind": "port",
            "sample": "CPU oracle (g++ -O2, std::thread x %d, the host-core share of one GPU; the host has %d cores "
                      "in its affinity mask, %d in total, %s): full emit (%.1f s) and kd-build (%.1f s) of the "
                      "1080p workload, then render of %d stratified rows of its 1920x1080x4 frame (%d radiance "
                      "queries in %.1f s), extrapolated to the frame"
                      % (cb["cores"], cb["cores_affinity"], cb["cores_total"], cb["cpu"], cb["t_emit"], cb["t_build"],
                         cb["rows"], cb["queries"], cb["t_render"]),
            "est_ms_per_frame": round(q_frame / cb["qps"] * 1e3 + (cb["t_emit"] + cb["t_build"]) * 1e3, 1),
            "host_cores_affinity": cb["cores_affinity"], "host_cores_total": cb["cores_total"],
            "cpu_model": cb["cpu"],
            "unit_room_256_in_full": cb["cfg1"],
        }
        if cb["ref"] is not None:
            ref = cb["ref"]
            out["plain_fp32_vs_fp64_oracle"] = row_errors(plain_rows, ref, cfg.exposure)
            for st in ns["settings"]:
                st["error_vs_fp64_oracle"] = row_errors(st["_rows"], ref, cfg.exposure)
                if st["_more_rows"]:
                    d = [st["error_vs_fp64_oracle"]["rmse_display_rel"]] + [
                        row_errors(r, ref, cfg.exposure)["rmse_display_rel"] for r in st["_more_rows"]]
                    st["display_rmse_vs_fp64_over_frames"] = {"frames": len(d), "mean": round(float(np.mean(d)), 6),
                                                              "max": round(float(np.max(d)), 6),
                                                              "each": [round(float(x), 6) for x in d]}
            ns["fp64_reference"] = ("oracle FP64 render (FP32 photon records, FP64 eye paths, march and estimates) "
                                    "of %d stratified rows x 1920 px x 4 spp (%.1f s on the host)"
                                    % (len(err_rows), cb["t_ref"]))
    if ns is not None:
        for st in ns["settings"]:
            del st["_rows"]
            del st["_more_rows"]
        out["cache_bf16"] = ns
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
