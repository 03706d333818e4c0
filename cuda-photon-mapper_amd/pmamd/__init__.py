"""Python binding of the MI355X photon mapper's C ABI (libpm.so), via ctypes.

The library is the product: HIP kernels for gfx950 behind the extern "C"
entry points declared in include/photon_mapper.h.  This module only mirrors
the C structs and wraps the calls; it never computes anything itself and it
raises if libpm.so is missing (there is no fallback path).

Reference interface mirrored: the three launchers of photonMappingKernel.cu
(:1523 launch_emit_photons_kernel, :1549 launch_photon_mapping_kernel,
:1569 launch_init_random_numbers_kernel), called in the order of
callbacksPBO.cpp:55-63 (init random numbers once, emit, render).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_ROOT, "libpm.so")

PM_OK = 0
PM_ERR_INVALID = -1
PM_ERR_HIP = -2
PM_ERR_STATE = -3
PM_ERR_CAPACITY = -4
PM_ERR_NOMEM = -5

PM_MAX_SPHERES = 256
PM_MAT_DIFFUSE, PM_MAT_MIRROR, PM_MAT_DIELECTRIC = 0, 1, 2
PM_MEDIUM_NONE, PM_MEDIUM_HOMOGENEOUS, PM_MEDIUM_NOISE = 0, 1, 2
PM_STORE_FP32, PM_STORE_BF16 = 0, 1
PM_EST_BOX, PM_EST_CONE = 0, 1
PM_MAP_SURFACE, PM_MAP_VOLUME = 0, 1
PM_OPT_GATHER_MODE, PM_OPT_BLOCKS_PER_CU, PM_OPT_CACHE_FUSED_INFER, PM_OPT_CACHE_INFER_CHUNK = 1, 2, 3, 4
PM_OPT_TILE_REGIONS = 5
PM_OPT_SAMPLE_LANES = 6
PM_OPT_RENDER_MEDIA = 7
PM_OPT_VOXEL_GRID = 8
PM_OPT_TOP_LEVELS = 9
PM_OPT_CACHE_NEAR_GROUPS = 10
PM_OPT_CACHE_EMA = 11
PM_OPT_CACHE_SEEDS = 12
VOXEL_TABLE_ENTRIES = 3 * 33 ** 3  # summed-volume table of one map's 32^3 voxel grid


class PmSphere(ctypes.Structure):
    _fields_ = [
        ("center", ctypes.c_float * 3),
        ("radius", ctypes.c_float),
        ("albedo", ctypes.c_float * 3),
        ("material", ctypes.c_int),
        ("ior", ctypes.c_float),
    ]


class PmConfig(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint32),
        ("wall_albedo", (ctypes.c_float * 3) * 5),
        ("n_spheres", ctypes.c_int),
        ("spheres", PmSphere * PM_MAX_SPHERES),
        ("light_pos", ctypes.c_float * 3),
        ("light_power", ctypes.c_float * 3),
        ("medium", ctypes.c_int),
        ("sigma_s", ctypes.c_float),
        ("sigma_a", ctypes.c_float),
        ("hg_g", ctypes.c_float),
        ("noise_res", ctypes.c_int),
        ("noise_freq", ctypes.c_int),
        ("n_emit", ctypes.c_uint32),
        ("cap_surface", ctypes.c_uint32),
        ("cap_volume", ctypes.c_uint32),
        ("max_bounces", ctypes.c_int),
        ("rr_survive", ctypes.c_float),
        ("leaf_size", ctypes.c_int),
        ("storage", ctypes.c_int),
        ("k_surface", ctypes.c_int),
        ("k_volume", ctypes.c_int),
        ("rcap_surface", ctypes.c_float),
        ("rcap_volume", ctypes.c_float),
        ("estimator", ctypes.c_int),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("spp", ctypes.c_int),
        ("march_step", ctypes.c_float),
        ("max_depth", ctypes.c_int),
        ("eye", ctypes.c_float * 3),
        ("focal", ctypes.c_float),
        ("exposure", ctypes.c_float),
        ("fp64_transmittance", ctypes.c_int),
        ("media_on", ctypes.c_int),
        ("anim_time", ctypes.c_float),
    ]


class PmStats(ctypes.Structure):
    _fields_ = [
        ("n_surface", ctypes.c_uint32),
        ("n_volume", ctypes.c_uint32),
        ("n_emit_surface", ctypes.c_uint32),
        ("n_emit_volume", ctypes.c_uint32),
        ("n_emit_total", ctypes.c_uint32),
        ("pad0", ctypes.c_uint32),
        ("emitted", ctypes.c_double * 3),
        ("absorbed", ctypes.c_double * 3),
        ("escaped", ctypes.c_double * 3),
        ("rr_gain", ctypes.c_double * 3),
        ("ms_emit", ctypes.c_double),
        ("ms_build", ctypes.c_double),
        ("ms_render", ctypes.c_double),
        ("ms_gather_kernel", ctypes.c_double),
        ("surface_queries", ctypes.c_uint64),
        ("volume_queries", ctypes.c_uint64),
    ]


class PmCacheStats(ctypes.Structure):
    _fields_ = [
        ("records", ctypes.c_uint64),
        ("fallbacks", ctypes.c_uint64),
        ("train_loss", ctypes.c_double),
        ("ms_record", ctypes.c_double),
        ("ms_train", ctypes.c_double),
        ("ms_infer", ctypes.c_double),
        ("ms_fallback", ctypes.c_double),
        ("ms_accumulate", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("frame", ctypes.c_uint32),
        ("inferred_volume", ctypes.c_uint32),
        ("inferred", ctypes.c_uint64),
    ]


class PmBatchTuning(ctypes.Structure):
    _fields_ = [
        ("leaf_size", ctypes.c_int),
        ("block_threads", ctypes.c_int),
        ("heap_slots", ctypes.c_int),
        ("heap_key_bytes", ctypes.c_int),
        ("gather_mode", ctypes.c_int),
        ("top_levels", ctypes.c_int),
    ]


class PmBatchStats(ctypes.Structure):
    _fields_ = [
        ("n_scenes", ctypes.c_uint32),
        ("work_items", ctypes.c_uint32),
        ("photons", ctypes.c_uint64),
        ("surface_records", ctypes.c_uint64),
        ("volume_records", ctypes.c_uint64),
        ("pixels", ctypes.c_uint64),
        ("surface_queries", ctypes.c_uint64),
        ("volume_queries", ctypes.c_uint64),
        ("ms_frame", ctypes.c_double),
        ("ms_setup", ctypes.c_double),
        ("tuning", PmBatchTuning),
        ("graph_nodes", ctypes.c_uint32),
        ("pad0", ctypes.c_uint32),
    ]


class PmError(RuntimeError):
    def __init__(self, status, what):
        super().__init__("%s failed with status %d" % (what, status))
        self.status = status


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "launch_init_random_numbers_kernel": (None, []),
    "launch_emit_photons_kernel": (None, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float,
                                          ctypes.c_bool, ctypes.c_bool]),
    "launch_photon_mapping_kernel": (None, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float,
                                            ctypes.c_bool, ctypes.c_bool]),
    "pm_config_preset": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(PmConfig)]),
    "pm_create": (ctypes.c_int, [ctypes.POINTER(PmConfig), ctypes.POINTER(ctypes.c_void_p)]),
    "pm_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pm_emit": (ctypes.c_int, [ctypes.c_void_p]),
    "pm_build": (ctypes.c_int, [ctypes.c_void_p]),
    "pm_render": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pm_render_crop": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p]),
    "pm_render_crop_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "pm_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PmStats)]),
    "pm_set_render": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_float, ctypes.c_float]),
    "pm_get_records": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_uint32]),
    "pm_set_records": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "pm_get_tree": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "pm_get_voxel_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]),
    "pm_legacy_context": (ctypes.c_void_p, []),
    "pm_knn": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                              ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pm_device_alloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    "pm_device_free": (ctypes.c_int, [ctypes.c_void_p]),
    "pm_memcpy_htod": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pm_memcpy_dtoh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pm_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "pm_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "pm_copy_kernel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pm_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "pm_cache_param_count": (ctypes.c_uint32, []),
    "pm_cache_reset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "pm_cache_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int]),
    "pm_cache_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                     ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "pm_cache_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_float, ctypes.POINTER(ctypes.c_double)]),
    "pm_cache_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "pm_cache_configure": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_float, ctypes.c_int, ctypes.c_int]),
    "pm_cache_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PmCacheStats)]),
    "pm_render_cache_crop": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "pm_batch_scene_config": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(PmConfig)]),
    "pm_batch_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "pm_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pm_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PmBatchTuning)]),
    "pm_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(PmBatchTuning),
                                         ctypes.POINTER(ctypes.c_double)]),
    "pm_batch_submit": (ctypes.c_int, [ctypes.c_void_p]),
    "pm_batch_time_phases": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "pm_batch_image": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "pm_batch_scene_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(PmStats)]),
    "pm_batch_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PmBatchStats)]),
    "pm_batch_get_records": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    "pm_batch_get_tree": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
}


def exported_symbols():
    """Names of every entry point include/photon_mapper.h declares."""
    return list(_SIGS)


def load(path=LIB_PATH):
    """Load libpm.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError("libpm.so not built: %s (run __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what):
    if status != PM_OK:
        lib = load()
        msg = lib.pm_status_string(status)
        raise PmError(status, "%s (%s)" % (what, msg.decode() if msg else "?"))


def preset(n):
    cfg = PmConfig()
    check(load().pm_config_preset(n, ctypes.byref(cfg)), "pm_config_preset")
    return cfg


def device_count():
    n = ctypes.c_int(0)
    st = load().pm_device_count(ctypes.byref(n))
    return n.value if st == PM_OK else 0


class DeviceBuffer:
    """A device allocation made through the library (hipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(load().pm_device_alloc(ctypes.byref(p), max(self.nbytes, 1)), "pm_device_alloc")
        self.ptr = p.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(load().pm_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "pm_memcpy_htod")

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(load().pm_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "pm_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            load().pm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PhotonMapper:
    """One library context: emit -> build -> render, plus inspection."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(load().pm_create(ctypes.byref(cfg), ctypes.byref(h)), "pm_create")
        self.h = h.value

    def close(self):
        if self.h:
            load().pm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, option, value):
        check(load().pm_set_option(self.h, option, value), "pm_set_option")

    def emit(self):
        check(load().pm_emit(self.h), "pm_emit")

    def build(self):
        check(load().pm_build(self.h), "pm_build")

    def stats(self):
        s = PmStats()
        check(load().pm_stats_get(self.h, ctypes.byref(s)), "pm_stats_get")
        return s

    def set_render(self, width, height, spp, k_s, k_v, rcap_s, rcap_v):
        check(load().pm_set_render(self.h, width, height, spp, k_s, k_v, rcap_s, rcap_v), "pm_set_render")

    def render(self, x0=0, y0=0, w=0, h=0, want_u8=False):
        if w <= 0:
            w, h, x0, y0 = self.cfg.width, self.cfg.height, 0, 0
        rgb = DeviceBuffer(w * h * 3 * 4)
        u8 = DeviceBuffer(w * h * 4) if want_u8 else None
        check(load().pm_render_crop(self.h, x0, y0, w, h, rgb.ptr, u8.ptr if u8 else None), "pm_render_crop")
        out = rgb.download(np.float32, w * h * 3).reshape(h, w, 3)
        if want_u8:
            return out, u8.download(np.uint8, w * h * 4).reshape(h, w, 4)
        return out

    def set_records(self, which, pos, power, dirs, n_emit):
        """Replace a map's records (host arrays [n, 3]); build() must follow."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        power = np.ascontiguousarray(power, np.float32).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        assert pos.shape == power.shape == dirs.shape
        check(load().pm_set_records(self.h, which, pos.ctypes.data, power.ctypes.data, dirs.ctypes.data,
                                    pos.shape[0], int(n_emit)), "pm_set_records")

    def records(self, which):
        s = self.stats()
        n = s.n_surface if which == PM_MAP_SURFACE else s.n_volume
        pos = np.empty((n, 3), np.float32)
        pw = np.empty((n, 3), np.float32)
        d = np.empty((n, 3), np.float32)
        check(load().pm_get_records(self.h, which, pos.ctypes.data, pw.ctypes.data, d.ctypes.data, n),
              "pm_get_records")
        return pos, pw, d

    def tree(self, which):
        s = self.stats()
        n = s.n_surface if which == PM_MAP_SURFACE else s.n_volume
        L = self.cfg.leaf_size
        nl = (n + L - 1) // L if n else 0
        ni = max(nl - 1, 0)
        perm = np.empty(n, np.uint32)
        split = np.empty(ni, np.float32)
        axis = np.empty(ni, np.uint32)
        check(load().pm_get_tree(self.h, which, perm.ctypes.data, split.ctypes.data, axis.ctypes.data, n, ni),
              "pm_get_tree")
        return perm, split, axis

    def voxel_table(self, which):
        """Summed-volume table [33, 33, 33, 3] (z, y, x, channel) of a map's voxel grid, uint64."""
        t = np.empty(VOXEL_TABLE_ENTRIES, np.uint64)
        check(load().pm_get_voxel_table(self.h, which, t.ctypes.data, t.size), "pm_get_voxel_table")
        return t.reshape(33, 33, 33, 3)

    # ---- learned radiance cache
    def cache_reset(self, seed):
        check(load().pm_cache_reset(self.h, seed), "pm_cache_reset")

    def cache_params(self, values=None):
        n = load().pm_cache_param_count()
        if values is None:
            out = np.empty(n, np.float32)
            check(load().pm_cache_params(self.h, out.ctypes.data, n, 0), "pm_cache_params")
            return out
        v = np.ascontiguousarray(values, np.float32)
        assert v.size == n
        check(load().pm_cache_params(self.h, v.ctypes.data, n, 1), "pm_cache_params")

    def cache_grad(self, inputs, targets):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        xb, tb = DeviceBuffer(x.nbytes), DeviceBuffer(t.nbytes)
        xb.upload(x)
        tb.upload(t)
        n = load().pm_cache_param_count()
        g = np.empty(n, np.float32)
        loss = ctypes.c_double()
        check(load().pm_cache_grad(self.h, xb.ptr, tb.ptr, x.shape[0], g.ctypes.data, n, ctypes.byref(loss)),
              "pm_cache_grad")
        return g, loss.value

    def cache_train(self, inputs, targets, batch, lr):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        xb, tb = DeviceBuffer(x.nbytes), DeviceBuffer(t.nbytes)
        xb.upload(x)
        tb.upload(t)
        loss = ctypes.c_double()
        check(load().pm_cache_train(self.h, xb.ptr, tb.ptr, x.shape[0], batch, lr, ctypes.byref(loss)),
              "pm_cache_train")
        return loss.value

    def cache_configure(self, threshold=0.05, train_samples=1 << 18, batch=1 << 14, lr=1e-3, march_stride=1,
                        cache_surface=0):
        check(load().pm_cache_configure(self.h, threshold, train_samples, batch, lr, march_stride, cache_surface),
              "pm_cache_configure")

    def cache_stats(self):
        s = PmCacheStats()
        check(load().pm_cache_stats_get(self.h, ctypes.byref(s)), "pm_cache_stats_get")
        return s

    def render_cache(self, x0=0, y0=0, w=0, h=0, want_u8=False):
        if w <= 0:
            w, h, x0, y0 = self.cfg.width, self.cfg.height, 0, 0
        rgb = DeviceBuffer(w * h * 3 * 4)
        u8 = DeviceBuffer(w * h * 4) if want_u8 else None
        check(load().pm_render_cache_crop(self.h, x0, y0, w, h, rgb.ptr, u8.ptr if u8 else None),
              "pm_render_cache_crop")
        out = rgb.download(np.float32, w * h * 3).reshape(h, w, 3)
        if want_u8:
            return out, u8.download(np.uint8, w * h * 4).reshape(h, w, 4)
        return out

    def cache_predict(self, inputs):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        xb, ob = DeviceBuffer(max(x.nbytes, 4)), DeviceBuffer(max(x.shape[0] * 12, 4))
        xb.upload(x)
        check(load().pm_cache_predict(self.h, xb.ptr, x.shape[0], ob.ptr), "pm_cache_predict")
        return ob.download(np.float32, x.shape[0] * 3).reshape(-1, 3)

    def knn(self, which, queries, k, rcap):
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, 3)
        nq = q.shape[0]
        qb = DeviceBuffer(max(q.nbytes, 4))
        qb.upload(q)
        ib = DeviceBuffer(nq * k * 4)
        db = DeviceBuffer(nq * k * 4)
        cb = DeviceBuffer(nq * 4)
        check(load().pm_knn(self.h, which, qb.ptr, nq, k, rcap, ib.ptr, db.ptr, cb.ptr), "pm_knn")
        return (ib.download(np.uint32, nq * k).reshape(nq, k), db.download(np.float32, nq * k).reshape(nq, k),
                cb.download(np.uint32, nq))


def batch_scene(seed, index):
    """Configuration of scene `index` of the batched workload drawn from `seed`."""
    cfg = PmConfig()
    check(load().pm_batch_scene_config(seed, index, ctypes.byref(cfg)), "pm_batch_scene_config")
    return cfg


class Batch:
    """Many scenes rendered per frame by one hipGraph launch (pm_batch_*)."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        arr = (PmConfig * len(self.cfgs))(*self.cfgs)
        h = ctypes.c_void_p()
        check(load().pm_batch_create(arr, len(self.cfgs), ctypes.byref(h)), "pm_batch_create")
        self.h = h.value

    def close(self):
        if self.h:
            load().pm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, leaf_size, block_threads, heap_slots, heap_key_bytes=2, gather_mode=0, top_levels=0):
        t = PmBatchTuning(leaf_size, block_threads, heap_slots, heap_key_bytes, gather_mode, top_levels)
        check(load().pm_batch_set_tuning(self.h, ctypes.byref(t)), "pm_batch_set_tuning")

    def autotune(self, sample_stride=16):
        t = PmBatchTuning()
        ms = ctypes.c_double()
        check(load().pm_batch_autotune(self.h, sample_stride, ctypes.byref(t), ctypes.byref(ms)), "pm_batch_autotune")
        return (t.leaf_size, t.block_threads, t.heap_slots, t.heap_key_bytes, t.gather_mode, t.top_levels), ms.value

    def submit(self):
        check(load().pm_batch_submit(self.h), "pm_batch_submit")

    def time_phases(self):
        """One frame launched kernel by kernel: (emit, build, render) in ms."""
        t = [ctypes.c_double() for _ in range(3)]
        check(load().pm_batch_time_phases(self.h, *[ctypes.byref(x) for x in t]), "pm_batch_time_phases")
        return tuple(x.value for x in t)

    def stats(self):
        s = PmBatchStats()
        check(load().pm_batch_stats_get(self.h, ctypes.byref(s)), "pm_batch_stats_get")
        return s

    def scene_stats(self, i):
        s = PmStats()
        check(load().pm_batch_scene_stats(self.h, i, ctypes.byref(s)), "pm_batch_scene_stats")
        return s

    def image(self, i):
        c = self.cfgs[i]
        n = c.width * c.height
        rgb = DeviceBuffer(n * 12)
        check(load().pm_batch_image(self.h, i, rgb.ptr, None), "pm_batch_image")
        return rgb.download(np.float32, n * 3).reshape(c.height, c.width, 3)

    def records(self, i, which):
        s = self.scene_stats(i)
        n = s.n_surface if which == PM_MAP_SURFACE else s.n_volume
        pos = np.empty((n, 3), np.float32)
        pw = np.empty((n, 3), np.float32)
        d = np.empty((n, 3), np.float32)
        check(load().pm_batch_get_records(self.h, i, which, pos.ctypes.data, pw.ctypes.data, d.ctypes.data, n),
              "pm_batch_get_records")
        return pos, pw, d

    def tree(self, i, which, leaf_size):
        s = self.scene_stats(i)
        n = s.n_surface if which == PM_MAP_SURFACE else s.n_volume
        nl = (n + leaf_size - 1) // leaf_size
        ni = max(nl - 1, 0)
        perm = np.empty(n, np.uint32)
        split = np.empty(max(ni, 1), np.float32)
        axis = np.empty(max(ni, 1), np.uint32)
        check(load().pm_batch_get_tree(self.h, i, which, perm.ctypes.data, split.ctypes.data, axis.ctypes.data, n, ni),
              "pm_batch_get_tree")
        return perm, split[:ni], axis[:ni]
