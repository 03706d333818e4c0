"""ctypes wrapper of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product path.

Each entry point restates part of the reference's photon-mapper pipeline
(photonMappingKernel.cu: emitPhotons :1215-1375, gatherPhotons :586-600,
integrateVolumePhotons :831-870, computePixelColor :926-1017) for the new
method; see oracle/pm_oracle.cpp for the per-function citations.
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_REPO, "cuda-photon-mapper_amd"))
from pmamd import PmConfig  # noqa: E402  (struct layout only)

# ORACLE_LIB selects another build of the same source (the sanitizer build).
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_SIGS = {
    "orc_config_preset": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(PmConfig)]),
    "orc_create": (ctypes.c_void_p, [ctypes.POINTER(PmConfig), ctypes.c_int]),
    "orc_destroy": (None, [_P]),
    "orc_set_render": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_float]),
    "orc_emit": (ctypes.c_int, [_P]),
    "orc_counts": (ctypes.c_int, [_P, _P]),
    "orc_tally": (ctypes.c_int, [_P, _P]),
    "orc_get_records": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_uint32]),
    "orc_set_records": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_uint32, ctypes.c_uint32]),
    "orc_build": (ctypes.c_int, [_P]),
    "orc_get_tree": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_uint32, ctypes.c_uint32]),
    "orc_knn": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                               _P, _P, _P]),
    "orc_render": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _P, _P]),
    "orc_philox": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32]),
    "orc_math": (ctypes.c_int, [ctypes.c_int, _P, _P, ctypes.c_uint32]),
    "orc_sample": (ctypes.c_int, [ctypes.c_int, _P, _P, _P, ctypes.c_uint32]),
    "orc_intersect": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, ctypes.c_uint32]),
    "orc_fresnel": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "orc_phase_hg": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "orc_density": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32]),
    "orc_noise_grid": (ctypes.c_int, [_P, _P, ctypes.c_uint32]),
    "orc_collision": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_int, _P, _P, ctypes.c_float, _P]),
    "orc_kd_leaves_under": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
    "orc_gather": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_int, _P]),
    "orc_tonemap": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_float, _P]),
    "orc_set_density": (ctypes.c_int, [_P, _P, ctypes.c_uint32]),
    "orc_march": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _P,
                                 _P, _P, _P]),
    "orc_render_const": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P, _P]),
    "orc_camera_ray": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _P, _P, _P]),
    "orc_validate": (ctypes.c_int, [ctypes.POINTER(PmConfig)]),
    "orc_voxel_tables": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P]),
    "orc_voxel_query": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, ctypes.c_uint32,
                                       _P]),
    "orc_render_voxel": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, _P, _P]),
}


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("oracle not built: %s" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for n, (r, a) in _SIGS.items():
            f = getattr(lib, n)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data


def preset(n):
    c = PmConfig()
    assert load().orc_config_preset(n, ctypes.byref(c)) == 0
    return c


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
    out = np.empty_like(ctr)
    load().orc_philox(_ptr(ctr), _ptr(key), _ptr(out), ctr.shape[0])
    return out


def math(fn, x):
    """fn: 0 sin(2 pi x), 1 cos(2 pi x), 2 log, 3 exp (the project's FP32 functions)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    assert load().orc_math(fn, _ptr(x), _ptr(y), x.size) == 0
    return y


def sample(kind, param, u):
    p = np.zeros(4, np.float32)
    p[: len(param)] = param
    u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
    out = np.empty((u.shape[0], 3), np.float32)
    load().orc_sample(kind, _ptr(p), _ptr(u), _ptr(out), u.shape[0])
    return out


def fresnel(cos_i, eta):
    return load().orc_fresnel(cos_i, eta)


def phase_hg(c, g):
    return load().orc_phase_hg(c, g)


def kd_leaves_under(c, nl):
    return load().orc_kd_leaves_under(c, nl)


def tonemap(v, exposure):
    """The frame-write rule of the library's headers (tonemap_u8) on FP32 values."""
    v = np.ascontiguousarray(v, np.float32)
    out = np.empty(v.shape, np.uint8)
    assert load().orc_tonemap(_ptr(v), v.size, exposure, _ptr(out)) == 0
    return out


def validate(cfg):
    return load().orc_validate(ctypes.byref(cfg))


class Oracle:
    def __init__(self, cfg, nthreads=None):
        if nthreads is None:
            nthreads = min(16, len(os.sched_getaffinity(0)))
        self.cfg = cfg
        self.nthreads = nthreads
        self.h = load().orc_create(ctypes.byref(cfg), nthreads)
        if not self.h:
            raise ValueError("invalid config")

    def __del__(self):
        try:
            if self.h:
                load().orc_destroy(self.h)
        except Exception:
            pass

    def set_render(self, w, h, spp, ks, kv, rs, rv):
        assert load().orc_set_render(self.h, w, h, spp, ks, kv, rs, rv) == 0
        self.cfg.width, self.cfg.height, self.cfg.spp = w, h, spp
        self.cfg.k_surface, self.cfg.k_volume = ks, kv
        self.cfg.rcap_surface, self.cfg.rcap_volume = rs, rv

    def emit(self):
        st = load().orc_emit(self.h)
        if st != 0:
            raise RuntimeError("orc_emit %d" % st)

    def counts(self):
        c = np.zeros(5, np.uint32)
        load().orc_counts(self.h, _ptr(c))
        return dict(n_surface=int(c[0]), n_volume=int(c[1]), n_emit_surface=int(c[2]), n_emit_volume=int(c[3]),
                    n_emit_total=int(c[4]))

    def tally(self):
        t = np.zeros(12, np.float64)
        load().orc_tally(self.h, _ptr(t))
        return dict(emitted=t[0:3], absorbed=t[3:6], escaped=t[6:9], rr_gain=t[9:12])

    def records(self, which):
        c = self.counts()
        n = c["n_surface"] if which == 0 else c["n_volume"]
        pos = np.empty((n, 3), np.float32)
        pw = np.empty((n, 3), np.float32)
        d = np.empty((n, 3), np.float32)
        assert load().orc_get_records(self.h, which, _ptr(pos), _ptr(pw), _ptr(d), n) == 0
        return pos, pw, d

    def set_records(self, which, pos, power, dirs, n_emit):
        pos = np.ascontiguousarray(pos, np.float32)
        power = np.ascontiguousarray(power, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        load().orc_set_records(self.h, which, _ptr(pos), _ptr(power), _ptr(dirs), pos.shape[0], n_emit)

    def build(self):
        load().orc_build(self.h)

    def tree(self, which):
        c = self.counts()
        n = c["n_surface"] if which == 0 else c["n_volume"]
        L = self.cfg.leaf_size
        nl = (n + L - 1) // L if n else 0
        ni = max(nl - 1, 0)
        perm = np.empty(n, np.uint32)
        split = np.empty(ni, np.float32)
        axis = np.empty(ni, np.uint32)
        assert load().orc_get_tree(self.h, which, _ptr(perm), _ptr(split), _ptr(axis), n, ni) == 0
        return perm, split, axis

    def knn(self, which, q, k, rcap, brute=False):
        q = np.ascontiguousarray(q, np.float32).reshape(-1, 3)
        nq = q.shape[0]
        idx = np.empty((nq, k), np.uint32)
        d2 = np.empty((nq, k), np.float32)
        cnt = np.empty(nq, np.uint32)
        st = load().orc_knn(self.h, which, _ptr(q), nq, k, rcap, 1 if brute else 0, _ptr(idx), _ptr(d2), _ptr(cnt))
        assert st == 0, st
        return idx, d2, cnt

    def render(self, x0=0, y0=0, w=0, h=0, fp64=False, bf16=False):
        if w <= 0:
            w, h, x0, y0 = self.cfg.width, self.cfg.height, 0, 0
        out = np.empty((h, w, 3), np.float64)
        nq = np.zeros(2, np.uint64)
        st = load().orc_render(self.h, x0, y0, w, h, 1 if fp64 else 0, 1 if bf16 else 0, _ptr(out), _ptr(nq))
        assert st == 0, st
        self.last_queries = (int(nq[0]), int(nq[1]))
        return out

    def intersect(self, o, d):
        o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        n = o.shape[0]
        t = np.empty(n, np.float32)
        kind = np.empty(n, np.int32)
        idx = np.empty(n, np.int32)
        load().orc_intersect(self.h, _ptr(o), _ptr(d), _ptr(t), _ptr(kind), _ptr(idx), n)
        return t, kind, idx

    def density(self, p):
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        out = np.empty(p.shape[0], np.float32)
        load().orc_density(self.h, _ptr(p), _ptr(out), p.shape[0])
        return out

    def noise_grid(self):
        r = self.cfg.noise_res
        out = np.empty(r ** 3, np.float32)
        assert load().orc_noise_grid(self.h, _ptr(out), out.size) == 0
        return out.reshape(r, r, r)

    def gather(self, which, x, v, albedo=(1.0, 1.0, 1.0), fp64=False):
        x = np.ascontiguousarray(x, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        a = np.ascontiguousarray(albedo, np.float32)
        out = np.zeros(3, np.float64)
        st = load().orc_gather(self.h, which, _ptr(x), _ptr(v), _ptr(a), 1 if fp64 else 0, _ptr(out))
        assert st == 0, st
        return out

    def set_density(self, grid):
        g = np.ascontiguousarray(grid, np.float32).ravel()
        assert load().orc_set_density(self.h, _ptr(g), g.size) == 0

    def march(self, o, d, t0, t1, xi, S, thr=(1.0, 1.0, 1.0), fp64=False):
        """march_segment with a constant in-scatter S: (L, outgoing throughput, queries)."""
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        S = np.ascontiguousarray(S, np.float32)
        thr = np.ascontiguousarray(thr, np.float32)
        L = np.empty(3, np.float32)
        tout = np.empty(3, np.float32)
        nq = np.zeros(1, np.uint32)
        st = load().orc_march(self.h, 1 if fp64 else 0, _ptr(o), _ptr(d), t0, t1, xi, _ptr(S), _ptr(thr), _ptr(L),
                              _ptr(tout), _ptr(nq))
        assert st == 0, st
        return L, tout, int(nq[0])

    def voxel_tables(self, bf16=False):
        """(grid_s, grid_v, table_s, table_v): uint64 [32,32,32,3] grids and [33,33,33,3] tables (z, y, x, ch)."""
        gs = np.empty((32, 32, 32, 3), np.uint64)
        gv = np.empty_like(gs)
        ts = np.empty((33, 33, 33, 3), np.uint64)
        tv = np.empty_like(ts)
        assert load().orc_voxel_tables(self.h, 1 if bf16 else 0, _ptr(gs), _ptr(gv), _ptr(ts), _ptr(tv)) == 0
        return gs, gv, ts, tv

    def voxel_query(self, kind, x, n=None, albedo=None, interp=False, bf16=False):
        """Voxel-grid estimate at points x (kind 0: surface with normals n and albedo, 1: volume), FP64 [n, 3]."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 3)
        m = x.shape[0]
        n = np.ascontiguousarray(np.zeros((m, 3)) if n is None else n, np.float32).reshape(-1, 3)
        albedo = np.ascontiguousarray(np.ones((m, 3)) if albedo is None else albedo, np.float32).reshape(-1, 3)
        out = np.empty((m, 3), np.float64)
        assert load().orc_voxel_query(self.h, kind, 1 if interp else 0, 1 if bf16 else 0, _ptr(x), _ptr(n),
                                      _ptr(albedo), m, _ptr(out)) == 0
        return out

    def render_voxel(self, x0=0, y0=0, w=0, h=0, interp=False, bf16=False):
        """FP32 render with the voxel-grid estimate (values as doubles), [h, w, 3]."""
        if w <= 0:
            w, h, x0, y0 = self.cfg.width, self.cfg.height, 0, 0
        out = np.empty((h, w, 3), np.float64)
        nq = np.zeros(2, np.uint64)
        assert load().orc_render_voxel(self.h, x0, y0, w, h, 1 if interp else 0, 1 if bf16 else 0, _ptr(out),
                                       _ptr(nq)) == 0
        self.last_queries = (int(nq[0]), int(nq[1]))
        return out

    def render_const(self, px, py, s, Sv, Ss):
        """One eye sample with a constant gather: (radiance, (surface, volume) queries)."""
        Sv = np.ascontiguousarray(Sv, np.float32)
        Ss = np.ascontiguousarray(Ss, np.float32)
        out = np.empty(3, np.float32)
        nq = np.zeros(2, np.uint32)
        assert load().orc_render_const(self.h, px, py, s, _ptr(Sv), _ptr(Ss), _ptr(out), _ptr(nq)) == 0
        return out, (int(nq[0]), int(nq[1]))

    def camera_ray(self, px, py, s):
        o = np.empty(3, np.float32)
        d = np.empty(3, np.float32)
        xi = ctypes.c_float()
        load().orc_camera_ray(self.h, px, py, s, _ptr(o), _ptr(d), ctypes.byref(xi))
        return o, d, xi.value

    def collision(self, pid, ev, o, d, tmax):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        t = ctypes.c_float()
        hit = load().orc_collision(self.h, pid, ev, _ptr(o), _ptr(d), tmax, ctypes.byref(t))
        return (t.value if hit else None)
