"""Pins of the oracle's voxel-grid estimate (the reference's 32^3 grid method,
photonMappingKernel.cu:260-267, :314-389, :442-583, :831-870, :1147-1183).

Each check is independent of the oracle's C++: the grids and summed-volume
tables against NumPy scatter-adds and cumulative sums, the window estimates
against a brute-force sum over the records in NumPy, and closed forms --
a uniform wall or medium gives the same estimate in every window, clamped or
not; a wall flux linear along one axis gives its value at the voxel corner
(floor(32 u) / 32) without the blend and at u - 1/64 with it; and the whole
render of a uniform scene equals the render with a constant gather.
"""
import math

import numpy as np
import pytest

import oracle as O

FIX = 2.0 ** 24
V = 32


def vox(p):
    """Voxel of coordinates in [0, 1] (FP32 product, truncation, clamped)."""
    v = (np.asarray(p, np.float32) * np.float32(V)).astype(np.int64)
    return np.clip(v, 0, V - 1)


def fixed(p):
    return np.floor(np.asarray(p, np.float32).astype(np.float64) * FIX + 0.5).astype(np.uint64)


def bf16_round(a):
    """FP32 -> BF16 -> FP32, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def grid_numpy(pos, power):
    g = np.zeros((V, V, V, 3), np.uint64)
    ix = vox(pos)
    q = fixed(power)
    for ch in range(3):
        np.add.at(g[..., ch], (ix[:, 2], ix[:, 1], ix[:, 0]), q[:, ch])
    return g


def table_numpy(g):
    t = np.zeros((V + 1, V + 1, V + 1, 3), np.uint64)
    t[1:, 1:, 1:] = g.cumsum(0, dtype=np.uint64).cumsum(1, dtype=np.uint64).cumsum(2, dtype=np.uint64)
    return t


def config(n_spheres=2, albedo=None):
    c = O.preset(1)
    c.n_spheres = n_spheres
    if albedo is not None:
        for w in range(5):
            for ch in range(3):
                c.wall_albedo[w][ch] = albedo
    c.width, c.height, c.spp = 48, 32, 1
    return c


def scale(c, n_emit):
    return np.array([np.float32(c.light_power[k]) / np.float32(n_emit) for k in range(3)], np.float64)


def random_records(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3), dtype=np.float32)
    power = (0.5 + rng.random((n, 3))).astype(np.float32)
    d = rng.normal(size=(n, 3))
    dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return pos, power, dirs


def wall_records(n, seed):
    """Records on the 5 walls of the room (x=1, y=0, x=0, y=1, z=1)."""
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3), dtype=np.float32)
    face = rng.integers(0, 5, n)
    axis = np.array([0, 1, 0, 1, 2])[face]
    val = np.array([1, 0, 0, 1, 1], np.float32)[face]
    pos[np.arange(n), axis] = val
    power = (0.5 + rng.random((n, 3))).astype(np.float32)
    return pos, power, np.tile(np.float32([0, 0, 1]), (n, 1))


@pytest.mark.parametrize("bf16", [False, True])
def test_grids_and_tables_match_numpy(bf16, ncpu):
    c = config()
    o = O.Oracle(c, ncpu)
    ps, ws, ds = random_records(20000, 1)
    pv, wv, dv = random_records(15000, 2)
    # coordinates on the voxel boundaries and the room's faces
    ps[:64] = np.float32(np.arange(64)[:, None] % 33) / np.float32(32)
    o.set_records(0, ps, ws, ds, 40000)
    o.set_records(1, pv, wv, dv, 40000)
    gs, gv, ts, tv = o.voxel_tables(bf16=bf16)
    rd = bf16_round if bf16 else (lambda a: a)
    assert np.array_equal(gs, grid_numpy(ps, rd(ws)))
    assert np.array_equal(gv, grid_numpy(pv, rd(wv)))
    assert np.array_equal(ts, table_numpy(gs))
    assert np.array_equal(tv, table_numpy(gv))


def surface_brute(pos, power, c, n_emit, x, axis, albedo, interp=False):
    """The window estimate at wall point x by summing the records in NumPy."""
    b, cc = (1, 2) if axis == 0 else ((0, 2) if axis == 1 else (0, 1))
    layer = V - 1 if x[axis] > 0.5 else 0
    ix = vox(pos)
    q = fixed(power).astype(np.float64)  # sums stay below 2^53 here
    on = ix[:, axis] == layer
    S = scale(c, n_emit)

    def window(vb, vc):
        b0, b1 = max(vb - 3, 0), min(vb + 3, V)
        c0, c1 = max(vc - 3, 0), min(vc + 3, V)
        m = on & (ix[:, b] >= b0) & (ix[:, b] < b1) & (ix[:, cc] >= c0) & (ix[:, cc] < c1)
        area = (b1 - b0) * (c1 - c0) / V ** 2
        return q[m].sum(0) / FIX * S / area

    if not interp:
        e = window(int(vox(x[b])), int(vox(x[cc])))
    else:
        ub, uc = float(x[b]) * V - 0.5, float(x[cc]) * V - 0.5
        ib, ic = math.floor(ub), math.floor(uc)
        fb, fc = min(max(ub - ib, 0), 1), min(max(uc - ic, 0), 1)
        cl = lambda v: min(max(v, 0), V - 1)
        w00, w10 = window(cl(ib), cl(ic)), window(cl(ib + 1), cl(ic))
        w01, w11 = window(cl(ib), cl(ic + 1)), window(cl(ib + 1), cl(ic + 1))
        lo, hi = w00 + fb * (w10 - w00), w01 + fb * (w11 - w01)
        e = lo + fc * (hi - lo)
    return np.asarray(albedo, np.float64) / math.pi * e


@pytest.mark.parametrize("interp", [False, True])
def test_surface_window_matches_brute_force(interp, ncpu):
    c = config()
    o = O.Oracle(c, ncpu)
    pos, power, dirs = wall_records(30000, 3)
    n_emit = 50000
    o.set_records(0, pos, power, dirs, n_emit)
    o.set_records(1, *random_records(10, 4), n_emit)
    rng = np.random.default_rng(5)
    faces = [(0, 1.0, -1), (1, 0.0, 1), (0, 0.0, 1), (1, 1.0, -1), (2, 1.0, -1)]
    xs, ns, albs, axes = [], [], [], []
    for axis, val, sign in faces:
        for _ in range(40):
            x = rng.random(3).astype(np.float32)
            x[axis] = val
            nrm = np.zeros(3, np.float32)
            nrm[axis] = sign
            xs.append(x); ns.append(nrm); axes.append(axis)
            albs.append(rng.random(3).astype(np.float32))
    # corners and edges of a wall: clamped windows and blends
    for x in ([1.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0, 0.01, 0.99], [1.0, 0.5, 0.015625]):
        xs.append(np.float32(x)); ns.append(np.float32([-1, 0, 0])); axes.append(0)
        albs.append(np.float32([1, 1, 1]))
    got = o.voxel_query(0, np.array(xs), np.array(ns), np.array(albs), interp=interp)
    for i, (x, a, al) in enumerate(zip(xs, axes, albs)):
        want = surface_brute(pos, power, c, n_emit, x, a, al, interp)
        np.testing.assert_allclose(got[i], want, rtol=1e-12, atol=0)


def test_surface_estimate_zero_off_walls(ncpu):
    """A normal that is not along an axis (a sphere's) gets no estimate."""
    c = config()
    o = O.Oracle(c, ncpu)
    pos, power, dirs = wall_records(1000, 6)
    o.set_records(0, pos, power, dirs, 1000)
    n = np.float32([[0.6, 0.8, 0.0], [0.0, 0.6, -0.8]])
    got = o.voxel_query(0, np.float32([[0.5, 0.5, 0.5]] * 2), n, np.ones((2, 3)))
    assert (got == 0).all()


def test_volume_window_matches_brute_force(ncpu):
    c = config()
    o = O.Oracle(c, ncpu)
    pv, wv, dv = random_records(40000, 7)
    n_emit = 70000
    o.set_records(1, pv, wv, dv, n_emit)
    o.set_records(0, *random_records(10, 8), n_emit)
    rng = np.random.default_rng(9)
    xs = rng.random((200, 3)).astype(np.float32)
    xs[:8] = np.float32([[0, 0, 0], [1, 1, 1], [0.02, 0.5, 0.98], [0.5, 0.5, 0.5],
                         [0.1, 0.1, 0.1], [0.9, 0.05, 0.3], [1 / 32, 2 / 32, 3 / 32], [0.99, 0.99, 0.01]])
    got = o.voxel_query(1, xs)
    ix = vox(pv)
    q = fixed(wv).astype(np.float64)
    S = scale(c, n_emit)
    for i, x in enumerate(xs):
        v = vox(x)
        lo = np.clip(v - 3, 1, V - 1)
        hi = np.clip(v + 3, 1, V - 1)
        m = np.all((ix >= lo) & (ix < hi), axis=1)
        vol = np.prod(hi - lo) / V ** 3
        want = q[m].sum(0) / FIX * S / vol / (4 * math.pi)
        np.testing.assert_allclose(got[i], want, rtol=1e-12, atol=0)


def cell_centres_on_wall(axis, val, power_of=None):
    """One record at the centre of each cell of a wall's boundary layer."""
    u = (np.arange(V) + 0.5) / V
    B, C = np.meshgrid(u, u, indexing="ij")
    pos = np.zeros((V * V, 3), np.float32)
    b, cc = (1, 2) if axis == 0 else ((0, 2) if axis == 1 else (0, 1))
    pos[:, axis] = val
    pos[:, b] = B.ravel()
    pos[:, cc] = C.ravel()
    p = np.ones(V * V) if power_of is None else power_of(B.ravel(), C.ravel())
    power = np.repeat(p[:, None], 3, axis=1).astype(np.float32)
    return pos, power, np.tile(np.float32([0, 0, 1]), (V * V, 1))


@pytest.mark.parametrize("interp", [False, True])
def test_uniform_wall_same_everywhere(interp, ncpu):
    """Equal power in every cell: every window, clamped or not, gives
    E = 1024 p Phi / N, so L = albedo / pi * E."""
    c = config()
    o = O.Oracle(c, ncpu)
    p = 0.25
    pos, power, dirs = cell_centres_on_wall(0, 1.0, lambda b, cc: p + 0 * b)
    n_emit = 1000
    o.set_records(0, pos, power, dirs, n_emit)
    rng = np.random.default_rng(10)
    xs = rng.random((100, 3)).astype(np.float32)
    xs[:, 0] = 1.0
    xs[:4, 1:] = np.float32([[0, 0], [1, 1], [0, 1], [0.999, 0.001]])
    ns = np.tile(np.float32([-1, 0, 0]), (100, 1))
    got = o.voxel_query(0, xs, ns, np.ones((100, 3)), interp=interp)
    want = V * V * p * scale(c, n_emit) / math.pi
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=1e-13)


@pytest.mark.parametrize("interp", [False, True])
def test_linear_wall_flux(interp, ncpu):
    """Cell power alpha * (i + 1/2) along y on the x=1 wall: an unclamped
    window's mean is alpha * v, so the estimate is 1024 alpha Phi / N times
    floor(32 y) without the blend and 32 y - 1/2 with it."""
    c = config()
    o = O.Oracle(c, ncpu)
    alpha = 1.0 / 64
    pos, power, dirs = cell_centres_on_wall(0, 1.0, lambda b, cc: alpha * (b * V))
    n_emit = 5000
    o.set_records(0, pos, power, dirs, n_emit)
    ys = np.linspace(3.6, 28.4, 57) / V
    xs = np.zeros((ys.size, 3), np.float32)
    xs[:, 0], xs[:, 1], xs[:, 2] = 1.0, ys, 0.5
    ns = np.tile(np.float32([-1, 0, 0]), (ys.size, 1))
    got = o.voxel_query(0, xs, ns, np.ones_like(xs), interp=interp)[:, 0]
    y32 = xs[:, 1].astype(np.float64) * V
    lin = (y32 - 0.5) if interp else np.floor(y32)
    want = V * V * alpha * lin * scale(c, n_emit)[0] / math.pi
    np.testing.assert_allclose(got, want, rtol=1e-6)


def test_uniform_volume_same_everywhere(ncpu):
    c = config()
    o = O.Oracle(c, ncpu)
    u = (np.arange(V) + 0.5) / V
    pos = np.stack(np.meshgrid(u, u, u, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    q = 0.125
    power = np.full_like(pos, q)
    n_emit = 100000
    o.set_records(1, pos, power, np.tile(np.float32([0, 0, 1]), (pos.shape[0], 1)), n_emit)
    rng = np.random.default_rng(11)
    xs = rng.random((200, 3)).astype(np.float32)
    xs[:3] = np.float32([[0, 0, 0], [1, 1, 1], [0.01, 0.99, 0.5]])
    got = o.voxel_query(1, xs)
    want = V ** 3 * q * scale(c, n_emit) / (4 * math.pi)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=1e-13)


def uniform_voxel_scene():
    """No spheres, every wall albedo 0.6, equal power in every cell of every
    wall layer and of the volume: the voxel estimates are the constants
    Ss = 0.6 / pi * 1024 p Phi / N and Sv = 32768 q Phi / N / (4 pi)."""
    c = config(n_spheres=0, albedo=0.6)
    # one record at the centre of every cell of the walls' boundary layers
    # (x = 0, x = 1, y = 0, y = 1 and z = 1; a cell on an edge belongs to two
    # layers and holds one record, so every layer is uniform)
    i = np.arange(V)
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    on = (I == 0) | (I == V - 1) | (J == 0) | (J == V - 1) | (K == V - 1)
    ps = ((np.stack([I[on], J[on], K[on]], -1) + 0.5) / V).astype(np.float32)
    p, q, n_emit = 0.5, 0.25, 20000
    ws = np.full_like(ps, p)
    ds = np.tile(np.float32([0, 0, 1]), (ps.shape[0], 1))
    u = (np.arange(V) + 0.5) / V
    pv = np.stack(np.meshgrid(u, u, u, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    wv = np.full_like(pv, q)
    dv = np.tile(np.float32([0, 0, 1]), (pv.shape[0], 1))
    S = scale(c, n_emit)
    Ss = (np.float32(0.6) / math.pi * V * V * p * S).astype(np.float32)
    Sv = (V ** 3 * q * S / (4 * math.pi)).astype(np.float32)
    return c, (ps, ws, ds), (pv, wv, dv), n_emit, Sv, Ss


def test_uniform_scene_render_equals_constant_gather(ncpu):
    c, srec, vrec, n_emit, Sv, Ss = uniform_voxel_scene()
    o = O.Oracle(c, ncpu)
    o.set_records(0, *srec, n_emit)
    o.set_records(1, *vrec, n_emit)
    img = o.render_voxel()
    np.testing.assert_allclose(o.render_voxel(interp=True), img, rtol=1e-6, atol=0)
    assert o.last_queries[0] == c.width * c.height and o.last_queries[1] > 0
    for py in range(0, c.height, 3):
        for px in range(0, c.width, 5):
            want, _ = o.render_const(px, py, 0, Sv, Ss)
            np.testing.assert_allclose(img[py, px], want, rtol=2e-6, atol=0)
