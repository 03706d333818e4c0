"""The oracle's canonical kd-tree and kNN, pinned by structure and brute force.

The kd-tree is checked against its definition (shape from n and the leaf
size, split axis = largest extent, median under (coordinate, index), leaf
order (x, index)) by independent numpy code; kNN by tree descent is checked
against exhaustive search, including exact ties, empty maps and fewer than k
photons inside the radius cap.
"""
import numpy as np
import pytest

import oracle as O


def leaves_under_brute(c, nl):
    cnt, stack = 0, [c]
    while stack:
        h = stack.pop()
        if h > 2 * nl - 2:
            continue
        if h >= nl - 1:
            cnt += 1
        else:
            stack += [2 * h + 1, 2 * h + 2]
    return cnt


@pytest.mark.parametrize("nl", [1, 2, 3, 5, 8, 13, 64, 100, 1000])
def test_leaves_under_closed_form(nl):
    for c in range(0, 2 * nl - 1):
        assert O.kd_leaves_under(c, nl) == leaves_under_brute(c, nl)


def fkey(v):
    u = np.float32(v).view(np.uint32).astype(np.uint64)
    return int(np.where(u & 0x80000000, ~u & 0xFFFFFFFF, u | 0x80000000))


def check_tree(pos, perm, split, axis, L):
    n = len(pos)
    nl = (n + L - 1) // L if n else 0
    assert sorted(perm.tolist()) == list(range(n))
    if nl <= 1:
        order = sorted(range(n), key=lambda i: (fkey(pos[i, 0]), i))
        assert perm.tolist() == order
        return

    def walk(h, lo, hi):
        if h >= nl - 1:
            assert hi - lo <= L
            seg = perm[lo:hi].tolist()
            assert seg == sorted(seg, key=lambda i: (fkey(pos[i, 0]), i)), "leaf order"
            return
        idx = perm[lo:hi]
        pts = pos[idx]
        ext = (pts.max(0) - pts.min(0)).astype(np.float32)
        a = int(axis[h])
        assert a == int(np.argmax(ext == ext.max())), "split axis = largest extent, lowest on ties"
        lc = L * O.kd_leaves_under(2 * h + 1, nl)
        keyed = sorted(idx.tolist(), key=lambda i: (fkey(pos[i, a]), i))
        med = keyed[lc]
        assert np.float32(split[h]).view(np.uint32) == np.float32(pos[med, a]).view(np.uint32)
        assert sorted(perm[lo:lo + lc].tolist()) == sorted(keyed[:lc])
        walk(2 * h + 1, lo, lo + lc)
        walk(2 * h + 2, lo + lc, hi)

    walk(0, 0, n)


def oracle_with_points(pos, L=8, which=0):
    c = O.preset(1)
    c.leaf_size = L
    o = O.Oracle(c, 4)
    n = len(pos)
    z = np.zeros((n, 3), np.float32)
    d = np.tile(np.array([0, -1, 0], np.float32), (n, 1))
    o.set_records(0, pos, np.ones((n, 3), np.float32), d, 1000)
    o.set_records(1, pos if which == 1 else z[:0], np.ones((n if which == 1 else 0, 3), np.float32),
                  d if which == 1 else z[:0], 1000)
    o.build()
    return o


@pytest.mark.parametrize("n,L", [(0, 8), (1, 8), (7, 8), (8, 8), (9, 8), (100, 8), (1000, 16), (3000, 32),
                                 (2049, 64)])
def test_tree_structure(n, L):
    rng = np.random.default_rng(n)
    pos = rng.random((n, 3)).astype(np.float32)
    pos[:, 2] *= 0.3  # unequal extents
    o = oracle_with_points(pos, L)
    perm, split, axis = o.tree(0)
    check_tree(pos, perm, split, axis, L)


def test_tree_with_duplicate_coordinates():
    rng = np.random.default_rng(5)
    pos = np.round(rng.random((2000, 3)) * 8) / 8  # many exact ties in every coordinate
    pos = pos.astype(np.float32)
    o = oracle_with_points(pos, 8)
    check_tree(pos, *o.tree(0), 8)


def brute(pos, q, k, rcap):
    d2 = ((pos.astype(np.float32)[None, :, :] - q.astype(np.float32)[:, None, :]) ** 2).astype(np.float32)
    d2 = (d2[..., 0] + d2[..., 1]) + d2[..., 2]
    out_i = np.full((len(q), k), 0xFFFFFFFF, np.uint32)
    out_d = np.full((len(q), k), np.inf, np.float32)
    cnt = np.zeros(len(q), np.uint32)
    r2 = np.float32(rcap) * np.float32(rcap)
    for j in range(len(q)):
        ok = np.nonzero(d2[j] <= r2)[0]
        order = sorted(ok.tolist(), key=lambda i: (d2[j, i], i))[:k]
        cnt[j] = len(order)
        out_i[j, :len(order)] = order
        out_d[j, :len(order)] = d2[j, order]
    return out_i, out_d, cnt


@pytest.mark.parametrize("k,rcap", [(1, 0.2), (8, 0.1), (32, 0.05), (64, 0.5), (64, 0.01)])
def test_knn_tree_equals_numpy_brute_force(k, rcap):
    rng = np.random.default_rng(k)
    pos = rng.random((1500, 3)).astype(np.float32)
    pos[:300] = np.round(pos[:300] * 16) / 16  # exact duplicates -> ties broken by index
    o = oracle_with_points(pos, 8)
    q = np.concatenate([rng.random((150, 3)), pos[:50], pos[300:320] + 1e-3]).astype(np.float32)
    ti, td, tc = o.knn(0, q, k, rcap)
    bi, bd, bc = brute(pos, q, k, rcap)
    assert np.array_equal(tc, bc)
    assert np.array_equal(ti, bi)
    assert np.array_equal(td.view(np.uint32), bd.view(np.uint32))
    oi, od, oc = o.knn(0, q, k, rcap, brute=True)  # the oracle's own brute force agrees too
    assert np.array_equal(oi, bi) and np.array_equal(oc, bc)


def test_knn_empty_map_and_no_neighbours():
    o = oracle_with_points(np.zeros((0, 3), np.float32), 8)
    i, d, c = o.knn(0, np.array([[0.5, 0.5, 0.5]], np.float32), 8, 0.1)
    assert c[0] == 0 and (i == 0xFFFFFFFF).all() and np.isinf(d).all()
    o = oracle_with_points(np.array([[0.9, 0.9, 0.9]], np.float32), 8)
    i, d, c = o.knn(0, np.array([[0.1, 0.1, 0.1]], np.float32), 8, 0.1)
    assert c[0] == 0
