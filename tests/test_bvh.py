"""MTBVH builder and traversal pins.

* The library's builder and the oracle's independent restatement of BVH::quickBuild +
  buildHitTable (src/accelerator/BVH.cpp:217-345) produce identical buffers.
* Structural invariants of the six threaded tables.
* The oracle's stackless closest hit equals a brute-force search over all triangles
  (numpy FP32 restatement of intersection.glsl:63-109) for rays of every dominant axis,
  and its t agrees with a textbook FP64 Moller-Trumbore.
"""
import numpy as np
import pytest

from zillumgl_amd import _lib, scene
from oracle import oracle as orc


def random_mesh(n, seed, degenerate=None):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 3))
    if degenerate == "same_centroid":
        c[:] = 0.25
    if degenerate == "plane":
        c[:, 2] = 0.0
    e = rng.normal(scale=0.05, size=(n, 3, 3))
    v = (c[:, None, :] + e).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return v, idx


CASES = [(1, 0, None), (2, 1, None), (3, 2, None), (17, 3, None), (1000, 4, None), (20000, 5, None),
         (64, 6, "same_centroid"), (500, 7, "plane")]


@pytest.mark.parametrize("n,seed,deg", CASES)
def test_builder_matches_oracle(n, seed, deg):
    v, idx = random_mesh(n, seed, deg)
    b1, t1 = _lib.build_mtbvh(v, idx)
    b2, t2 = orc.build_mtbvh(v, idx)
    assert np.array_equal(b1.view(np.uint32), b2.view(np.uint32))
    assert np.array_equal(t1, t2)


def subtree_sizes(table):
    """Subtree size of every row of a preorder table from its miss links."""
    return table[:, 2] - np.arange(len(table))


@pytest.mark.parametrize("n,seed,deg", CASES)
def test_hit_table_invariants(n, seed, deg):
    v, idx = random_mesh(n, seed, deg)
    bounds, tables = orc.build_mtbvh(v, idx)
    size = 2 * n - 1
    assert tables.shape == (6, size, 3)
    cent = (bounds[:, :3] + bounds[:, 3:]) * 0.5
    for t in range(6):
        T = tables[t]
        # every node exactly once; every primitive in exactly one leaf
        assert np.array_equal(np.sort(T[:, 0]), np.arange(size))
        leaves = T[T[:, 1] >= 0]
        assert len(leaves) == n
        assert np.array_equal(np.sort(leaves[:, 1]), np.arange(n))
        sz = subtree_sizes(T)
        assert np.all(sz[T[:, 1] >= 0] == 1)
        assert np.all(T[:, 2] <= size) and np.all(T[:, 2] > np.arange(size))
        # internal rows: children are the next row and the row after the first child's subtree
        axis, greater = t // 2, (t % 2 == 0)
        for k in np.nonzero(T[:, 1] < 0)[0][:2000]:
            a = k + 1
            b = a + sz[a]
            assert sz[k] == 1 + sz[a] + sz[b]
            na, nb, nk = T[a, 0], T[b, 0], T[k, 0]
            lo, hi = bounds[nk, :3], bounds[nk, 3:]
            for ch in (na, nb):
                assert np.all(bounds[ch, :3] >= lo) and np.all(bounds[ch, 3:] <= hi)
            if greater:
                assert cent[na, axis] >= cent[nb, axis]
            else:
                assert cent[na, axis] <= cent[nb, axis]
        # leaf bounds are the triangle's AABB
        for k in np.nonzero(T[:, 1] >= 0)[0][:500]:
            tri = v[idx[T[k, 1]]]
            assert np.array_equal(bounds[T[k, 0], :3], tri.min(0)) and np.array_equal(bounds[T[k, 0], 3:], tri.max(0))


def brute_force(v, idx, o, d):
    """numpy FP32 restatement of intersectTriangle over all triangles; min t, first index on ties."""
    f = np.float32
    a, b, c = v[idx[:, 0]], v[idx[:, 1]], v[idx[:, 2]]
    ab, ac = (b - a).astype(f), (c - a).astype(f)

    def cross(x, y):
        return np.stack([x[..., 1] * y[..., 2] - x[..., 2] * y[..., 1], x[..., 2] * y[..., 0] - x[..., 0] * y[..., 2],
                         x[..., 0] * y[..., 1] - x[..., 1] * y[..., 0]], -1).astype(f)

    def dot(x, y):
        return ((x[..., 0] * y[..., 0] + x[..., 1] * y[..., 1]) + x[..., 2] * y[..., 2]).astype(f)

    n = o.shape[0]
    best_t = np.full(n, f(1e8), f)
    best_p = np.full(n, -1, np.int64)
    for i in range(idx.shape[0]):
        p = cross(d, np.broadcast_to(ac[i], d.shape))
        det = dot(np.broadcast_to(ab[i], p.shape), p)
        ao = (o - a[i]).astype(f)
        neg = det < 0
        ao = np.where(neg[:, None], -ao, ao)
        det = np.where(neg, -det, det)
        u = dot(ao, p)
        q = cross(ao, np.broadcast_to(ab[i], ao.shape))
        vv = dot(d, q)
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (dot(np.broadcast_to(ac[i], q.shape), q) / det).astype(f)
        ok = (np.abs(det) >= f(1e-6)) & (u >= 0) & (u <= det) & (vv >= 0) & ((u + vv).astype(f) <= det) & (t > 0)
        better = ok & (t < best_t)
        best_t = np.where(better, t, best_t)
        best_p = np.where(better, i, best_p)
    return best_p, best_t


def test_closest_hit_equals_brute_force_all_tables():
    d = scene.config1(8, 8)
    ps = scene.pack(d)
    rng = np.random.default_rng(11)
    n = 6000
    o = rng.uniform(0.02, 0.98, (n, 3)).astype(np.float32)
    dirs = rng.normal(size=(n, 3))
    # force every dominant axis / sign so that all six tables are exercised
    for k in range(6):
        sl = slice(k * 1000, (k + 1) * 1000)
        dirs[sl, k // 2] = (5.0 if k % 2 == 0 else -5.0)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dirs = dirs.astype(np.float32)
    rays = np.ascontiguousarray(np.concatenate([o.T, dirs.T]).astype(np.float32))
    prim, t, _, _ = orc.trace_closest(ps, rays)
    bp, bt = brute_force(ps.vertices, ps.indices, o, dirs)
    _check_against_brute_force(prim, t, bp, bt, ps.vertices, ps.indices, o, dirs)


def _check_against_brute_force(prim, t, bp, bt, v=None, idx=None, o=None, d=None):
    """Exact agreement except where two triangles are coplanar (the boxes stand on the
    floor): there the reference's pruning test `boxDist > dist` on the rounded slab
    distance may skip a triangle whose t is a few ulp smaller, keeping the first one found."""
    diff = t.view(np.uint32) != bt.view(np.uint32)
    assert diff.mean() < 1e-3
    ulps = np.abs(t.view(np.int32).astype(np.int64) - bt.view(np.int32).astype(np.int64))
    assert np.all(ulps[diff] <= 4)
    assert np.all(bt[diff] <= t[diff])
    # exact ties in t (shared edges, the boxes' bottoms lying on the floor) may report
    # either triangle: the reported one must hit at exactly the same t
    other = np.nonzero((prim != bp) & ~diff)[0]
    for i in other:
        p1, t1 = brute_force(v, idx[prim[i]:prim[i] + 1], o[i:i + 1], d[i:i + 1])
        assert p1[0] == 0 and t1[0] == t[i]


def test_axis_aligned_and_grazing_rays():
    """boxHit's special branches (|d.i| > 1-1e-6 and |d.i| < 1e-6) against brute force."""
    ps = scene.pack(scene.config1(8, 8))
    rng = np.random.default_rng(12)
    o = rng.uniform(0.05, 0.95, (600, 3)).astype(np.float32)
    dirs = np.zeros((600, 3), np.float32)
    for k in range(6):
        dirs[k * 100:(k + 1) * 100, k // 2] = 1.0 if k % 2 == 0 else -1.0
    # half of them slightly off axis with one exactly-zero component
    dirs[::2, (np.arange(300) % 3)] += np.float32(1e-3)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.concatenate([o.T, dirs.T]).astype(np.float32))
    prim, t, _, _ = orc.trace_closest(ps, rays)
    bp, bt = brute_force(ps.vertices, ps.indices, o, dirs)
    _check_against_brute_force(prim, t, bp, bt, ps.vertices, ps.indices, o, dirs)


def test_intersection_t_matches_fp64_textbook():
    rng = np.random.default_rng(13)
    n = 2000
    a = rng.uniform(-1, 1, (n, 3))
    b = a + rng.normal(scale=0.5, size=(n, 3))
    c = a + rng.normal(scale=0.5, size=(n, 3))
    tgt = (a + b + c) / 3.0
    o = tgt + rng.normal(size=(n, 3)) * 2.0
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    v = np.stack([a, b, c], 1).reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    bp, bt = brute_force_single(v, idx, o.astype(np.float32), d.astype(np.float32))
    nrm = np.cross(b - a, c - a)
    t64 = np.einsum("ij,ij->i", nrm, a - o) / np.einsum("ij,ij->i", nrm, d)
    ok = bp >= 0
    assert ok.mean() > 0.99
    assert np.max(np.abs(bt[ok] - t64[ok]) / np.abs(t64[ok])) < 1e-4


def brute_force_single(v, idx, o, d):
    """One triangle per ray."""
    out_p = np.full(len(o), -1)
    out_t = np.full(len(o), np.float32(1e8))
    for i in range(len(o)):
        p, t = brute_force(v, idx[i:i + 1], o[i:i + 1], d[i:i + 1])
        if p[0] >= 0:
            out_p[i], out_t[i] = i, t[0]
    return out_p, out_t
