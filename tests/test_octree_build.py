"""Octree build and point location (no GPU needed for the build: the library's
host restatement runs on the CPU).

* the oracle is pinned by the reference's own self-demo known answer
  (octree_fem_interpolator_optimized.py:413-433: field = node coordinates,
  query (0.25,0.25,0.25) returns itself), linear precision on a jittered
  mesh, brute-force containing-element search, element-partition invariants
  and NumPy's own mean for the centroids;
* the library's C++ build (jt_octree_build) must reproduce the oracle's
  arrays exactly."""
import ctypes as C

import numpy as np
import pytest

from jaxtrace_amd import _lib as L
from jaxtrace_amd import workloads as W
from oracle import octree as OO

F32 = np.float32


def test_reference_self_demo_known_answer():
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                    [1, 1, 1], [1, 0, 1], [1, 1, 0], [0, 1, 1]], F32)
    conn = np.array([[0, 1, 2, 3], [4, 5, 6, 7]], np.int32)
    tree = OO.build(pts, conn, max_elements_per_leaf=1, max_depth=5)
    val, e = OO.interpolate(tree, pts, conn, pts, np.array([0.25, 0.25, 0.25], F32), 5)
    assert e == 0
    np.testing.assert_allclose(val, [0.25, 0.25, 0.25], atol=1e-7)


def test_centroid_is_numpy_mean():
    pts, conn = W.kuhn_mesh(6, 0.2, 1)
    tree = OO.build(pts, conn, 8, 12)
    want = np.stack([pts[c].mean(axis=0) for c in conn[:500]])
    np.testing.assert_array_equal(tree["element_centroids"][:500], want)


def test_partition_and_linear_precision():
    pts, conn = W.kuhn_mesh(6, 0.2, 2)
    tree = OO.build(pts, conn, 8, 12)
    leaves = tree["nodes_is_leaf"]
    elems = tree["nodes_elements"][leaves]
    ids = np.sort(elems[elems >= 0])
    np.testing.assert_array_equal(ids, np.arange(len(conn)))      # each element in exactly one leaf
    assert (tree["nodes_elem_counts"][~leaves] == 0).all()
    # pre-order: every child index is larger than its parent
    ch = tree["nodes_children"]
    par = np.repeat(np.arange(len(ch)), 8).reshape(-1, 8)
    assert (ch[ch >= 0] > par[ch >= 0]).all()
    rng = np.random.default_rng(3)
    Q = rng.uniform(0.02, 0.98, (300, 3)).astype(F32)
    hits = 0
    for q in Q:
        v, e = OO.interpolate(tree, pts, conn, pts, q, 12)
        if e >= 0:
            hits += 1
            np.testing.assert_allclose(v, q, atol=2e-6)
            inside = [OO.barycentric(q, pts[c])[0] for c in conn]
            assert inside[e]
    assert hits > 100          # centroid binning misses some points by design


def _lib_build(pts, conn, leaf, depth):
    lib = L.load()
    h = C.c_void_p()
    L.check(lib.jt_octree_build(pts.ctypes.data, len(pts), conn.ctypes.data, len(conn), leaf, depth,
                                C.byref(h)), "build")
    nn, me, dep = C.c_int64(), C.c_int32(), C.c_int32()
    lib.jt_octree_info(h, C.byref(nn), C.byref(me), C.byref(dep))
    n, m, M = nn.value, me.value, len(conn)
    out = {"nodes_min": np.empty((n, 3), F32), "nodes_max": np.empty((n, 3), F32),
           "nodes_elements": np.empty((n, m), np.int32), "nodes_elem_counts": np.empty(n, np.int32),
           "nodes_children": np.empty((n, 8), np.int32), "nodes_is_leaf": np.empty(n, np.uint8),
           "element_bounds": np.empty((M, 2, 3), F32), "element_centroids": np.empty((M, 3), F32)}
    lib.jt_octree_export(h, *[out[k].ctypes.data for k in (
        "nodes_min", "nodes_max", "nodes_elements", "nodes_elem_counts", "nodes_children", "nodes_is_leaf",
        "element_bounds", "element_centroids")])
    lib.jt_octree_free(h)
    return out


@pytest.mark.parametrize("n,leaf,depth,jit", [(6, 8, 12, 0.2), (8, 32, 12, 0.2), (7, 4, 3, 0.0), (10, 16, 12, 0.3)])
def test_library_build_matches_oracle(n, leaf, depth, jit):
    pts, conn = W.kuhn_mesh(n, jit, 5)
    want = OO.build(pts, conn, leaf, depth)
    got = _lib_build(pts, conn, leaf, depth)
    for k, v in want.items():
        np.testing.assert_array_equal(got[k], v.astype(got[k].dtype), err_msg=k)


def test_library_build_rejects_bad_connectivity():
    pts = np.zeros((4, 3), F32)
    conn = np.array([[0, 1, 2, 9]], np.int32)
    lib = L.load()
    h = C.c_void_p()
    rc = lib.jt_octree_build(pts.ctypes.data, 4, conn.ctypes.data, 1, 32, 12, C.byref(h))
    assert rc == L.JT_EINVAL and b"out of range" in lib.jt_last_error()
