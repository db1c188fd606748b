"""Oracle: octree build, point location and P1 interpolation on tetrahedra.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows jaxtrace/fields/octree_fem_interpolator_optimized.py
  * :41-215  ``build_octree_mesh_optimized`` (f32 element AABBs and centroids,
             centroid-only octant binning with >= tests, DFS pre-order node
             numbering, children in octant order 0..7, flattening)
  * :218-250 ``point_in_tetrahedron_fast`` (Cramer barycentrics, |det| < 1e-10
             -> 1, inside iff all b >= -1e-6 and all b <= 1 - 1e-6)
  * :253-380 ``interpolate_octree_optimized`` (fixed-depth descent that stays
             on leaves and missing children, first hit in stored leaf order,
             fallback to the nearest of the 4 nodes of candidate slot 0,
             where an empty slot (-1) indexes the LAST element)
and jaxtrace/fields/octree_fem_time_series_optimized.py:87-113 (blend every
node in time, then interpolate).

The 3x3 determinant is the explicit cofactor sum that ``jnp.linalg.det``
uses for 3x3 inputs, evaluated left to right; XLA may fuse its products into
FMAs, so barycentrics are pinned to rounding only (parity unpinned at the ulp
level; point-location ties within 1e-6 of a face are the only place it can
show).  The element centroid is NumPy's mean of a (4,3) float32 array, i.e.
((p0+p1)+p2)+p3 then /4, which tests confirm against NumPy itself.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def build(points, connectivity, max_elements_per_leaf=32, max_depth=12):
    """Returns a dict of the flattened arrays of OctreeMeshOptimized."""
    pts = np.asarray(points, dtype=F32)
    conn = np.asarray(connectivity, dtype=np.int32)
    M = conn.shape[0]
    ep = pts[conn]                                          # (M,4,3)
    ebounds = np.stack([ep.min(axis=1), ep.max(axis=1)], axis=1).astype(F32)
    cent = ((((ep[:, 0] + ep[:, 1]) + ep[:, 2]) + ep[:, 3]) / F32(4)).astype(F32)
    bmin, bmax = pts.min(axis=0), pts.max(axis=0)
    nodes = []   # (min, max, elems, children(list), is_leaf)

    def subdivide(mn, mx, elems, depth):
        idx = len(nodes)
        if len(elems) <= max_elements_per_leaf or depth >= max_depth:
            nodes.append([mn, mx, np.asarray(elems, np.int32), [-1] * 8, True])
            return idx
        c = ((mn + mx) / F32(2.0)).astype(F32)
        ob = [
            (mn, c),
            (np.array([c[0], mn[1], mn[2]], F32), np.array([mx[0], c[1], c[2]], F32)),
            (np.array([mn[0], c[1], mn[2]], F32), np.array([c[0], mx[1], c[2]], F32)),
            (np.array([c[0], c[1], mn[2]], F32), np.array([mx[0], mx[1], c[2]], F32)),
            (np.array([mn[0], mn[1], c[2]], F32), np.array([c[0], c[1], mx[2]], F32)),
            (np.array([c[0], mn[1], c[2]], F32), np.array([mx[0], c[1], mx[2]], F32)),
            (np.array([mn[0], c[1], c[2]], F32), np.array([c[0], mx[1], mx[2]], F32)),
            (c, mx),
        ]
        e = np.asarray(elems)
        ce = cent[e]
        octant = (ce[:, 0] >= c[0]).astype(int) + 2 * (ce[:, 1] >= c[1]) + 4 * (ce[:, 2] >= c[2])
        node = [mn, mx, np.zeros(0, np.int32), [-1] * 8, False]
        nodes.append(node)
        for o in range(8):
            sub = e[octant == o]
            if len(sub):
                node[3][o] = subdivide(ob[o][0], ob[o][1], sub, depth + 1)
        return idx

    subdivide(bmin, bmax, np.arange(M), 0)
    n = len(nodes)
    max_elems = max(len(nd[2]) for nd in nodes)
    out = {
        "nodes_min": np.array([nd[0] for nd in nodes], F32).reshape(n, 3),
        "nodes_max": np.array([nd[1] for nd in nodes], F32).reshape(n, 3),
        "nodes_elements": np.full((n, max_elems), -1, np.int32),
        "nodes_elem_counts": np.array([len(nd[2]) for nd in nodes], np.int32),
        "nodes_children": np.array([nd[3] for nd in nodes], np.int32).reshape(n, 8),
        "nodes_is_leaf": np.array([nd[4] for nd in nodes], bool),
        "element_bounds": ebounds,
        "element_centroids": cent,
    }
    for i, nd in enumerate(nodes):
        out["nodes_elements"][i, :len(nd[2])] = nd[2]
    return out


def det3(c0, c1, c2):
    """jnp.linalg.det of the 3x3 matrix with columns c0, c1, c2 (cofactor sum,
    left to right)."""
    a00, a10, a20 = c0[..., 0], c0[..., 1], c0[..., 2]
    a01, a11, a21 = c1[..., 0], c1[..., 1], c1[..., 2]
    a02, a12, a22 = c2[..., 0], c2[..., 1], c2[..., 2]
    return (((((a00 * a11) * a22 + (a01 * a12) * a20) + (a02 * a10) * a21)
             - (a02 * a11) * a20) - (a00 * a12) * a21) - (a01 * a10) * a22


def barycentric(q, v):
    """:218-250 for one point q (3,) and tet vertices v (4,3)."""
    v0p, v01, v02, v03 = q - v[0], v[1] - v[0], v[2] - v[0], v[3] - v[0]
    d = det3(v01, v02, v03)
    d = F32(1.0) if abs(d) < 1e-10 else d
    b1 = F32(det3(v0p, v02, v03) / d)
    b2 = F32(det3(v01, v0p, v03) / d)
    b3 = F32(det3(v01, v02, v0p) / d)
    b0 = F32(F32(F32(F32(1.0) - b1) - b2) - b3)
    b = np.array([b0, b1, b2, b3], F32)
    inside = bool(np.all(b >= F32(-1e-6)) and np.all(b <= F32(1.0 + -1e-6)))
    return inside, b


def locate(tree, q, max_depth):
    """:278-302: fixed-depth descent."""
    node = 0
    for _ in range(max_depth):
        if tree["nodes_is_leaf"][node]:
            continue
        c = ((tree["nodes_min"][node] + tree["nodes_max"][node]) / F32(2.0)).astype(F32)
        o = int(q[0] >= c[0]) + 2 * int(q[1] >= c[1]) + 4 * int(q[2] >= c[2])
        ch = tree["nodes_children"][node, o]
        node = ch if ch >= 0 else node
    return node


def interpolate(tree, points, conn, F, q, max_depth):
    """:253-380 for one query; F is the (already time-blended) nodal field.
    Returns (value (3,), found element or -1)."""
    pts = np.asarray(points, F32)
    q = np.asarray(q, F32)
    node = locate(tree, q, max_depth)
    cands = tree["nodes_elements"][node]
    for e in cands:
        if e < 0:
            continue
        bmin, bmax = tree["element_bounds"][e]
        if not (np.all(q >= bmin) and np.all(q <= bmax)):
            continue
        ni = conn[e]
        inside, b = barycentric(q, pts[ni])
        if inside:
            v = F[ni]
            val = ((b[0] * v[0] + b[1] * v[1]) + b[2] * v[2]) + b[3] * v[3]
            return val.astype(F32), int(e)
    first = cands[0]
    ni = conn[first]                       # -1 -> last element (negative indexing)
    d = pts[ni] - q
    dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return np.array(F[ni[int(np.argmin(dd))]], F32), -1


def sample(tree, points, conn, data, times, Q, t, max_depth, interpolation="linear",
           extrapolation="constant"):
    """OctreeFEMTimeSeriesFieldOptimized.sample_at_positions (:87-113)."""
    from .fields import sample_all_nodes_at_time
    F = sample_all_nodes_at_time(data, times, t, interpolation, extrapolation)
    return np.stack([interpolate(tree, points, conn, F, q, max_depth)[0] for q in np.asarray(Q, F32)])
