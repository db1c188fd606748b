"""Pins the oracle's temporal blend, trilinear and nearest-node sampling
(oracle/fields.py) against closed forms, invariants and brute force."""
import numpy as np
import pytest

from oracle import fields as F

F32 = np.float32


def _grid(n, lo=0.0, hi=1.0, endpoint=True):
    g = np.linspace(lo, hi, n, endpoint=endpoint).astype(F32)
    return g


def multilinear(P, c):
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    return (c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z
            + c[6] * z * x + c[7] * x * y * z)


def test_trilinear_reproduces_multilinear_fields():
    rng = np.random.default_rng(0)
    gx, gy, gz = _grid(9, 0, 2), _grid(7, -1, 1), _grid(5, 0.5, 1.5)
    meta = F.GridMeta(gx, gy, gz)
    X, Y, Z = np.meshgrid(gx.astype(np.float64), gy.astype(np.float64), gz.astype(np.float64),
                          indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    coefs = rng.standard_normal((3, 8))
    V = np.stack([multilinear(nodes, c) for c in coefs], 1).reshape(9, 7, 5, 3).astype(F32)
    P = np.stack([rng.uniform(0, 2, 500), rng.uniform(-1, 1, 500), rng.uniform(0.5, 1.5, 500)],
                 1).astype(F32)
    got = F.trilinear(V, meta, P, "clamp")
    want = np.stack([multilinear(P.astype(np.float64), c) for c in coefs], 1)
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-5)


def test_trilinear_periodic_wrap_and_seam():
    rng = np.random.default_rng(1)
    n = 16
    g = (np.arange(n) / n).astype(F32)
    meta = F.GridMeta(g, g, g)
    V = rng.standard_normal((n, n, n, 3)).astype(F32)
    # dyadic points so x+1 and x-2 are exact in f32
    P = (rng.integers(0, 4 * n, (200, 3)) / (4 * n)).astype(F32)
    a = F.trilinear(V, meta, P, "periodic")
    b = F.trilinear(V, meta, (P + F32(1)).astype(F32), "periodic")
    c = F.trilinear(V, meta, (P - F32(2)).astype(F32), "periodic")
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, c)
    # seam: halfway between node n-1 and node 0 along x
    q = np.array([[(n - 0.5) / n, 0, 0]], F32)
    got = F.trilinear(V, meta, q, "periodic")[0]
    np.testing.assert_allclose(got, 0.5 * V[n - 1, 0, 0] + 0.5 * V[0, 0, 0], rtol=1e-6, atol=1e-6)
    # grid nodes return node values exactly
    idx = rng.integers(0, n, (50, 3))
    np.testing.assert_array_equal(F.trilinear(V, meta, (idx / n).astype(F32), "periodic"),
                                  V[idx[:, 0], idx[:, 1], idx[:, 2]])


def test_trilinear_out_of_bounds_modes():
    g = _grid(5, 0, 1)
    meta = F.GridMeta(g, g, g)
    V = np.arange(5 * 5 * 5 * 3, dtype=F32).reshape(5, 5, 5, 3)
    P = np.array([[-0.5, 0.5, 0.5], [1.0, 1.0, 1.0], [0.5, 1.5, 0.25], [0.0, 0.0, 0.0]], F32)
    cl = F.trilinear(V, meta, P, "clamp")
    np.testing.assert_array_equal(cl[0], F.trilinear(V, meta, np.array([[0.0, 0.5, 0.5]], F32))[0])
    np.testing.assert_array_equal(cl[1], V[4, 4, 4])           # inclusive upper bound
    ze = F.trilinear(V, meta, P, "zero")
    assert (ze[0] == 0).all() and (ze[2] == 0).all()
    np.testing.assert_array_equal(ze[1], V[4, 4, 4])
    na = F.trilinear(V, meta, P, "nan")
    assert np.isnan(na[0]).all() and np.isnan(na[2]).all() and not np.isnan(na[3]).any()


def test_blend_linear_in_time_and_snapshot_identity():
    rng = np.random.default_rng(2)
    times = np.array([0.0, 0.3, 0.5, 1.25], F32)
    a, b = rng.standard_normal((2, 40, 3))
    data = np.stack([a + b * float(t) for t in times]).astype(F32)
    for t in rng.uniform(0, 1.25, 20).astype(F32):
        got = F.sample_all_nodes_at_time(data, times, t)
        np.testing.assert_allclose(got, a + b * float(t), rtol=0, atol=2e-6 * (1 + np.abs(b).max()))
    for k, t in enumerate(times):
        np.testing.assert_array_equal(F.sample_all_nodes_at_time(data, times, t), data[k])


def test_blend_extrapolation_modes():
    times = np.array([1.0, 2.0, 4.0], F32)
    a = np.array([[1.0, -2.0, 0.5]], F32)
    b = np.array([[0.25, 1.0, -1.0]], F32)
    data = np.stack([a + b * float(t) for t in times]).astype(F32)
    np.testing.assert_array_equal(F.sample_all_nodes_at_time(data, times, 0.0, "linear", "constant"), data[0])
    np.testing.assert_array_equal(F.sample_all_nodes_at_time(data, times, 9.0, "linear", "constant"), data[-1])
    assert (F.sample_all_nodes_at_time(data, times, 9.0, "linear", "zero") == 0).all()
    assert np.isnan(F.sample_all_nodes_at_time(data, times, -1.0, "linear", "nan")).all()
    # linear extrapolation continues the first / last segment exactly for linear data
    np.testing.assert_allclose(F.sample_all_nodes_at_time(data, times, 0.0, "linear", "linear"), a, atol=1e-6)
    np.testing.assert_allclose(F.sample_all_nodes_at_time(data, times, 6.0, "linear", "linear"),
                               a + 6 * b, atol=1e-5)
    # nearest: first minimum of |times - t|
    np.testing.assert_array_equal(F.sample_all_nodes_at_time(data, times, 2.9, "nearest"), data[1])
    np.testing.assert_array_equal(F.sample_all_nodes_at_time(data, times, 3.0, "nearest"), data[1])
    np.testing.assert_array_equal(F.sample_all_nodes_at_time(data, times, 3.1, "nearest"), data[2])


def test_corner_blend_equals_all_node_blend_bitwise():
    rng = np.random.default_rng(3)
    n = 12
    g = (np.arange(n) / n).astype(F32)
    meta = F.GridMeta(g, g, g)
    data = rng.standard_normal((5, n, n, n, 3)).astype(F32)
    times = np.sort(rng.uniform(0, 1, 5)).astype(F32)
    P = rng.uniform(-0.2, 1.2, (300, 3)).astype(F32)
    for mode in ("clamp", "periodic", "zero", "nan"):
        for t in (times[0], 0.37, times[-1] + 0.1):
            a = F.structured_series_sample(data, times, meta, P, t, mode, "linear", "linear")
            b = F.structured_series_sample_corners(data, times, meta, P, t, mode, "linear", "linear")
            np.testing.assert_array_equal(a, b)


def test_nearest_node_brute_force_with_ties():
    rng = np.random.default_rng(4)
    nodes = rng.uniform(0, 1, (30, 3)).astype(F32)
    nodes[17] = nodes[5]                      # duplicate: ties go to the lower index
    Q = np.concatenate([rng.uniform(0, 1, (50, 3)), nodes[[5, 17]]]).astype(F32)
    got = F.nearest_node_index(nodes, Q)
    for i, q in enumerate(Q):
        d = [float(np.sum((q.astype(np.float64) - p) ** 2)) for p in nodes]
        assert got[i] == int(np.argmin(d))
    assert got[-1] == 5 and got[-2] == 5
