"""Pins the density oracle (oracle/density.py): hash-grid candidates by
brute-force enumeration, the x32 key truncation, the SPH cubic-spline
closed form, brute-force SPH sums and the KDE closed form."""
import math

import numpy as np
import pytest

from oracle import density as D

F32 = np.float32


def brute_candidates(P, q, h, cap, key_mode):
    """Independent enumeration: for each stencil offset in lexicographic
    order, the particles whose (key-mode) cell equals the probed cell, by
    increasing particle index, first ``cap``; empty slots -> dummy."""
    inv = F32(F32(1) / F32(h))
    origin = P.min(0)
    C = np.floor((P - origin) * inv).astype(np.int64)
    C -= C.min(0)
    cq = np.floor((q - origin) * inv).astype(np.int64) - np.floor((P - origin) * inv).astype(np.int64).min(0)
    order_first = np.argsort(
        (C[:, 1] * 4096 + C[:, 2]) if key_mode == "x32" else ((C[:, 0] * 4096 + C[:, 1]) * 4096 + C[:, 2]),
        kind="stable")[0]
    pts, valid = [], []
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                c = cq + [di, dj, dk]
                if key_mode == "x32":
                    m = (C[:, 1] == c[1]) & (C[:, 2] == c[2]) & (c[1] >= 0) & (c[2] >= 0)
                else:
                    m = (C == c).all(1) & (c >= 0).all()
                members = list(np.nonzero(m)[0][:cap])
                pts += members + [order_first] * (cap - len(members))
                valid += [True] * len(members) + [False] * (cap - len(members))
    return np.array(pts), np.array(valid)


@pytest.mark.parametrize("key_mode", ["x32", "x64"])
def test_candidates_match_brute_force(key_mode):
    rng = np.random.default_rng(0)
    P = rng.random((600, 3)).astype(F32)
    h = 0.125
    g = D.HashGrid(P, h, per_cell_cap=8, key_mode=key_mode)
    for q in rng.uniform(-0.1, 1.1, (40, 3)).astype(F32):
        cp, cv = g.candidates(q)
        bp, bv = brute_candidates(P, q, h, 8, key_mode)
        np.testing.assert_array_equal(cv, bv)
        np.testing.assert_array_equal(cp, bp)


def test_x32_key_drops_x_and_triplicates_columns():
    rng = np.random.default_rng(1)
    P = rng.random((500, 3)).astype(F32)
    g = D.HashGrid(P, 0.1, per_cell_cap=16, key_mode="x32")
    cp, cv = g.candidates(np.array([0.5, 0.5, 0.5], F32))
    blocks = cp.reshape(3, 9, 16), cv.reshape(3, 9, 16)
    for k in (1, 2):
        np.testing.assert_array_equal(blocks[0][0], blocks[0][k])
    # particles far away in x are candidates in x32 mode
    assert np.abs(P[cp[cv], 0] - 0.5).max() > 0.3


def test_query_sort_stable_and_dummy_slots():
    P = np.array([[0.1, 0.1, 0.1], [0.1, 0.1, 0.1], [0.15, 0.1, 0.1], [0.9, 0.9, 0.9]], F32)
    g = D.HashGrid(P, 0.2, per_cell_cap=4, key_mode="x64")
    idx, dist = g.query_many(np.array([[0.1, 0.1, 0.1]], F32), 0.2, 6)
    # equal distances keep candidate order (lower particle index first)
    np.testing.assert_array_equal(idx[0, :3], [0, 1, 2])
    np.testing.assert_allclose(dist[0, :3], [0, 0, 0.05], atol=1e-7)
    assert np.isinf(dist[0, 3:]).all()
    # invalid slots carry pt_indices_sorted[0]
    assert (idx[0, 3:] == g.order[0]).all()


def test_cubic_spline_normalisation_and_shape():
    h = F32(2 ** -7)
    r = np.linspace(0, 2.5 * float(h), 200001)
    w = D.cubic_spline(r.astype(F32), h).astype(np.float64)
    integral = np.trapezoid(w * 4 * np.pi * r * r, r)
    assert abs(integral - 1.0) < 1e-4
    np.testing.assert_allclose(D.cubic_spline(F32(0), h), 1 / (np.pi * float(h) ** 3), rtol=1e-6)
    assert D.cubic_spline(F32(2.0) * h, h) == 0 and D.cubic_spline(F32(np.inf), h) == 0
    # continuity at q = 1
    a = D.cubic_spline(h * F32(0.9999999), h)
    b = D.cubic_spline(h * F32(1.0000001), h)
    assert abs(float(a) - float(b)) / float(a) < 1e-5


def test_sph_matches_brute_force_sum():
    rng = np.random.default_rng(2)
    P = rng.random((3000, 3)).astype(F32)
    Q = rng.random((50, 3)).astype(F32)
    h = 0.05
    rho, idx, dist = D.sph_density(P, Q, h, per_cell_cap=64, max_neighbors=512, key_mode="x64",
                                   normalize=False)
    inv = F32(1) / F32(h)
    origin = P.min(0)
    C = np.floor((P - origin) * inv).astype(np.int64)
    for m, q in enumerate(Q):
        cq = np.floor((q - origin) * inv).astype(np.int64)
        near = (np.abs(C - cq) <= 1).all(1)
        d = np.sqrt(((P[near].astype(np.float64) - q) ** 2).sum(1))
        d = d[d <= 2 * h]
        hh = float(h)
        qq = d / hh
        sig = 1 / (math.pi * hh ** 3)
        w = np.where(qq <= 1, sig * (1 - 1.5 * qq ** 2 + 0.75 * qq ** 3), sig * 0.25 * (2 - qq) ** 3)
        np.testing.assert_allclose(rho[m], w.sum(), rtol=2e-5)


def test_kde_closed_form_and_scott():
    P = np.array([[0.5, 0.5, 0.5]], F32)
    Q = np.array([[0.5, 0.5, 0.5], [0.6, 0.5, 0.5]], F32)
    h = 0.1
    got = D.kde_eval(Q, P, h)
    np.testing.assert_allclose(got, (2 * np.pi) ** -1.5 * np.exp([0.0, -0.5]), rtol=1e-6)
    x = np.random.default_rng(3).normal(0, 0.2, (200_000, 3)).astype(F32)
    assert abs(D.scott_bandwidth(x) / (0.2 * 200_000 ** (-1 / 7)) - 1) < 0.01
    X, Y, Z = D.make_grid_3d(x[:10], 4, ((0, 1),) * 3)
    assert X.shape == (4, 4, 4) and X[0, 1, 0] > X[0, 0, 0] and Y[1, 0, 0] > Y[0, 0, 0]
