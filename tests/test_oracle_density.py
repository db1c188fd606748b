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


# ---------------------------------------------------------------------------
# Closed forms and invariants of the SPH kernels, the bandwidth rules and the
# normalisations (sph.py:227-262, 365-378; kernels.py:30-70; kde.py:205-239).
# Each test is chosen so that a dropped factor, a wrong power of h or pi, or
# a swapped dimension fails it.
# ---------------------------------------------------------------------------

# W(0) * h^D for each kernel and dimension: the textbook normalisations
W0 = {("cubic_spline", 3): 1 / math.pi, ("cubic_spline", 2): 10 / (7 * math.pi),
      ("wendland", 3): 21 / (2 * math.pi), ("wendland", 2): 7 / math.pi,
      ("gaussian", 3): math.pi ** -1.5, ("gaussian", 2): 1 / math.pi}
SUPPORT = {"cubic_spline": 2.0, "wendland": 1.0, "gaussian": 8.0}


@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland", "gaussian"])
@pytest.mark.parametrize("dim", [2, 3])
def test_sph_kernel_integrates_to_one(kernel, dim):
    """Radial integral of W over R^D is 1 (shell measure 4 pi r^2 in 3-D,
    2 pi r in 2-D), at two smoothing lengths."""
    for h in (2.0 ** -7, 0.3):
        r = np.linspace(0.0, SUPPORT[kernel] * h, 400001)
        w = D.SPH_KERNELS[kernel](r.astype(F32), F32(h), dim).astype(np.float64)
        shell = 4 * np.pi * r * r if dim == 3 else 2 * np.pi * r
        assert abs(np.trapezoid(w * shell, r) - 1.0) < 2e-5, (kernel, dim, h)


@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland", "gaussian"])
@pytest.mark.parametrize("dim", [2, 3])
def test_sph_kernel_value_at_zero_and_support(kernel, dim):
    h = F32(0.05)
    np.testing.assert_allclose(float(D.SPH_KERNELS[kernel](F32(0), h, dim)), W0[(kernel, dim)] / float(h) ** dim,
                               rtol=2e-6)
    if kernel == "wendland":
        # compact support at q = 1, and the C2 shape (1-q)^4 (1+4q) inside
        for q in (1.0, 1.5, 3.0):
            assert D.wendland(F32(q) * h, h, dim) == 0
        q = 0.5
        np.testing.assert_allclose(float(D.wendland(F32(q) * h, h, dim)),
                                   W0[(kernel, dim)] / float(h) ** dim * (1 - q) ** 4 * (1 + 4 * q), rtol=2e-6)
    elif kernel == "gaussian":
        # exp(-q^2) shape: W(h) = W(0)/e, W(2h) = W(0)/e^4
        w0 = float(D.gaussian_sph(F32(0), h, dim))
        np.testing.assert_allclose(float(D.gaussian_sph(h, h, dim)), w0 / math.e, rtol=2e-6)
        np.testing.assert_allclose(float(D.gaussian_sph(F32(2) * h, h, dim)), w0 * math.exp(-4), rtol=4e-6)
    else:
        assert D.cubic_spline(F32(2) * h, h, dim) == 0
        # the two pieces meet at q = 1 with value sigma/4
        np.testing.assert_allclose(float(D.cubic_spline(h, h, dim)), 0.25 * W0[(kernel, dim)] / float(h) ** dim,
                                   rtol=2e-6)


def test_silverman_closed_form():
    """Silverman = (4/(d+2))^(1/(d+4)) sigma n^(-1/(d+4)): equal to Scott's
    rule in 2-D (factor 1), and on an isotropic Gaussian sample close to the
    population value in 3-D."""
    rng = np.random.default_rng(30)
    x2 = rng.normal(0, 0.3, (50_000, 2)).astype(F32)
    assert D.silverman_bandwidth(x2) == D.scott_bandwidth(x2)
    x3 = rng.normal(0, 0.2, (200_000, 3)).astype(F32)
    want = (4 / 5) ** (1 / 7) * 0.2 * 200_000 ** (-1 / 7)
    assert abs(D.silverman_bandwidth(x3) / want - 1) < 0.01
    np.testing.assert_allclose(D.silverman_bandwidth(x3) / D.scott_bandwidth(x3), (4 / 5) ** (1 / 7), rtol=3e-7)


def test_kde_normalize_3d_invariants():
    """After kde.py:229-239 the grid integrates to the domain volume
    (sum D cell_vol = vol), the result does not depend on the scale of the
    input, and a constant field maps to vol / (cell_vol * points)."""
    x = np.linspace(-0.5, 1.5, 9, dtype=F32)
    y = np.linspace(0.0, 0.5, 7, dtype=F32)
    z = np.linspace(2.0, 3.0, 5, dtype=F32)
    X, Y, Z = np.meshgrid(x, y, z, indexing="xy")
    Dd = np.random.default_rng(31).random(X.shape).astype(F32) + F32(0.1)
    out = D.kde_normalize_3d(X, Y, Z, Dd, 12345)
    cell = (2.0 / 8) * (0.5 / 6) * (1.0 / 4)
    vol = 2.0 * 0.5 * 1.0
    np.testing.assert_allclose((out.astype(np.float64) * cell).sum(), vol, rtol=1e-5)
    np.testing.assert_allclose(D.kde_normalize_3d(X, Y, Z, Dd * F32(64), 12345), out, rtol=1e-6)
    np.testing.assert_allclose(D.kde_normalize_3d(X, Y, Z, np.ones_like(Dd), 7), vol / (cell * Dd.size),
                               rtol=1e-6)


def test_kde_normalize_2d_invariants():
    x = np.linspace(0.25, 1.0, 11, dtype=F32)
    y = np.linspace(-1.0, 1.0, 11, dtype=F32)
    X, Y = np.meshgrid(x, y, indexing="xy")
    Z = np.random.default_rng(32).random(X.shape).astype(F32) + F32(0.5)
    out = D.kde_normalize_2d(X, Y, Z, 999)
    cell = (0.75 / 10) * (2.0 / 10)
    np.testing.assert_allclose((out.astype(np.float64) * cell).sum(), 0.75 * 2.0, rtol=1e-5)
    np.testing.assert_allclose(D.kde_normalize_2d(X, Y, Z * F32(3), 999), out, rtol=1e-6)


def _lattice(n, dim):
    a = (np.arange(n) + 0.5) / n
    G = np.meshgrid(*([a] * dim), indexing="ij")
    return np.stack([g.ravel() for g in G], 1)


def test_sph_normalised_unit_lattice_density_is_one_3d():
    """A cubic lattice of spacing a = 1/n filling the unit box has number
    density N / vol = 1/a^3.  The Wendland kernel with h = 4a sums to 1/a^3
    within 4e-4 at interior points (lattice quadrature of a unit-integral
    kernel), and its support (r <= h) lies inside the 27-cell stencil of
    cell size h.  (The cubic spline's support 2h does not: the reference's
    stencil reaches only one cell, h, beyond the query's cell, so its sums
    drop the neighbours between h and 2h in some directions.)
    sph.py:365-378 divides by N / vol(bbox of the queries), so with queries
    spanning the unit box the normalised density is 1, and with queries
    spanning half of it 1/2."""
    n = 32
    P = _lattice(n, 3).astype(F32)
    h = 4.0 / n
    rng = np.random.default_rng(33)
    Q = np.concatenate([[[0, 0, 0], [1, 1, 1]], rng.uniform(5 / n, 1 - 5 / n, (24, 3))]).astype(F32)
    rho, _, _ = D.sph_density(P, Q, h, 64, 512, "x64", True, kernel="wendland")
    np.testing.assert_allclose(rho[2:], 1.0, atol=5e-4)
    raw, _, _ = D.sph_density(P, Q, h, 64, 512, "x64", False, kernel="wendland")
    np.testing.assert_allclose(raw[2:] / n ** 3, 1.0, atol=5e-4)
    Qh = Q.copy()
    Qh[1] = [0.5, 1, 1]
    Qh[2:, 0] = rng.uniform(5 / n, 0.5, 24).astype(F32)
    rho_h, _, _ = D.sph_density(P, Qh, h, 64, 512, "x64", True, kernel="wendland")
    np.testing.assert_allclose(rho_h[2:], 0.5, atol=3e-4)


def test_sph_normalised_unit_lattice_density_is_one_2d():
    """dimensions='2d': the slab |z - 0.5| <= 0.05 of a 3-D point set holds
    one n x n layer (sph.py:164-177); its area density is 1/a^2 and the 2-D
    Wendland kernel (h = 4a) sums to it within 1.2e-3, so the density
    normalised by N / area(bbox of the queries) (sph.py:366-371) is 1.
    Queries given with three columns are projected on the plane."""
    n = 32
    L2 = _lattice(n, 2)
    layers = [np.concatenate([L2, np.full((len(L2), 1), zv)], 1) for zv in (0.2, 0.5, 0.8)]
    P3 = np.concatenate(layers).astype(F32)
    rng = np.random.default_rng(34)
    Q = np.concatenate([[[0, 0, 0.9], [1, 1, 0.1]], np.c_[rng.uniform(5 / n, 1 - 5 / n, (40, 2)), rng.random(40)]])
    rho, _, _ = D.sph_density(P3, Q.astype(F32), 4.0 / n, 64, 512, "x64", True, kernel="wendland",
                              dimensions="2d", position=0.5, slab_thickness=0.1)
    np.testing.assert_allclose(rho[2:], 1.0, atol=2e-3)


def test_extract_2d_slab_and_fallback():
    P3 = np.array([[0.1, 0.2, 0.5], [0.3, 0.4, 0.9], [0.5, 0.6, 0.46]], F32)
    np.testing.assert_array_equal(D.extract_2d(P3, "xy", 0.5, 0.1), P3[[0, 2]][:, :2])
    np.testing.assert_array_equal(D.extract_2d(P3, "yz", 0.3, 0.01), P3[[1]][:, [1, 2]])
    # empty slab -> every point projected
    np.testing.assert_array_equal(D.extract_2d(P3, "xz", 5.0, 0.1), P3[:, [0, 2]])
