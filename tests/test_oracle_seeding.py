"""CPU checks of the seeding oracle and the host side of uniform_grid_seeds."""
import numpy as np
import pytest

from jaxtrace_amd import seeding as S
from oracle import seeding as O

F32 = np.float32


def test_expand_matches_triple_loop():
    rng = np.random.default_rng(0)
    xc, yc, zc = (np.sort(rng.standard_normal(k)).astype(F32) for k in (4, 3, 5))
    rows = [(x, y, z) for x in xc for y in yc for z in zc]      # i slowest, k fastest
    np.testing.assert_array_equal(O.expand(xc, yc, zc), np.array(rows, dtype=F32))


@pytest.mark.parametrize("n", [1, 2, 7, 33])
def test_integer_lattices(n):
    got = O.uniform_grid_seeds(n, [[0, 0, 0], [max(n - 1, 1)] * 3])
    i = np.arange(n, dtype=F32) if n > 1 else np.zeros(1, F32)
    assert got.shape == (n ** 3, 3)
    np.testing.assert_array_equal(got[:, 0], np.repeat(i, n * n))
    np.testing.assert_array_equal(got[:, 2], np.tile(i, n * n))
    inner = O.uniform_grid_seeds((n, n, n), [[0, 0, 0], [n + 1] * 3], include_boundaries=False)
    np.testing.assert_array_equal(np.unique(inner[:, 1]), np.arange(1, n + 1, dtype=F32))


def test_corners_and_planar_resolution():
    b = [[-0.3, 0.1, 2.0], [1.7, 0.9, 2.5]]
    g = O.uniform_grid_seeds((5, 4, 3), b)
    np.testing.assert_array_equal(g[0], np.asarray(b[0], F32))
    np.testing.assert_array_equal(g[-1], np.asarray(b[1], F32))
    p = O.uniform_grid_seeds((6, 2), b)
    assert p.shape == (12, 3) and np.all(p[:, 2] == F32(2.0))


def test_host_axes_equal_oracle_axes():
    for res, inc in [(9, True), ((5, 4, 3), False), ((6, 2), True), ((6, 2), False)]:
        b = [[-0.3, 0.1, 2.0], [1.7, 0.9, 2.5]]
        for got, want in zip(S.grid_axes(res, b, inc), O.axes(res, b, inc)):
            np.testing.assert_array_equal(got, want)


def test_bounds_forms_and_errors():
    want = np.array([[0, 1, 2], [3, 4, 5]], F32)
    np.testing.assert_array_equal(S._validate_bounds([0, 3, 1, 4, 2, 5]), want)
    np.testing.assert_array_equal(S._validate_bounds([[0, 3], [1, 4], [2, 5]]), want)
    # the 2-D forms set z min = z max = 0, which the reference's own min < max
    # check rejects (seeding.py:104-107, :116-119, :133-134)
    with pytest.raises(ValueError):
        S._validate_bounds([[0, 3], [1, 4]])
    with pytest.raises(ValueError):
        S._validate_bounds([0, 3, 1, 4])
    with pytest.raises(ValueError):
        S._validate_bounds([[1, 0, 0], [0, 1, 1]])
    with pytest.raises(ValueError):
        S._validate_bounds([0, 1, 2])
    with pytest.raises(ValueError):
        S.grid_axes((2, 2, 2, 2), [[0, 0, 0], [1, 1, 1]])
