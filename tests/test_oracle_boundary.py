"""Pins the oracle's boundary conditions (oracle/boundary.py) against
closed forms and invariants."""
import numpy as np
import pytest

from oracle import boundary as B

F32 = np.float32


def test_periodic_closed_forms():
    x = np.array([-0.25, 1.25, 0.5, 3.0, -2.75], F32)
    np.testing.assert_array_equal(B.periodic(x, 0.0, 1.0), np.array([0.75, 0.25, 0.5, 0.0, 0.25], F32))
    np.testing.assert_array_equal(B.periodic(np.array([5.0], F32), -1.0, 3.0), [1.0])
    # x - lo = -eps rounds up to exactly hi
    assert B.periodic(np.array([-1e-9], F32), 0.0, 1.0)[0] == 1.0


def test_reflective_closed_forms():
    x = np.array([1.25, -0.25, 2.25, 0.5, 1.0, -1.0], F32)
    np.testing.assert_array_equal(B.reflective(x, 0.0, 1.0), np.array([0.75, 0.25, 0.25, 0.5, 1.0, 1.0], F32))


def test_clamp_and_absorb():
    x = np.array([-1.0, 0.5, 2.0, np.nan], F32)
    got = B.clamp(x, 0.0, 1.0)
    np.testing.assert_array_equal(got[:3], [0.0, 0.5, 1.0])
    assert np.isnan(got[3])
    rows = np.array([[0.5, 0.5, 0.5], [0.5, 1.5, 0.5]], F32)
    r = B.absorb_rows(rows, [[0, 0, 0], [1, 1, 1]])
    np.testing.assert_array_equal(r[0], rows[0])
    assert np.isnan(r[1]).all()
    m = B.per_axis(rows, [[0, 0, 0], [1, 1, 1]], ["none", "absorbing", "none"])
    assert np.isnan(m[1, 1]) and m[1, 0] == 0.5 and m[1, 2] == 0.5


@pytest.mark.parametrize("kind", ["periodic", "reflective", "clamping"])
def test_range_and_idempotence(kind):
    rng = np.random.default_rng(0)
    b = [[-1.0, 0.0, 2.0], [1.0, 0.5, 3.5]]
    x = rng.uniform(-6, 6, (2000, 3)).astype(F32)
    y = B.per_axis(x, b, [kind] * 3)
    lo, hi = np.array(b[0], F32), np.array(b[1], F32)
    assert ((y >= lo) & (y <= hi)).all()
    inner = y[((y > lo) & (y < hi)).all(1)]
    np.testing.assert_array_equal(B.per_axis(inner, b, [kind] * 3), inner)


def test_bounds_parsing():
    np.testing.assert_array_equal(B.ensure_bounds([0, 1, 2, 3, 4, 5]), [[0, 2, 4], [1, 3, 5]])
    np.testing.assert_array_equal(B.ensure_bounds([[0, 1], [2, 3], [4, 5]]), [[0, 2, 4], [1, 3, 5]])
    with pytest.raises(ValueError):
        B.ensure_bounds([[1, 1, 1], [0, 2, 2]])


def test_continuous_inlet_grid_snap():
    bounds = [[0, 0, 0], [1, 1, 1]]
    x = np.array([[1.2, 0.3012, 0.996], [0.5, -0.1, 1.2], [-0.2, 0.5, 0.5], [0.4, 0.4, 0.4]], F32)
    y = B.continuous_inlet(x, bounds, "x", "positive", "grid", {"x": 64, "y": 64, "z": 64})
    grid = np.linspace(F32(0), F32(1), 64, dtype=F32)
    # exited: x -> inlet, y/z -> nearest inlet-grid node (k/63)
    np.testing.assert_array_equal(y[0], [0.0, grid[19], grid[63]])
    assert round(0.3012 * 63) == 19
    # reflected off-axis coordinates, x untouched
    np.testing.assert_allclose(y[1], [0.5, 0.1, 0.8], atol=1e-7)
    # backstop for particles behind the inlet
    assert y[2, 0] == F32(0.001)
    np.testing.assert_array_equal(y[3], x[3])


def test_oracle_random_inlet_invariants():
    """The random replacements of the oracle's continuous inlet
    (boundary.py:1552-1571): exited particles land on the inlet face with
    off-axis coordinates inside the face ([lo, hi) for 'uniform', clipped to
    [lo, hi] for 'gaussian'); the others are untouched; exactly 2 * n_exited
    values are drawn from the global np.random state, and the lower off-axis
    receives the first n_exited of them (checked against a second draw from
    the same seed); with no exit nothing is drawn."""
    b = [[0.0, -1.0, 0.5], [1.0, 1.0, 2.5]]
    rng = np.random.default_rng(3)
    x = rng.uniform([0.1, -0.9, 0.6], [0.9, 0.9, 2.4], (400, 3)).astype(np.float32)
    x[::7, 0] = 1.25                                    # 58 exit through x > 1
    n_ex = int((x[:, 0] > 1.0).sum())
    for dist in ("uniform", "gaussian"):
        np.random.seed(11)
        got = B.continuous_inlet(x, b, "x", "positive", dist)
        after = np.random.random()
        np.random.seed(11)
        if dist == "uniform":
            first = np.random.uniform(0.0, 1.0, 2 * n_ex)
        else:
            first = np.random.normal(0.0, 1.0, 2 * n_ex)
        assert np.random.random() == after              # 2 * n_ex draws, no more
        ex = x[:, 0] > 1.0
        np.testing.assert_array_equal(got[~ex], x[~ex])
        assert np.all(got[ex, 0] == 0.0)
        lo, hi = np.array(b[0][1:]), np.array(b[1][1:])
        assert np.all(got[ex, 1:] >= lo) and np.all(got[ex, 1:] <= hi)
        # the y column comes from the first n_ex standard draws, z from the rest
        if dist == "uniform":
            np.testing.assert_allclose(got[ex, 1], -1.0 + 2.0 * first[:n_ex], rtol=0, atol=1e-6)
            np.testing.assert_allclose(got[ex, 2], 0.5 + 2.0 * first[n_ex:], rtol=0, atol=1e-6)
        else:
            np.testing.assert_allclose(got[ex, 1], np.clip(0.0 + (2.0 / 6) * first[:n_ex], -1, 1), atol=1e-6)
            np.testing.assert_allclose(got[ex, 2], np.clip(1.5 + (2.0 / 6) * first[n_ex:], 0.5, 2.5), atol=1e-6)
    np.random.seed(5)
    before = np.random.get_state()[1].copy()
    inside = x.copy()
    inside[:, 0] = 0.5
    B.continuous_inlet(inside, b, "x", "positive", "uniform")
    np.testing.assert_array_equal(np.random.get_state()[1], before)
