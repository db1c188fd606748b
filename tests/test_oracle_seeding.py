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


# ---- random_seeds (seeding.py:140-182) and the inlet-face generator --------

def test_threefry2x32_known_answers():
    """Random123's known-answer vectors for threefry2x32 with 20 rounds
    (key, counter) -> output."""
    cases = [((0, 0), (0, 0), (0x6B200159, 0x99BA4EFE)),
             ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
             ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0))]
    for (k0, k1), (c0, c1), (y0, y1) in cases:
        a, b = O.threefry2x32(k0, k1, [c0], [c1])
        assert (int(a[0]), int(b[0])) == (y0, y1)


def test_jax_uniform_layout():
    """jax.random.uniform(PRNGKey(0)) = 0.41845703 (the JAX <= 0.4 value: the
    first threefry word of counter (0, 0) under key (0, 0), top 23 bits as a
    mantissa in [1, 2), minus 1); the halves layout: word i and word h + i
    come from one block; values in [0, 1) with 2^-23 resolution."""
    assert O.uniform01(O.prng_key(0), ()) == np.float32(0.41845703)
    m = 7
    bits = O.random_bits32(O.prng_key(5), m)
    y0, y1 = O.threefry2x32(0, 5, np.arange(4, dtype=np.uint32), np.array([4, 5, 6, 0], np.uint32))
    np.testing.assert_array_equal(bits, np.concatenate([y0, y1[:3]]))
    u = O.uniform01(O.prng_key(123), (1000, 3))
    assert u.min() >= 0 and u.max() < 1 and np.all(u * 2 ** 23 == np.floor(u * 2 ** 23))
    assert abs(float(u.mean()) - 0.5) < 0.02
    assert O.prng_key(2 ** 32 + 7) == (1, 7)


def test_random_seeds_scaled_to_bounds():
    b = [[-1.0, 2.0, 0.5], [3.0, 2.5, 0.75]]
    x = O.random_seeds(5000, b, rng_seed=9)
    assert x.shape == (5000, 3) and x.dtype == np.float32
    lo, hi = np.array(b[0], np.float32), np.array(b[1], np.float32)
    assert np.all(x >= lo) and np.all(x <= hi)
    u = O.uniform01(O.prng_key(9), (5000, 3))
    np.testing.assert_array_equal(x, lo + u * (hi - lo))
    assert O.random_seeds(0, b).shape == (0, 3)


def test_inlet_grid_particles_closed_forms():
    """boundary.py:1321-1441: fewer particles than face points are systematic
    samples i * total // n of the 'ij' face grid (distinct, increasing); more
    particles use every face point once, then the extras; the flow axis holds
    the inlet position; off-axis values are face-grid coordinates."""
    b = [[0, 0, 0], [1, 2, 4]]
    conc = {"x": 5, "y": 3, "z": 5}
    p = O.inlet_grid_particles(7, 0.25, 1, b, conc)           # face axes x (5) and z (5): 25 points
    assert np.all(p[:, 1] == np.float32(0.25))
    xs, zs = np.linspace(0, 1, 5, dtype=np.float32), np.linspace(0, 4, 5, dtype=np.float32)
    want = [(xs[k // 5], zs[k % 5]) for k in (int(i * 25 / 7) for i in range(7))]
    np.testing.assert_array_equal(p[:, [0, 2]], np.array(want, np.float32))
    q = O.inlet_grid_particles(30, 1.0, 0, b, conc)           # face axes y (3) and z (5): 15 points
    face = {(float(y), float(z)) for y in np.linspace(0, 2, 3, dtype=np.float32)
            for z in np.linspace(0, 4, 5, dtype=np.float32)}
    assert {(float(a), float(c)) for a, c in q[:15, 1:]} == face
    assert len(O.inlet_grid_indices(15, 15)) == 15 and list(O.inlet_grid_indices(15, 15)) == list(range(15))
