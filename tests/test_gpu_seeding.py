"""GPU parity of uniform_grid_seeds (jt_grid_seeds) against the oracle's
np.meshgrid expansion on the same coordinate arrays: BITWISE, at small
ragged sizes, in the planar and interior variants, and at 256^3 seeds."""
import numpy as np
import pytest
import torch

import jaxtrace_amd as jt
from oracle import seeding as O

pytestmark = pytest.mark.gpu
F32 = np.float32
B = [[-0.3, 0.1, 2.0], [1.7, 0.9, 2.5]]


@pytest.mark.parametrize("res,inc", [(1, True), (2, True), (7, False), ((5, 4, 3), True), ((5, 4, 3), False),
                                     ((6, 2), True), ((1, 9, 1), True), ((33, 17, 65), False)])
def test_grid_seeds_bitwise(gpu, res, inc):
    got = jt.uniform_grid_seeds(res, B, include_boundaries=inc)
    want = O.uniform_grid_seeds(res, B, include_boundaries=inc)
    assert got.dtype == np.float32 and got.shape == want.shape
    np.testing.assert_array_equal(got, want)


def test_grid_seeds_empty_and_device_output(gpu):
    assert jt.uniform_grid_seeds((0, 4, 4), B).shape == (0, 3)
    d = jt.uniform_grid_seeds(8, [[0, 0, 0], [1, 1, 1]], return_device=True)
    assert d.is_cuda and tuple(d.shape) == (512, 3)
    np.testing.assert_array_equal(d.cpu().numpy(), O.uniform_grid_seeds(8, [[0, 0, 0], [1, 1, 1]]))


def test_grid_seeds_full_size(gpu):
    """256^3 = 2^24 seeds in the unit box (the tracking workload's particle
    count), bitwise against the oracle."""
    d = jt.uniform_grid_seeds(256, [[0, 0, 0], [1, 1, 1]], return_device=True)
    want = O.uniform_grid_seeds(256, [[0, 0, 0], [1, 1, 1]])
    assert torch.equal(d.cpu(), torch.from_numpy(want))


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 42), (1000, 7), (333_333, 2 ** 33 + 5), (1 << 22, 3)])
def test_random_seeds_bitwise(gpu, n, seed):
    """random_seeds (seeding.py:140-182): the jt_random_seeds kernel's
    threefry words, uniforms and scaled positions equal the oracle's bitwise
    (odd and even word counts, a seed above 2^32)."""
    b = [[-1.0, 0.0, 2.0], [1.0, 0.5, 6.0]]
    got = jt.random_seeds(n, b, rng_seed=seed)
    np.testing.assert_array_equal(got, O.random_seeds(n, b, rng_seed=seed))


def test_random_seeds_edges(gpu):
    assert jt.random_seeds(0, [[0, 0, 0], [1, 1, 1]]).shape == (0, 3)
    t = jt.random_seeds(10, [0, 1, 0, 2, 0, 3], rng_seed=1, return_device=True)
    assert t.is_cuda and t.shape == (10, 3)
    with pytest.raises(ValueError):
        jt.random_seeds(4, [[0, 0, 0], [0, 1, 1]])


@pytest.mark.parametrize("n,axis,conc", [(7, 1, {"x": 5, "y": 3, "z": 5}), (4096, 0, {"x": 64, "y": 64, "z": 64}),
                                         (5000, 2, {"x": 40, "y": 50, "z": 1}), (1, 0, {"x": 1, "y": 1, "z": 1}),
                                         (100, 2, {"x": 7, "y": 9, "z": 3})])
def test_inlet_grid_particles_bitwise(gpu, n, axis, conc):
    """_generate_inlet_grid_particles (boundary.py:1321-1441): fewer, equal
    and more particles than inlet-face points, each flow axis."""
    import jaxtrace_amd as jt
    b = [[0.0, -1.0, 0.5], [2.0, 1.0, 1.5]]
    got = jt.generate_inlet_grid_particles(n, 0.125, axis, b, conc)
    np.testing.assert_array_equal(got, O.inlet_grid_particles(n, 0.125, axis, b, conc))
