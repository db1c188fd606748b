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
