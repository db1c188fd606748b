"""GPU parity on the benchmark's headline tracking workload at full size.

Random Fourier field on a 256^3 periodic lattice x 16 snapshots at
t_j = 2j/15 (seed 2), 2^24 particles uniform in the unit box (torch CPU
generator, seed 3), RK4, dt = 2e-3, x periodic / y,z reflective walls, in
float32, bfloat16 and float16 snapshot storage.  100 steps from t = 0 run as
four 25-step launches with the cell sort between them (the library
defaults at this size), and the snapshot pair changes at t = 2/15 (between
steps 66 and 67).  For 4096 sampled particles every step is compared BITWISE
with the oracle's step from the GPU's own previous state (the kernels do the
oracle's float32 operations in its order), so a wrong corner, weight, blend
or boundary anywhere in the run fails at the step where it happens.
"""
import numpy as np
import pytest
import torch

import jaxtrace_amd as jt
from jaxtrace_amd import workloads as W
from oracle import boundary as OB
from oracle import fields as OF
from oracle import integrators as OI
from oracle import tracker as OT

pytestmark = pytest.mark.gpu
F32 = np.float32
N_PART = 1 << 24
KINDS = ["periodic", "reflective", "reflective"]
BOX = [[0, 0, 0], [1, 1, 1]]


@pytest.fixture(scope="module")
def headline(gpu):
    data, times, g = W.random_fourier(n=256, n_snap=16, t_end=2.0, seed=2, device=gpu)
    gen = torch.Generator(device="cpu").manual_seed(3)
    x0 = torch.rand((N_PART, 3), generator=gen, dtype=torch.float32).to(gpu)
    yield data, times, g, x0
    del data
    torch.cuda.empty_cache()


@pytest.mark.parametrize("storage", ["float32", "bfloat16", "float16"])
def test_headline_workload_sampled_per_step_bitwise(headline, storage):
    data, times, g, x0 = headline
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic", storage=storage,
                                     layout="halo")
    bc = jt.mixed_boundary(BOX, dict(zip("xyz", KINDS)))
    tr = jt.create_tracker("rk4", f, bc, return_device=True)
    traj = tr.track_particles(x0, (0.0, 0.2), 101, dt=2e-3)
    assert traj.metadata["n_particles"] == N_PART
    idx = np.sort(np.random.default_rng(41).choice(N_PART, 4096, replace=False))
    sample = traj.positions[:, torch.from_numpy(idx).to(traj.positions.device)].cpu().numpy()
    p = traj.positions
    assert bool(((p >= 0) & (p <= 1)).all())
    del traj, p, tr
    torch.cuda.empty_cache()
    rounded = {"float32": torch.float32, "bfloat16": torch.bfloat16, "float16": torch.float16}[storage]
    host = data.to(rounded).to(torch.float32).cpu().numpy()
    meta = OF.GridMeta(g, g, g)
    field_fn = lambda q, t: OF.structured_series_sample_corners(host, times, meta, q, t, "periodic")
    tgrid, dt = OT.time_grid((0.0, 0.2), 101, 2e-3)
    pairs = set()
    for s in range(100):
        pairs.add(OF.blend_params(times, tgrid[s])[1:3])
        nxt = OB.per_axis(OI.rk4_step(sample[s], tgrid[s], dt, field_fn), BOX, KINDS)
        np.testing.assert_array_equal(sample[s + 1], nxt, err_msg=f"{storage} step {s}")
    assert pairs == {(0, 1), (1, 2)}          # the run crosses a snapshot change
