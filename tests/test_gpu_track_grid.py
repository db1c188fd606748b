"""GPU parity of the structured-grid hot path (jt_grid_sample, jt_apply_boundary,
jt_track_grid) against the float32 oracle on identical seeded inputs.

Bar: the kernels perform the oracle's IEEE float32 operations in the same
order (library built with -ffp-contract=off), so sampled velocities, boundary
results and per-step positions are compared BITWISE; whole trajectories are
additionally held to 1e-6 relative per step from identical states."""
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

STORAGES = [("float32", "packed"), ("float32", "padded"), ("bfloat16", "packed"), ("float16", "packed"),
            ("float32", "halo"), ("float32", "halo3"), ("float32", "halo4"), ("bfloat16", "halo"),
            ("float16", "halo")]


def _rounded(data, storage):
    t = torch.from_numpy(data)
    if storage == "bfloat16":
        return t.to(torch.bfloat16).to(torch.float32).numpy()
    if storage == "float16":
        return t.to(torch.float16).to(torch.float32).numpy()
    return data


def _series(seed=0, T=5, shape=(9, 7, 6)):
    rng = np.random.default_rng(seed)
    data = rng.standard_normal((T,) + shape + (3,)).astype(F32)
    times = np.sort(rng.uniform(0, 2, T)).astype(F32)
    gx = (np.arange(shape[0]) * 0.25 - 0.5).astype(F32)
    gy = (np.arange(shape[1]) / shape[1]).astype(F32)
    gz = np.linspace(0.1, 0.9, shape[2]).astype(F32)
    return data, times, (gx, gy, gz)


@pytest.mark.parametrize("storage,layout", STORAGES)
@pytest.mark.parametrize("mode", ["clamp", "zero", "nan", "periodic"])
def test_sample_bitwise(gpu, storage, layout, mode):
    data, times, (gx, gy, gz) = _series()
    f = jt.StructuredTimeSeriesField(data, times, gx, gy, gz, bounds_handling=mode,
                                     interpolation="linear", extrapolation="linear",
                                     storage=storage, layout=layout)
    meta = OF.GridMeta(gx, gy, gz)
    rng = np.random.default_rng(1)
    P = np.stack([rng.uniform(-1.2, 2.0, 2000), rng.uniform(-0.3, 1.3, 2000),
                  rng.uniform(-0.2, 1.1, 2000)], 1).astype(F32)
    P[:50] = np.stack([gx[rng.integers(0, 9, 50)], gy[rng.integers(0, 7, 50)],
                       gz[rng.integers(0, 6, 50)]], 1)           # exact nodes
    ref_data = _rounded(data, storage)
    for t in [times[0] - 0.5, times[0], 0.5 * (times[1] + times[2]), times[3], times[-1] + 0.25]:
        got = f.sample_at_positions(P, float(t))
        want = OF.structured_series_sample(ref_data, times, meta, P, F32(t), mode, "linear", "linear")
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("interp,extrap", [("linear", "constant"), ("linear", "zero"),
                                           ("linear", "nan"), ("nearest", "constant")])
def test_sample_time_modes_bitwise(gpu, interp, extrap):
    data, times, (gx, gy, gz) = _series(seed=3)
    f = jt.StructuredTimeSeriesField(data, times, gx, gy, gz, "clamp", interp, extrap)
    meta = OF.GridMeta(gx, gy, gz)
    P = np.random.default_rng(2).uniform(-0.5, 1.5, (500, 3)).astype(F32)
    for t in [-3.0, times[1], 0.77, times[-1], 5.0]:
        want = OF.structured_series_sample(data, times, meta, P, F32(t), "clamp", interp, extrap)
        np.testing.assert_array_equal(f.sample_at_positions(P, t), want)


def test_time_independent_sampler(gpu):
    data, _, (gx, gy, gz) = _series(seed=4, T=1)
    s = jt.StructuredGridSampler(data[0], gx, gy, gz, bounds_handling="periodic")
    P = np.random.default_rng(5).uniform(-2, 2, (777, 3)).astype(F32)
    np.testing.assert_array_equal(s.sample(P), OF.trilinear(data[0], OF.GridMeta(gx, gy, gz), P, "periodic"))


BOUNDS = [[-1.0, 0.0, 0.25], [1.0, 0.5, 2.0]]


@pytest.mark.parametrize("kinds", [("periodic",) * 3, ("reflective",) * 3, ("clamping",) * 3,
                                   ("periodic", "reflective", "absorbing"),
                                   ("absorbing", "clamping", "periodic")])
def test_boundary_bitwise(gpu, kinds):
    x = np.random.default_rng(6).uniform(-4, 4, (5000, 3)).astype(F32)
    x[:3] = [[1.0, 0.5, 2.0], [-1.0, 0.0, 0.25], [-1e-9, 1e-9, 0.25]]
    if len(set(kinds)) == 1:
        bc = {"periodic": jt.periodic_boundary, "reflective": jt.reflective_boundary,
              "clamping": jt.clamping_boundary}[kinds[0]](BOUNDS)
    else:
        bc = jt.mixed_boundary(BOUNDS, dict(zip("xyz", kinds)))
    np.testing.assert_array_equal(bc(x), OB.per_axis(x, BOUNDS, list(kinds)))


def test_absorbing_rows_bitwise(gpu):
    x = np.random.default_rng(7).uniform(-2, 2, (3000, 3)).astype(F32)
    bc = jt.absorbing_boundary_factory(BOUNDS)
    np.testing.assert_array_equal(bc(x), OB.absorb_rows(x, BOUNDS))


@pytest.mark.parametrize("direction", ["positive", "negative"])
def test_continuous_inlet_bitwise(gpu, direction):
    b = [[0, 0, 0], [1, 1, 1]]
    conc = {"x": 64, "y": 64, "z": 48}
    x = np.random.default_rng(8).uniform(-0.3, 1.3, (4000, 3)).astype(F32)
    bc = jt.continuous_inlet_boundary_factory(b, "y", direction, "grid", conc)
    np.testing.assert_array_equal(bc(x), OB.continuous_inlet(x, b, "y", direction, "grid", conc))


def _cfg1(n=10_000):
    data, times, g = W.abc_flow()
    x0 = W.uniform_particles(n, 42)
    return data, times, g, x0


@pytest.mark.parametrize("storage,layout", [("float32", "packed"), ("bfloat16", "packed"), ("float16", "packed"),
                                            ("float32", "halo"), ("bfloat16", "halo"), ("float16", "halo")])
def test_cfg1_rk4_periodic_per_step_parity(gpu, storage, layout):
    """Config: ABC flow 32^3 x 4 snapshots, 10,000 particles (seed 42), RK4,
    dt = 0.01, 100 steps, periodic unit box."""
    data, times, g, x0 = _cfg1()
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic", storage=storage,
                                     layout=layout)
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    traj = jt.create_tracker("rk4", f, bc).track_particles(x0, (0.0, 1.0), 101, dt=0.01)
    assert traj.positions.shape == (101, 10_000, 3)
    meta = OF.GridMeta(g, g, g)
    ref = _rounded(data, storage)
    field_fn = lambda p, t: OF.structured_series_sample_corners(ref, times, meta, p, t, "periodic")
    tgrid, dt = OT.time_grid((0.0, 1.0), 101, 0.01)
    np.testing.assert_array_equal(traj.times, tgrid)
    worst = 0.0
    for i in range(100):
        nxt = OB.per_axis(OI.rk4_step(traj.positions[i], tgrid[i], dt, field_fn), [[0, 0, 0], [1, 1, 1]],
                          ["periodic"] * 3)
        d = np.abs(traj.positions[i + 1].astype(np.float64) - nxt)
        worst = max(worst, float((d / np.maximum(np.abs(nxt), 1e-6)).max()))
        np.testing.assert_array_equal(traj.positions[i + 1], nxt)
    assert worst <= 1e-6


@pytest.mark.parametrize("integ", ["euler", "rk2"])
def test_cfg1_euler_rk2_bitwise(gpu, integ):
    data, times, g, x0 = _cfg1(2000)
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic")
    bc = jt.reflective_boundary([[0, 0, 0], [1, 1, 1]])
    traj = jt.create_tracker(integ, f, bc).track_particles(x0, (0.0, 0.2), 21)
    meta = OF.GridMeta(g, g, g)
    field_fn = lambda p, t: OF.structured_series_sample_corners(data, times, meta, p, t, "periodic")
    tt, pos = OT.track(x0, (0.0, 0.2), 21, field_fn,
                       lambda p: OB.per_axis(p, [[0, 0, 0], [1, 1, 1]], ["reflective"] * 3), integ)
    np.testing.assert_array_equal(traj.positions, pos)


def test_recording_options_and_edges(gpu):
    data, times, g, x0 = _cfg1(1000)
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic")
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    full = jt.create_tracker("rk4", f, bc).track_particles(x0, (0, 1), 21).positions
    every5 = jt.create_tracker("rk4", f, bc, record_every=5).track_particles(x0, (0, 1), 21)
    np.testing.assert_array_equal(every5.positions, full[::5])
    assert every5.times.shape == (5,)
    fin = jt.create_tracker("rk4", f, bc, record_final_only=True, steps_per_launch=3)
    fin = fin.track_particles(x0, (0, 1), 21)
    np.testing.assert_array_equal(fin.positions, full[[0, 20]])
    one = jt.create_tracker("rk4", f, bc).track_particles(x0[:1], (0, 1), 1)
    assert one.positions.shape == (1, 1, 3)
    empty = jt.create_tracker("rk4", f, bc).track_particles(np.zeros((0, 3), F32), (0, 1), 11)
    assert empty.positions.shape == (11, 0, 3)
    dev = jt.create_tracker("rk4", f, bc, return_device=True, record_velocities=True)
    tr = dev.track_particles(torch.from_numpy(x0[:, :2].copy()), (0, 1), 6)   # 2-D input -> z = 0
    assert tr.positions.is_cuda and float(tr.positions[0, :, 2].abs().max()) == 0.0
    np.testing.assert_array_equal(tr.velocities[3].cpu().numpy(),
                                  f.sample_at_positions(tr.positions[3].cpu().numpy(), float(tr.times[3])))


def test_large_mixed_bc_properties(gpu):
    """2^22 particles in a 128^3 x 8 random Fourier field, mixed x-periodic /
    y,z-reflective walls: bounds hold, reruns are bitwise identical, and 4096
    sampled particles match the oracle step bitwise."""
    dev = torch.device("cuda")
    data, times, g = W.random_fourier(n=128, n_snap=8, t_end=1.0, seed=2, device=dev)
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic")
    x0 = torch.rand((1 << 22, 3), generator=torch.Generator().manual_seed(3)).numpy()
    bc = jt.mixed_boundary([[0, 0, 0], [1, 1, 1]], {"x": "periodic", "y": "reflective", "z": "reflective"})
    tr = jt.create_tracker("rk4", f, bc, return_device=True)
    a = tr.track_particles(x0, (0.0, 0.04), 21, dt=2e-3)
    b = tr.track_particles(x0, (0.0, 0.04), 21, dt=2e-3)
    assert torch.equal(a.positions, b.positions)
    p = a.positions
    assert bool(((p >= 0) & (p <= 1)).all())
    host = data.cpu().numpy()
    meta = OF.GridMeta(g, g, g)
    idx = np.random.default_rng(9).choice(1 << 22, 4096, replace=False)
    field_fn = lambda q, t: OF.structured_series_sample_corners(host, times, meta, q, t, "periodic")
    tgrid, dt = OT.time_grid((0.0, 0.04), 21, 2e-3)
    for s in (0, 7, 19):
        cur = p[s].cpu().numpy()[idx]
        nxt = OB.per_axis(OI.rk4_step(cur, tgrid[s], dt, field_fn), [[0, 0, 0], [1, 1, 1]],
                          ["periodic", "reflective", "reflective"])
        np.testing.assert_array_equal(p[s + 1].cpu().numpy()[idx], nxt)


@pytest.mark.parametrize("layout", ["packed", "halo", "halo3"])
def test_cell_sorting_is_invisible(gpu, layout):
    """Re-ordering particles by cell every k steps (and writing trajectories
    back through the permutation) leaves every recorded position bitwise
    unchanged."""
    data, times, g, x0 = _cfg1(5000)
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic", layout=layout)
    bc = jt.mixed_boundary([[0, 0, 0], [1, 1, 1]], {"x": "periodic", "y": "reflective", "z": "reflective"})
    a = jt.create_tracker("rk4", f, bc).track_particles(x0, (0, 1), 41)
    b = jt.create_tracker("rk4", f, bc, sort_every=7, steps_per_launch=5).track_particles(x0, (0, 1), 41)
    np.testing.assert_array_equal(a.positions, b.positions)


@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_cell_sorting_odd_grid_final_only(gpu, storage):
    """Cell sort on a 45 x 19 x 70 grid (keys of 7 bits per axis), launches
    longer than the sort interval, final positions only: the result equals
    the unsorted run bitwise and comes back in the caller's order."""
    rng = np.random.default_rng(31)
    gx, gy, gz = (np.linspace(0, 1, m, dtype=F32) for m in (45, 19, 70))
    times = np.array([0.0, 0.5, 1.0], F32)
    data = (0.3 * rng.standard_normal((3, 45, 19, 70, 3))).astype(F32)
    f = jt.StructuredTimeSeriesField(data, times, gx, gy, gz, bounds_handling="clamp", storage=storage,
                                     layout="halo")
    bc = jt.reflective_boundary([[0, 0, 0], [1, 1, 1]])
    x0 = W.uniform_particles(70_001, 5)
    a = jt.create_tracker("rk4", f, bc, record_final_only=True, sort_every=0).track_particles(x0, (0, 1), 33)
    for se, spl in ((3, 10), (8, 8), (5, 1)):
        b = jt.create_tracker("rk4", f, bc, record_final_only=True, sort_every=se,
                              steps_per_launch=spl).track_particles(x0, (0, 1), 33)
        np.testing.assert_array_equal(a.positions, b.positions)


def test_time_series_field_nearest_node(gpu):
    """TimeSeriesField (time_series.py:82-473): blend every node, take the
    nearest node's value; exhaustive argmin with ties to the lowest index."""
    rng = np.random.default_rng(12)
    nodes = rng.uniform(0, 1, (3000, 3)).astype(F32)
    nodes[100] = nodes[7]                       # duplicate node: tie -> index 7
    times = np.array([0.0, 0.4, 1.0], F32)
    data = rng.standard_normal((3, 3000, 3)).astype(F32)
    f = jt.TimeSeriesField(data, times, nodes)
    Q = np.concatenate([rng.uniform(-0.1, 1.1, (700, 3)), nodes[[7, 100]]]).astype(F32)
    for t in (-1.0, 0.0, 0.25, 1.0, 3.0):
        want = OF.nearest_node_sample(data, times, nodes, Q, F32(t))
        np.testing.assert_array_equal(f.sample_at_positions(Q, t), want)
    # tracking through the nearest-node field, bitwise per step
    bc = jt.clamping_boundary([[0, 0, 0], [1, 1, 1]])
    x0 = W.uniform_particles(500, 13)
    traj = jt.create_tracker("rk4", f, bc).track_particles(x0, (0.0, 0.5), 11)
    fn = lambda p, t: OF.nearest_node_sample(data, times, nodes, p, t)
    tt, pos = OT.track(x0, (0.0, 0.5), 11, fn, lambda p: OB.per_axis(p, [[0, 0, 0], [1, 1, 1]], ["clamping"] * 3))
    np.testing.assert_array_equal(traj.positions, pos)


@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_periodic_fast_path_edges(gpu, storage):
    """The periodic sampler on a power-of-two grid with the z-halo layout
    takes the kernel's periodic fast path; the packed layout takes the general
    index code.  Start points outside the box (wrapped by the sampler only:
    the boundary is open), a hair below 0 (jnp.mod rounds to n), at 1, beyond
    two periods (the general fmod path), -0.0 and NaN, tracked through RK4:
    both layouts must reproduce the oracle bitwise, step by step."""
    from jaxtrace_amd.boundary import BoundaryCondition
    data, times, g, x0 = _cfg1(3000)
    x0[:10] = [[-1e-9, 0.5, 0.25], [1.0, 1.0 - 2 ** -24, 0.0], [2.75, -3.5, 7.0], [-0.0, -0.0, -0.0],
               [-1.0 + 1e-7, 2.0, -2.0 + 1e-7], [np.nan, 0.5, 0.5], [0.5, np.nan, 0.5], [1e6, -1e6, 3.0],
               [31.0 / 32, 1.0 / 32, 0.5 - 2 ** -26], [-2 ** -30, 1.0 + 2 ** -23, -0.5]]
    ref = _rounded(data, storage)
    meta = OF.GridMeta(g, g, g)
    field_fn = lambda p, t: OF.structured_series_sample_corners(ref, times, meta, p, t, "periodic")
    open_bc = BoundaryCondition("_open_bc", np.array([[0, 0, 0], [1, 1, 1]], F32))
    tt, pos = OT.track(x0, (0.0, 0.1), 11, field_fn, lambda p: p, "rk4")
    for layout in ("halo", "halo3", "packed"):
        f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic", storage=storage,
                                         layout=layout)
        traj = jt.create_tracker("rk4", f, open_bc).track_particles(x0, (0.0, 0.1), 11)
        np.testing.assert_array_equal(traj.positions, pos, err_msg=layout)


@pytest.mark.parametrize("extrapolation", ["constant", "linear", "zero", "nan"])
@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_fast_path_leaves_snapshot_range(gpu, storage, extrapolation):
    """Tracking on the fast path (periodic power-of-two grid, z-halo layout)
    from inside the snapshot times to beyond the last one, so the run switches
    mid-way from 'interp' steps to extrapolated ones (copy / linear / zero /
    NaN blends) and back through the per-lane corner cache: every step
    bitwise equal to the oracle."""
    data, times, g, x0 = _cfg1(3000)
    ref = _rounded(data, storage)
    meta = OF.GridMeta(g, g, g)
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic", storage=storage,
                                     extrapolation=extrapolation)
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    traj = jt.create_tracker("rk4", f, bc).track_particles(x0, (0.9, 1.15), 26)
    field_fn = lambda p, t: OF.structured_series_sample_corners(ref, times, meta, p, t, "periodic",
                                                                extrapolation=extrapolation)
    tt, pos = OT.track(x0, (0.9, 1.15), 26, field_fn,
                       lambda p: OB.per_axis(p, [[0, 0, 0], [1, 1, 1]], ["periodic"] * 3), "rk4")
    np.testing.assert_array_equal(traj.positions, pos)


@pytest.mark.parametrize("dist", ["uniform", "gaussian"])
@pytest.mark.parametrize("direction", ["positive", "negative"])
def test_continuous_inlet_random_bitwise(gpu, dist, direction):
    """'uniform'/'gaussian' inlets draw their replacements from the host
    np.random state: the boundary call and a tracked run (step with an open
    boundary, then jt_inlet_exits / np.random / jt_inlet_place) reproduce the
    oracle under the same seed, step by step."""
    b = [[0, 0, 0], [1, 1, 1]]
    x = np.random.default_rng(9).uniform(-0.3, 1.3, (5000, 3)).astype(F32)
    bc = jt.continuous_inlet_boundary_factory(b, "z", direction, dist)
    np.random.seed(21)
    got = bc(x)
    np.random.seed(21)
    np.testing.assert_array_equal(got, OB.continuous_inlet(x, b, "z", direction, dist))
    data, times, g, x0 = _cfg1(4000)
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic")
    bcx = jt.continuous_inlet_boundary_factory(b, "x", direction, dist)
    np.random.seed(5)
    traj = jt.create_tracker("rk4", f, bcx).track_particles(x0, (0.0, 0.3), 31)
    meta = OF.GridMeta(g, g, g)
    field_fn = lambda p, t: OF.structured_series_sample_corners(data, times, meta, p, t, "periodic")
    np.random.seed(5)
    tt, pos = OT.track(x0, (0.0, 0.3), 31, field_fn,
                       lambda p: OB.continuous_inlet(p, b, "x", direction, dist), "rk4")
    np.testing.assert_array_equal(traj.positions, pos)
    assert (pos[1:, :, 0] == (0.0 if direction == "positive" else 1.0)).any()   # replacements happened


@pytest.mark.parametrize("mode", ["periodic", "clamp"])
def test_pair_records_layout(gpu, mode):
    """jt_grid_pair_records: for each snapshot pair (s, s+1) and node (i, j, k)
    the xyz of V_s(k), V_s(k+1), V_s+1(k), V_s+1(k+1), with k+1 = Nz the halo
    plane (plane 0 for periodic bounds, Nz-1 otherwise); an odd grid."""
    rng = np.random.default_rng(3)
    data = rng.standard_normal((3, 5, 6, 7, 3)).astype(F32)
    gx, gy, gz = (np.linspace(0, 1, n, dtype=F32) for n in (5, 6, 7))
    f = jt.StructuredTimeSeriesField(data, np.array([0, 0.5, 1], F32), gx, gy, gz, bounds_handling=mode,
                                     pair_records=True)
    rec = f._pairs.cpu().numpy()
    nxt = np.concatenate([data[:, :, :, 1:], data[:, :, :, (0 if mode == "periodic" else -1):][:, :, :, :1]], 3)
    want = np.concatenate([data[:-1], nxt[:-1], data[1:], nxt[1:]], axis=-1)
    np.testing.assert_array_equal(rec, want)
    with pytest.raises(ValueError):
        jt.StructuredTimeSeriesField(data, np.array([0, 0.5, 1], F32), gx, gy, gz, storage="bfloat16",
                                     pair_records=True)
    assert jt.StructuredTimeSeriesField(data, np.array([0, 0.5, 1], F32), gx, gy, gz, bounds_handling=mode,
                                        pair_records=False)._desc.pair_data is None


@pytest.mark.parametrize("pairs", [True, False])
@pytest.mark.parametrize("span,n", [((0.0, 0.5), 51), ((0.9, 1.15), 26)])
def test_pair_records_tracking_bitwise(gpu, pairs, span, n):
    """RK4 on the periodic fast path with and without the pair records, over
    snapshot changes (0 .. 0.5) and out past the last snapshot time (0.9 ..
    1.15: steps that blend (s, s) or extrapolate take the per-snapshot
    layout, the others the records): both bitwise equal to the oracle at
    every step, with the cell sort on."""
    data, times, g, x0 = _cfg1(3000)
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic", pair_records=pairs)
    assert (f._desc.pair_data is not None) == pairs
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    tr = jt.create_tracker("rk4", f, bc, sort_every=5, steps_per_launch=5)
    traj = tr.track_particles(x0, span, n)
    meta = OF.GridMeta(g, g, g)
    field_fn = lambda p, t: OF.structured_series_sample_corners(data, times, meta, p, t, "periodic")
    tt, pos = OT.track(x0, span, n, field_fn, lambda p: OB.per_axis(p, [[0, 0, 0], [1, 1, 1]], ["periodic"] * 3),
                       "rk4")
    np.testing.assert_array_equal(traj.positions, pos)
