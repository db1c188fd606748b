"""GPU parity of the hash grid, neighbour lists, SPH and KDE against the
float32 oracle on identical seeded inputs.

Bars: sort order, cell tables and neighbour lists (indices AND distances)
bit-exact; SPH and KDE densities within 1e-5 relative (the reference leaves
the 512-term sums' order to XLA)."""
import numpy as np
import pytest
import torch

import jaxtrace_amd as jt
from jaxtrace_amd import _lib as L
from jaxtrace_amd import workloads as W
from oracle import density as OD

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.mark.parametrize("key_mode", ["x32", "x64"])
def test_build_order_and_cells(gpu, key_mode):
    P = W.gaussian_mixture(50_000, seed=1)
    h = 2.0 / 128
    g = jt.HashGridNeighbors(P, h, 64, key_mode)
    o = OD.HashGrid(P, h, 64, key_mode)
    np.testing.assert_array_equal(g.pt_indices_sorted.cpu().numpy(), o.order)
    np.testing.assert_array_equal(g.origin, o.origin)
    cs, ce = g.cell_start.cpu().numpy(), g.cell_end.cpu().numpy()
    cnt = ce - cs
    assert int(cnt.sum()) == len(P) and int((cnt > 0).sum()) == o.unique_ids.size
    np.testing.assert_array_equal(np.sort(cnt[cnt > 0]), np.sort(o.counts))
    np.testing.assert_array_equal(np.sort(cs[cnt > 0]), o.starts)


# (points, dim, cell size, key mode): the counting sort's 1, 2, 3 and 4
# passes of <= 8 key bits, one point, ragged last tiles of 4,096 points.
BUILD_CASES = [
    (1, 3, 2.0 / 128, "x32"),
    (4095, 2, 1.0 / 64, "x32"),          # 6 key bits: one pass
    (4097, 3, 2.0 / 128, "x32"),         # 12 bits: two passes
    (3 * 4096 + 17, 3, 1.0 / 64, "x64"), # 18 bits: three passes
    (300_001, 3, 1.0 / 128, "x64"),      # 21 bits: three passes
    (200_003, 2, 1.0 / 20000, "x64"),    # 29 bits: four passes
    (1 << 20, 3, 2.0 / 256, "x32"),      # the density workload's cell size
]


@pytest.mark.parametrize("n,dim,h,key_mode", BUILD_CASES)
def test_build_counting_sort(gpu, n, dim, h, key_mode):
    P = W.gaussian_mixture(n, seed=11)[:, :dim].copy()
    g = jt.HashGridNeighbors(P, h, 64, key_mode)
    o = OD.HashGrid(P, h, 64, key_mode)
    order = g.pt_indices_sorted.cpu().numpy()
    np.testing.assert_array_equal(order, o.order)
    np.testing.assert_array_equal(g.sorted_positions.cpu().numpy(), P[o.order])
    # dense cell ids of the oracle's present cells: their [start, end) rows
    C = np.floor((P - o.origin) * o.inv).astype(np.int64) - o.Cmin
    dims = [int(v) for v in g.dims]
    if dim == 3:
        dense = (C[:, 0] * dims[1] + C[:, 1]) * dims[2] + C[:, 2] if key_mode == "x64" else C[:, 1] * dims[2] + C[:, 2]
    else:
        dense = C[:, 0] * dims[1] + C[:, 1] if key_mode == "x64" else C[:, 1]
    ids = torch.from_numpy(dense[o.order][o.starts]).to(g.cell_start.device)
    np.testing.assert_array_equal(g.cell_start[ids].cpu().numpy(), o.starts)
    np.testing.assert_array_equal(g.cell_end[ids].cpu().numpy(), o.starts + o.counts)
    cnt = g.cell_end.long() - g.cell_start.long()
    assert int(cnt.sum()) == n and int((cnt != 0).sum()) == o.unique_ids.size


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("n,offset", [(1, 0), (5, 1), (4095, 0), (70_001, 1), (1 << 20, 0), (1 << 20, 1)])
def test_points_bounds(gpu, dim, n, offset):
    """jt_points_bounds on 16-byte aligned arrays (vector loads and a tail)
    and on a view one point into its buffer (the scalar loads)."""
    P = W.gaussian_mixture(n + offset, seed=13)[:, :dim].copy()
    Pd = torch.from_numpy(P).to(gpu)[offset:]
    assert Pd.is_contiguous()
    lib = L.load()
    mm = torch.empty(2 * dim, dtype=torch.float32, device=gpu)
    ws = torch.empty(L.JT_BOUNDS_WS_BYTES, dtype=torch.uint8, device=gpu)
    L.check(lib.jt_points_bounds(L.ptr(Pd), n, dim, L.ptr(mm), L.ptr(ws), L.JT_BOUNDS_WS_BYTES,
                                 L.stream_handle()), "jt_points_bounds")
    got = mm.cpu().numpy()
    np.testing.assert_array_equal(got[:dim], P[offset:].min(0))
    np.testing.assert_array_equal(got[dim:], P[offset:].max(0))


@pytest.mark.parametrize("key_mode", ["x32", "x64"])
@pytest.mark.parametrize("cap,K", [(64, 512), (8, 20), (3, 100)])
def test_query_many_bitwise(gpu, key_mode, cap, K):
    P = W.gaussian_mixture(20_000, seed=2)
    h = 1.0 / 64
    g = jt.HashGridNeighbors(P, h, cap, key_mode)
    o = OD.HashGrid(P, h, cap, key_mode)
    rng = np.random.default_rng(3)
    Q = np.concatenate([rng.uniform(-0.05, 1.05, (300, 3)), P[rng.choice(len(P), 200)]]).astype(F32)
    gi, gd = g.query_many(Q, 2 * h, K)
    oi, od = o.query_many(Q, 2 * h, K)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("key_mode", ["x32", "x64"])
@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland", "gaussian"])
def test_sph_density(gpu, key_mode, kernel):
    P = W.gaussian_mixture(100_000, seed=4)
    h = 2.0 / 256
    rng = np.random.default_rng(5)
    Q = np.concatenate([rng.uniform(0, 1, (400, 3)), P[rng.choice(len(P), 400)]]).astype(F32)
    est = jt.SPHDensityEstimator(P, dimensions="3d", smoothing_length=h, kernel_type=kernel,
                                 per_cell_cap=64, max_neighbors=512, key_mode=key_mode)
    got = est.evaluate(Q)
    want, _, _ = OD.sph_density(P, Q, h, 64, 512, key_mode, True, kernel=kernel)
    floor = 1e-3 * np.abs(want).max()
    m = np.abs(want) >= floor
    np.testing.assert_allclose(got[m], want[m], rtol=1e-5)
    np.testing.assert_allclose(got[~m], want[~m], rtol=1e-5, atol=1e-5 * floor)


@pytest.mark.parametrize("key_mode", ["x32", "x64"])
@pytest.mark.parametrize("K", [100, 128, 512])
def test_sph_truncation_path(gpu, key_mode, K):
    """Dense cluster: more than max_neighbors valid candidates, so the
    wave takes the stable first-K selection (x32: K not a multiple of the
    stencil's 3 copies of each cell, so the cut falls inside a copy group)."""
    rng = np.random.default_rng(6)
    P = (0.5 + 0.002 * rng.standard_normal((20_000, 3))).astype(F32)
    h = 0.004
    Q = (0.5 + 0.003 * rng.standard_normal((256, 3))).astype(F32)
    est = jt.SPHDensityEstimator(P, dimensions="3d", smoothing_length=h, key_mode=key_mode, max_neighbors=K,
                                 normalize=False)
    want, idx, _ = OD.sph_density(P, Q, h, 64, K, key_mode, False)
    np.testing.assert_allclose(est.evaluate(Q), want, rtol=1e-5)


def test_sph_truncation_per_point_masses(gpu):
    """Per-point masses keep the sorted first-K selection (which tied
    candidates are kept matters then)."""
    rng = np.random.default_rng(16)
    P = (0.5 + 0.002 * rng.standard_normal((20_000, 3))).astype(F32)
    masses = rng.uniform(0.5, 2.0, len(P)).astype(F32)
    h = 0.004
    Q = (0.5 + 0.003 * rng.standard_normal((128, 3))).astype(F32)
    for key_mode in ("x32", "x64"):
        est = jt.SPHDensityEstimator(P, dimensions="3d", smoothing_length=h, key_mode=key_mode, max_neighbors=100,
                                     normalize=False, masses=masses)
        want, _, _ = OD.sph_density(P, Q, h, 64, 100, key_mode, False, masses=masses)
        np.testing.assert_allclose(est.evaluate(Q), want, rtol=1e-5)


@pytest.mark.parametrize("key_mode", ["x32", "x64"])
def test_sph_truncation_ties(gpu, key_mode):
    """The K-th nearest inside a group of equal distances.  Points on a
    lattice of spacing 2^-9 around lattice-point queries have bit-identical
    float32 distances per shell (|k|^2 = 9 holds 30 points, and the 100th
    nearest of the centre query is the 7th of them), so the first-K cut
    (stable order, neighbors.py:153-161) splits a tie group."""
    d = F32(2.0 ** -9)
    k = np.arange(-5, 6, dtype=F32)
    X, Y, Z = np.meshgrid(k, k, k, indexing="ij")
    P = (F32(0.5) + d * np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)).astype(F32)
    Q = (F32(0.5) + d * np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, -1, 2], [2, 2, 2]], F32)).astype(F32)
    h = float(4 * d)
    for K in (100, 57, 123):
        est = jt.SPHDensityEstimator(P, dimensions="3d", smoothing_length=h, key_mode=key_mode, max_neighbors=K,
                                     normalize=False)
        want, _, _ = OD.sph_density(P, Q, h, 64, K, key_mode, False)
        np.testing.assert_allclose(est.evaluate(Q), want, rtol=1e-5)


def test_kde_exhaustive_and_cutoff(gpu):
    P = W.gaussian_mixture(1 << 14, seed=7)
    kde = jt.KDEEstimator(P, resolution=24, bounds=((0, 1),) * 3)
    h = kde.h
    np.testing.assert_allclose(h, OD.scott_bandwidth(P), rtol=0)
    X, Y, Z = OD.make_grid_3d(P, 24, ((0, 1),) * 3)
    G = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    want = OD.kde_eval(G, P, h, dtype=np.float64)
    got = kde.evaluate(G)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * want.max())
    cut = jt.KDEEstimator(P, resolution=24, bounds=((0, 1),) * 3, cutoff_sigmas=7.0)
    d_cut = cut.density_grid_3d().cpu().numpy().ravel()
    floor = 1e-3 * want.max()
    m = want >= floor
    np.testing.assert_allclose(d_cut[m], want[m], rtol=1e-5)
    # normalised evaluate_3d follows kde.py:229-239 on the same densities
    Xg, Yg, Zg, D = kde.evaluate_3d()
    np.testing.assert_allclose(D, OD.kde_normalize_3d(Xg, Yg, Zg, want.reshape(Xg.shape).astype(F32), len(P)),
                               rtol=2e-5, atol=1e-6)


def test_kde_cutoff_tiles_sampled(gpu):
    """Cut-off KDE on a 40^3 grid (16 x 16 x 32 tiles, the last ones 8 points
    deep in x, y and z) over 2^16 mixture particles: 4096 sampled grid points
    against the float64 exhaustive sum (kde.py:243-253) at the 1e-5 bar."""
    P = W.gaussian_mixture(1 << 16, seed=11)
    cut = jt.KDEEstimator(P, resolution=40, bounds=((0, 1),) * 3, cutoff_sigmas=7.0)
    d = cut.density_grid_3d().cpu().numpy().ravel()
    X, Y, Z = OD.make_grid_3d(P, 40, ((0, 1),) * 3)
    G = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    idx = np.random.default_rng(12).choice(len(G), 4096, replace=False)
    want = OD.kde_eval(G[idx], P, cut.h, dtype=np.float64)
    m = want >= 1e-3 * want.max()
    assert m.sum() > 1000
    np.testing.assert_allclose(d[idx][m], want[m], rtol=1e-5)


@pytest.mark.parametrize("shape", [(37, 21, 70), (5, 3, 2), (64, 64, 64)])
def test_kde_cutoff_workspace_path(gpu, shape):
    """jt_kde_grid_cutoff_ws (per-block slices summed in a fixed order) on
    grids with ragged tiles in every axis: the same bits on two calls, within
    2e-6 of the float-atomics entry point, and within 1e-5 of the float64
    exhaustive sum (kde.py:243-253) at 2048 sampled points above the floor."""
    import ctypes as C
    lib = L.load()
    P = W.gaussian_mixture(1 << 15, seed=40)
    Pd = torch.from_numpy(P).cuda()
    h = float(OD.scott_bandwidth(P))
    R = F32(7.0 * h)
    hg = jt.HashGridNeighbors(Pd, float(R) / 2, 1, "x64", "cuda")
    nx, ny, nz = shape
    axes = [np.linspace(0.05, 0.95, k, dtype=F32) for k in shape]
    gx, gy, gz = (torch.from_numpy(a).cuda() for a in axes)
    ws_bytes = int(lib.jt_kde_grid_cutoff_workspace_size(nx, ny, nz))
    assert ws_bytes >= nx * ny * nz * 4
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    outs = []
    for _ in range(2):
        out = torch.full((ny, nx, nz), -1.0, dtype=torch.float32, device="cuda")
        L.check(lib.jt_kde_grid_cutoff_ws(C.byref(hg._desc), len(P), L.ptr(gx), nx, L.ptr(gy), ny, L.ptr(gz), nz,
                                          C.c_float(F32(h)), C.c_float(R), L.ptr(out), L.ptr(ws), ws_bytes,
                                          L.stream_handle()), "jt_kde_grid_cutoff_ws")
        outs.append(out.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    at = torch.empty((ny, nx, nz), dtype=torch.float32, device="cuda")
    L.check(lib.jt_kde_grid_cutoff(C.byref(hg._desc), len(P), L.ptr(gx), nx, L.ptr(gy), ny, L.ptr(gz), nz,
                                   C.c_float(F32(h)), C.c_float(R), L.ptr(at), L.stream_handle()),
            "jt_kde_grid_cutoff")
    at = at.cpu().numpy()
    np.testing.assert_allclose(outs[0], at, rtol=2e-6, atol=1e-7 * np.abs(at).max())
    with pytest.raises(MemoryError, match="workspace too small"):
        L.check(lib.jt_kde_grid_cutoff_ws(C.byref(hg._desc), len(P), L.ptr(gx), nx, L.ptr(gy), ny, L.ptr(gz), nz,
                                          C.c_float(F32(h)), C.c_float(R), L.ptr(out), L.ptr(ws), ws_bytes - 512,
                                          L.stream_handle()), "jt_kde_grid_cutoff_ws")
    X, Y, Z = np.meshgrid(*axes, indexing="xy")
    G = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    idx = np.random.default_rng(41).choice(len(G), min(2048, len(G)), replace=False)
    want = OD.kde_eval(G[idx], P, F32(h), dtype=np.float64)
    m = want >= 1e-3 * want.max()
    np.testing.assert_allclose(outs[0].ravel()[idx][m], want[m], rtol=1e-5)


@pytest.mark.parametrize("n,shape", [(1, (5, 3)), (63, (37, 21)), (4097, (130, 70)), (70_001, (128, 64)),
                                     (200_003, (300, 129))])
def test_kde_grid2d_all_pairs(gpu, n, shape):
    """jt_kde_grid2d (every pair, on the matrix cores) on grids with ragged
    tiles and particle counts off the 64-particle passes: within 1e-5 of the
    float64 all-pairs sum (kde.py:244-270) above 1e-3 of the maximum, the
    same bits on two calls, within 1e-5 of the one-thread-per-point kernel,
    and MemoryError on a short workspace."""
    import ctypes as C
    lib = L.load()
    rng = np.random.default_rng(50 + n)
    P = (0.5 + 0.15 * rng.standard_normal((n, 2))).astype(F32)
    h = float(OD.scott_bandwidth(P)) if n > 1 else 0.05
    nx, ny = shape
    x = np.linspace(0.0, 1.0, nx, dtype=F32)
    y = np.linspace(0.1, 0.9, ny, dtype=F32)
    Pd, gx, gy = (torch.from_numpy(a).cuda() for a in (P, x, y))
    ws_bytes = int(lib.jt_kde_grid2d_workspace_size(n, nx, ny))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    outs = []
    for _ in range(2):
        out = torch.full((ny, nx), -1.0, dtype=torch.float32, device="cuda")
        L.check(lib.jt_kde_grid2d(L.ptr(Pd), n, L.ptr(gx), nx, L.ptr(gy), ny, C.c_float(F32(h)), L.ptr(out),
                                  L.ptr(ws), ws_bytes, L.stream_handle()), "jt_kde_grid2d")
        outs.append(out.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    X, Y = np.meshgrid(x, y, indexing="xy")
    G = np.stack([X.ravel(), Y.ravel()], 1)
    idx = np.random.default_rng(51).choice(len(G), min(1500, len(G)), replace=False)
    want = OD.kde_eval(G[idx], P, F32(h), dtype=np.float64)
    got = outs[0].ravel()[idx]
    m = want >= 1e-3 * want.max()
    np.testing.assert_allclose(got[m], want[m], rtol=1e-5)
    np.testing.assert_allclose(got[~m], want[~m], rtol=0, atol=1e-5 * 1e-3 * want.max())
    ref = torch.empty(nx * ny, dtype=torch.float32, device="cuda")
    Gd = torch.from_numpy(np.ascontiguousarray(G.astype(F32))).cuda()
    L.check(lib.jt_kde_eval(L.ptr(Pd), n, 2, L.ptr(Gd), nx * ny, C.c_float(F32(h)), L.ptr(ref), L.stream_handle()),
            "jt_kde_eval")
    ref = ref.cpu().numpy().reshape(ny, nx)
    np.testing.assert_allclose(outs[0], ref, rtol=1e-5, atol=1e-8 * ref.max())
    with pytest.raises(MemoryError, match="workspace too small"):
        L.check(lib.jt_kde_grid2d(L.ptr(Pd), n, L.ptr(gx), nx, L.ptr(gy), ny, C.c_float(F32(h)), L.ptr(out),
                                  L.ptr(ws), ws_bytes - 512, L.stream_handle()), "jt_kde_grid2d")


@pytest.mark.parametrize("n,shape", [(1, (3, 5, 2)), (255, (17, 33, 40)), (70_001, (40, 40, 40)),
                                     (300_007, (33, 20, 70))])
def test_kde_grid3d_all_pairs(gpu, n, shape):
    """jt_kde_grid3d (every pair, the tile kernel over chunks of the points)
    on grids with ragged tiles: within 1e-5 of the float64 all-pairs sum
    (kde.py:244-270) above 1e-3 of the maximum, the same bits on two calls,
    within 1e-5 of the one-thread-per-point kernel, MemoryError on a short
    workspace."""
    import ctypes as C
    lib = L.load()
    rng = np.random.default_rng(60 + n)
    P = (0.5 + 0.15 * rng.standard_normal((n, 3))).astype(F32)
    h = float(OD.scott_bandwidth(P)) if n > 1 else 0.05
    nx, ny, nz = shape
    axes = [np.linspace(0.0, 1.0, nx, dtype=F32), np.linspace(0.1, 0.9, ny, dtype=F32),
            np.linspace(0.2, 0.8, nz, dtype=F32)]
    Pd = torch.from_numpy(P).cuda()
    gx, gy, gz = (torch.from_numpy(a).cuda() for a in axes)
    ws_bytes = int(lib.jt_kde_grid3d_workspace_size(n, nx, ny, nz))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    outs = []
    for _ in range(2):
        out = torch.full((ny, nx, nz), -1.0, dtype=torch.float32, device="cuda")
        L.check(lib.jt_kde_grid3d(L.ptr(Pd), n, L.ptr(gx), nx, L.ptr(gy), ny, L.ptr(gz), nz, C.c_float(F32(h)),
                                  L.ptr(out), L.ptr(ws), ws_bytes, L.stream_handle()), "jt_kde_grid3d")
        outs.append(out.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    X, Y, Z = np.meshgrid(*axes, indexing="xy")
    G = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    idx = np.random.default_rng(61).choice(len(G), min(1500, len(G)), replace=False)
    want = OD.kde_eval(G[idx], P, F32(h), dtype=np.float64)
    got = outs[0].ravel()[idx]
    m = want >= 1e-3 * want.max()
    np.testing.assert_allclose(got[m], want[m], rtol=1e-5)
    np.testing.assert_allclose(got[~m], want[~m], rtol=0, atol=1e-5 * 1e-3 * want.max())
    ref = torch.empty(len(G), dtype=torch.float32, device="cuda")
    Gd = torch.from_numpy(np.ascontiguousarray(G.astype(F32))).cuda()
    L.check(lib.jt_kde_eval(L.ptr(Pd), n, 3, L.ptr(Gd), len(G), C.c_float(F32(h)), L.ptr(ref), L.stream_handle()),
            "jt_kde_eval")
    ref = ref.cpu().numpy().reshape(ny, nx, nz)
    np.testing.assert_allclose(outs[0], ref, rtol=1e-5, atol=1e-8 * ref.max())
    with pytest.raises(MemoryError, match="workspace too small"):
        L.check(lib.jt_kde_grid3d(L.ptr(Pd), n, L.ptr(gx), nx, L.ptr(gy), ny, L.ptr(gz), nz, C.c_float(F32(h)),
                                  L.ptr(out), L.ptr(ws), ws_bytes - 512, L.stream_handle()), "jt_kde_grid3d")


@pytest.mark.parametrize("n,m,dim", [(1, 1, 3), (1000, 7, 2), (200_003, 2500, 3), (1 << 20, 300, 2),
                                     (70_001, 70_000, 3)])
def test_kde_eval_chunked(gpu, n, m, dim):
    """jt_kde_eval_ws (the particles split over blocks when the queries are
    few, float64 chunk sums added in chunk order) against jt_kde_eval (one
    float64 sum per query) to 2 float32 ulps, and against the float64 sum of
    the oracle (kde.py:244-270) to 1e-5 above 1e-3 of the maximum; the same
    bits on two calls; MemoryError on a short workspace."""
    import ctypes as C
    lib = L.load()
    rng = np.random.default_rng(70 + n)
    P = (0.5 + 0.15 * rng.standard_normal((n, dim))).astype(F32)
    Q = rng.uniform(0.0, 1.0, (m, dim)).astype(F32)
    h = F32(0.03)
    Pd, Qd = torch.from_numpy(P).cuda(), torch.from_numpy(Q).cuda()
    ws_bytes = int(lib.jt_kde_eval_workspace_size(n, m))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    outs = []
    for _ in range(2):
        out = torch.full((m,), -1.0, dtype=torch.float32, device="cuda")
        L.check(lib.jt_kde_eval_ws(L.ptr(Pd), n, dim, L.ptr(Qd), m, C.c_float(h), L.ptr(out), L.ptr(ws), ws_bytes,
                                   L.stream_handle()), "jt_kde_eval_ws")
        outs.append(out.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    ref = torch.empty(m, dtype=torch.float32, device="cuda")
    L.check(lib.jt_kde_eval(L.ptr(Pd), n, dim, L.ptr(Qd), m, C.c_float(h), L.ptr(ref), L.stream_handle()),
            "jt_kde_eval")
    ref = ref.cpu().numpy()
    np.testing.assert_allclose(outs[0], ref, rtol=2.5e-7, atol=0)
    sel = np.random.default_rng(71).choice(m, min(m, 400), replace=False)
    want = OD.kde_eval(Q[sel], P, h, dtype=np.float64)
    got = outs[0][sel]
    mk = want >= max(1e-3 * want.max(), 1e-30)           # float32 holds the value
    np.testing.assert_allclose(got[mk], want[mk], rtol=1e-5)
    if ws_bytes > 0:
        with pytest.raises(MemoryError, match="workspace too small"):
            L.check(lib.jt_kde_eval_ws(L.ptr(Pd), n, dim, L.ptr(Qd), m, C.c_float(h), L.ptr(out), L.ptr(ws),
                                       ws_bytes // 4, L.stream_handle()), "jt_kde_eval_ws")


def test_full_size_kde_grid3d(gpu):
    """evaluate_3d's default (no cut-off: every pair) at the reference's
    default resolution 100 over the 2^24 mixture particles; 16 sampled grid
    points against the float64 sum."""
    P = W.gaussian_mixture(1 << 24, seed=6)
    kde = jt.KDEEstimator(torch.from_numpy(P).cuda(), resolution=100, bounds=((0, 1),) * 3, normalize=False)
    D = kde.density_grid_3d().cpu().numpy()
    x, y, z = kde._axes_3d()
    rng = np.random.default_rng(62)
    iy, ix, iz = (rng.integers(0, 100, 16) for _ in range(3))
    G = np.stack([x[ix], y[iy], z[iz]], 1)
    want = np.concatenate([OD.kde_eval(G, P[k:k + (1 << 21)], F32(kde.h), dtype=np.float64) * (1 << 21)
                           for k in range(0, 1 << 24, 1 << 21)]).reshape(8, 16).sum(0) / (1 << 24)
    m = want >= 1e-3 * want.max()
    assert m.sum() >= 8
    np.testing.assert_allclose(D[iy, ix, iz][m], want[m], rtol=1e-5)


def test_full_size_kde_grid2d(gpu):
    """The 2-D benchmark workload: 2^24 mixture particles' (x, y) on a 256^2
    grid, every pair; 24 sampled grid points against the float64 sum."""
    P = np.ascontiguousarray(W.gaussian_mixture(1 << 24, seed=6)[:, :2])
    kde = jt.KDEEstimator(torch.from_numpy(P).cuda(), resolution=256, bounds=((0, 1), (0, 1)), normalize=False)
    x = np.linspace(0, 1, 256, dtype=F32)
    D = kde.density_grid_2d(x, x, P).cpu().numpy()
    rng = np.random.default_rng(52)
    iy, ix = rng.integers(0, 256, 24), rng.integers(0, 256, 24)
    G = np.stack([x[ix], x[iy]], 1)
    want = np.concatenate([OD.kde_eval(G, P[k:k + (1 << 21)], F32(kde.h), dtype=np.float64) * (1 << 21)
                           for k in range(0, 1 << 24, 1 << 21)]).reshape(8, 24).sum(0) / (1 << 24)
    m = want >= 1e-3 * want.max()
    assert m.sum() >= 12
    np.testing.assert_allclose(D[iy, ix][m], want[m], rtol=1e-5)


@pytest.mark.parametrize("dims", ["3d", "2d"])
def test_sph_query_tiles_match_per_query_path(gpu, dims):
    """x32 keys, uniform masses: the column-tiled evaluation
    (jt_sph_eval_tiled) and one wave per query (jt_sph_eval) sum the same
    terms.  Queries: a block of grid points (many per column), random points
    in and around the particle bounds (columns probed partly or not at all),
    and a dense cluster (more than max_neighbors valid candidates: the top-K
    pass).  Both paths against the oracle too."""
    rng = np.random.default_rng(31)
    P = np.concatenate([W.gaussian_mixture(150_000, seed=30),
                        (0.5 + 0.002 * rng.standard_normal((20_000, 3)))]).astype(F32)
    h = 2.0 / 256
    a = np.linspace(0.3, 0.6, 40, dtype=F32)
    X, Y, Z = np.meshgrid(a, a, a, indexing="xy")
    Q = np.concatenate([np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1), rng.uniform(-0.1, 1.1, (3000, 3)),
                        (0.5 + 0.003 * rng.standard_normal((500, 3)))]).astype(F32)
    kw = dict(plane="xy", position=0.5, slab_thickness=0.1) if dims == "2d" else {}
    res = {}
    for tiled in (True, False):
        est = jt.SPHDensityEstimator(P, dimensions=dims, smoothing_length=h, key_mode="x32", max_neighbors=300,
                                     normalize=False, tiled_queries=tiled, **kw)
        res[tiled] = est.evaluate(Q)
    np.testing.assert_allclose(res[True], res[False], rtol=2e-6, atol=1e-6 * np.abs(res[False]).max())
    sel = np.concatenate([np.arange(0, len(a) ** 3, 97), np.arange(len(a) ** 3, len(Q), 7)])
    want, _, _ = OD.sph_density(P, Q[sel], h, 64, 300, "x32", False, dimensions=dims, **kw)
    floor = 1e-3 * np.abs(want).max()
    for tiled in (True, False):
        got = res[tiled][sel]
        m = np.abs(want) >= floor
        np.testing.assert_allclose(got[m], want[m], rtol=1e-5)
        np.testing.assert_allclose(got[~m], want[~m], rtol=1e-5, atol=1e-5 * floor)


def test_full_size_sph_tiles(gpu):
    """The benchmark's SPH workload (2^24 mixture particles, h = 2/256,
    x32, 512 neighbours) on a 64 x 64 x 64 block of its 256^3 query grid:
    the tiled and the per-query paths agree, and 256 sampled queries match
    the oracle."""
    P = W.gaussian_mixture(1 << 24, seed=6)
    h = 2.0 / 256
    a = np.linspace(0.0, 1.0, 256, dtype=F32)[96:160]
    X, Y, Z = np.meshgrid(a, a, a, indexing="xy")
    Q = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(F32)
    Pd = torch.from_numpy(P).cuda()
    res = {}
    for tiled in (True, False):
        est = jt.SPHDensityEstimator(Pd, dimensions="3d", smoothing_length=h, key_mode="x32", normalize=False,
                                     tiled_queries=tiled)
        res[tiled] = est.evaluate(Q)
        del est
    np.testing.assert_allclose(res[True], res[False], rtol=2e-6, atol=1e-6 * np.abs(res[False]).max())
    sel = np.random.default_rng(32).choice(len(Q), 256, replace=False)
    want, _, _ = OD.sph_density(P, Q[sel], h, 64, 512, "x32", False)
    m = np.abs(want) >= 1e-3 * np.abs(want).max()
    np.testing.assert_allclose(res[True][sel][m], want[m], rtol=1e-5)


def test_full_size_sph_sampled_queries(gpu):
    """2^24 Gaussian-mixture particles, h = 2/256: neighbour lists of 2048
    sampled grid queries bit-exact with the oracle in both key modes."""
    P = W.gaussian_mixture(1 << 24, seed=6)
    h = 2.0 / 256
    a = np.linspace(0.0, 1.0, 256, dtype=F32)
    rng = np.random.default_rng(8)
    ii = rng.integers(0, 256, (2048, 3))
    Q = np.stack([a[ii[:, 0]], a[ii[:, 1]], a[ii[:, 2]]], 1).astype(F32)
    for mode in ("x32", "x64"):
        g = jt.HashGridNeighbors(P, h, 64, mode)
        o = OD.HashGrid(P, h, 64, mode)
        gi, gd = g.query_many(Q[:512], 2 * h, 512)
        oi, od = o.query_many(Q[:512], 2 * h, 512)
        np.testing.assert_array_equal(gi, oi)
        np.testing.assert_array_equal(gd, od)
        del g


# ---- 2-D mode: the reference's default SPH dimensionality (sph.py:90, 112-117),
# 2-D hash keys (neighbors.py:78-81) and KDEEstimator.evaluate_2d (kde.py:205-221)

@pytest.mark.parametrize("key_mode", ["x32", "x64"])
def test_hashgrid_2d_build_and_queries_bitwise(gpu, key_mode):
    P = np.ascontiguousarray(W.gaussian_mixture(30_000, seed=21)[:, :2])
    h = 1.0 / 64
    g = jt.HashGridNeighbors(P, h, 16, key_mode)
    o = OD.HashGrid(P, h, 16, key_mode)
    np.testing.assert_array_equal(g.pt_indices_sorted.cpu().numpy(), o.order)
    rng = np.random.default_rng(22)
    Q = np.concatenate([rng.uniform(-0.05, 1.05, (200, 2)), P[rng.choice(len(P), 200)]]).astype(F32)
    for K in (64, 300):
        gi, gd = g.query_many(Q, 2 * h, K)
        oi, od = o.query_many(Q, 2 * h, K)
        np.testing.assert_array_equal(gd, od)
        np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("key_mode", ["x32", "x64"])
@pytest.mark.parametrize("kernel", ["cubic_spline", "wendland", "gaussian"])
def test_sph_density_2d(gpu, key_mode, kernel):
    """dimensions='2d' on (N,3) positions: the slab |y - 0.5| <= 0.05 in the
    'xz' plane, queries given with three columns (projected), area
    normalisation.  In x32 mode the 2-D key keeps only the second axis, so
    cells are whole columns and the 64-point cap and the 512-neighbour cut
    both bite."""
    P3 = W.gaussian_mixture(200_000, seed=23)
    h = 2.0 / 256
    rng = np.random.default_rng(24)
    Q = np.concatenate([rng.uniform(0, 1, (300, 3)), P3[rng.choice(len(P3), 300)]]).astype(F32)
    kw = dict(plane="xz", position=0.5, slab_thickness=0.1)
    est = jt.SPHDensityEstimator(P3, dimensions="2d", smoothing_length=h, kernel_type=kernel, key_mode=key_mode,
                                 **kw)
    got = est.evaluate(Q)
    want, _, _ = OD.sph_density(P3, Q, h, 64, 512, key_mode, True, kernel=kernel, dimensions="2d", **kw)
    floor = 1e-3 * np.abs(want).max()
    m = np.abs(want) >= floor
    assert m.sum() > 100                             # Wendland's support is h: fewer queries reach the floor
    np.testing.assert_allclose(got[m], want[m], rtol=1e-5)
    np.testing.assert_allclose(got[~m], want[~m], rtol=1e-5, atol=1e-5 * floor)


def test_sph_density_2d_reference_defaults(gpu):
    """The reference's default estimator (dimensions='2d', plane 'xy',
    position 0, slab 0.1, cubic spline), given a smoothing length, on
    (M,2) queries."""
    P3 = W.gaussian_mixture(100_000, seed=25)
    P3[:, 2] *= 0.1                                  # put points in the default slab |z| <= 0.05
    h = 1.0 / 128
    Q = np.random.default_rng(26).uniform(0, 1, (500, 2)).astype(F32)
    got = jt.SPHDensityEstimator(P3, smoothing_length=h).evaluate(Q)
    want, _, _ = OD.sph_density(P3, Q, h, dimensions="2d")
    m = np.abs(want) >= 1e-3 * np.abs(want).max()
    np.testing.assert_allclose(got[m], want[m], rtol=1e-5)


def test_kde_evaluate_2d_and_projections(gpu):
    """KDEEstimator.evaluate_2d on 2-D points and on the 'yz' slab of 3-D
    points (kde.py:112-124, 173-186, 205-221), and evaluate() with 2-D
    queries on a 3-D estimator (kde.py:160-166), against the float64
    exhaustive sum and the oracle's normalisation (pinned by its integral)."""
    P3 = W.gaussian_mixture(1 << 14, seed=27)
    P2 = np.ascontiguousarray(P3[:, :2])
    for pts_in, kw in ((P2, {}), (P3, dict(plane="yz", slab_position=0.5, slab_thickness=0.2))):
        kde = jt.KDEEstimator(pts_in, resolution=32, **kw)
        np.testing.assert_allclose(kde.h, OD.scott_bandwidth(pts_in), rtol=0)
        X, Y, Z = kde.evaluate_2d()
        pts = pts_in if pts_in.shape[1] == 2 else OD.extract_2d(pts_in, kw["plane"], kw["slab_position"],
                                                                  kw["slab_thickness"])
        Xo, Yo = OD.make_grid_2d(pts, 32)
        np.testing.assert_array_equal(X, Xo)
        np.testing.assert_array_equal(Y, Yo)
        G = np.stack([Xo.ravel(), Yo.ravel()], 1)
        dens = OD.kde_eval(G, pts, kde.h, dtype=np.float64).reshape(Xo.shape)
        want = OD.kde_normalize_2d(Xo, Yo, dens, len(pts_in))
        m = want >= 1e-3 * want.max()
        np.testing.assert_allclose(Z[m], want[m], rtol=2e-5)
    kde3 = jt.KDEEstimator(P3, plane="xy", slab_position=0.4, slab_thickness=0.1)
    Q2 = np.random.default_rng(28).uniform(0, 1, (400, 2)).astype(F32)
    want = OD.kde_eval(Q2, OD.extract_2d(P3, "xy", 0.4, 0.1), kde3.h, dtype=np.float64)
    got = kde3.evaluate(Q2)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5 * want.max())
