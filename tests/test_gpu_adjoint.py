"""GPU adjoint RK4 (jt_track_grid_adjoint behind torch.autograd) against the
float64 autograd restatement (oracle/adjoint_ref.py, itself pinned by FD) and
against float64 central finite differences on sampled entries."""
import numpy as np
import pytest
import torch

import jaxtrace_amd as jt
from jaxtrace_amd import adjoint as ADJ
from jaxtrace_amd import workloads as W
from oracle import adjoint_ref as A
from oracle import fields as OF

pytestmark = pytest.mark.gpu
KINDS = ["periodic", "reflective", "reflective"]


def _problem(n=12, T=4, N=64, steps=30, seed=3):
    data, times, g = W.random_fourier(n=n, n_snap=T, t_end=1.0, n_modes=8, kmax=3, seed=seed)
    rng = np.random.default_rng(seed + 1)
    x0 = (np.floor(rng.uniform(0.05, 0.95, (N, 3)) * n) + rng.uniform(0.2, 0.8, (N, 3))) / n
    target = rng.uniform(0, 1, (N, 3))
    return data, times, g, x0.astype(np.float32), target.astype(np.float32), steps


@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_adjoint_matches_fp64_autograd(gpu, storage):
    data, times, g, x0, target, steps = _problem()
    dev = torch.device("cuda")
    bc = jt.mixed_boundary([[0, 0, 0], [1, 1, 1]], dict(zip("xyz", KINDS)))
    U = data.to(dev).requires_grad_(True)
    X = torch.tensor(x0, device=dev, requires_grad=True)
    xT = ADJ.track_rk4(U, times, g, g, g, X, (0.0, steps * 2e-3), steps + 1, dt=2e-3, boundary=bc,
                       storage=storage, checkpoint_every=8)
    # forward equals the plain tracker bitwise
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic", storage=storage)
    ref_xT = jt.create_tracker("rk4", f, bc, record_final_only=True).track_particles(
        x0, (0.0, steps * 2e-3), steps + 1, dt=2e-3).positions[-1]
    np.testing.assert_array_equal(xT.detach().cpu().numpy(), ref_xT)
    tgt = torch.tensor(target, device=dev)
    loss = ((xT - tgt) ** 2).sum(1).mean()
    loss.backward()
    # float64 autograd on the stored (possibly rounded) snapshot values
    Us = data.to(torch.bfloat16).to(torch.float32) if storage == "bfloat16" else data
    U64 = Us.double().requires_grad_(True)
    X64 = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    meta = OF.GridMeta(g, g, g)
    tg = np.linspace(0.0, steps * 2e-3, steps + 1, dtype=np.float32)
    xT64 = A.track(U64, times, meta, X64, tg[:-1], float(np.float32(2e-3)), KINDS)
    l64 = ((xT64 - torch.tensor(target, dtype=torch.float64)) ** 2).sum(1).mean()
    l64.backward()
    gx, gx64 = X.grad.cpu().numpy(), X64.grad.numpy()
    gu, gu64 = U.grad.cpu().numpy(), U64.grad.numpy()
    np.testing.assert_allclose(gx, gx64, rtol=1e-3, atol=1e-4 * np.abs(gx64).max())
    np.testing.assert_allclose(gu, gu64, rtol=1e-3, atol=1e-4 * np.abs(gu64).max())


def test_adjoint_against_fp64_finite_differences(gpu):
    data, times, g, x0, target, steps = _problem(N=32, steps=20, seed=7)
    dev = torch.device("cuda")
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    U = data.to(dev).requires_grad_(True)
    X = torch.tensor(x0, device=dev, requires_grad=True)
    xT = ADJ.track_rk4(U, times, g, g, g, X, (0.0, steps * 2e-3), steps + 1, dt=2e-3, boundary=bc,
                       checkpoint_every=6)
    ((xT - torch.tensor(target, device=dev)) ** 2).sum(1).mean().backward()
    meta = OF.GridMeta(g, g, g)
    tg = np.linspace(0.0, steps * 2e-3, steps + 1, dtype=np.float32)
    dt = float(np.float32(2e-3))
    U64, x64, t64 = data.double(), torch.tensor(x0, dtype=torch.float64), torch.tensor(target, dtype=torch.float64)

    def L(Uv, xv):
        xT = A.track(Uv, times, meta, xv, tg[:-1], dt, ["periodic"] * 3)
        return float(((xT - t64) ** 2).sum(1).mean())

    rng = np.random.default_rng(9)
    for _ in range(16):
        i, a = int(rng.integers(0, len(x0))), int(rng.integers(0, 3))
        xp, xm = x64.clone(), x64.clone()
        xp[i, a] += 1e-6
        xm[i, a] -= 1e-6
        fd = (L(U64, xp) - L(U64, xm)) / 2e-6
        np.testing.assert_allclose(float(X.grad[i, a]), fd, rtol=2e-3, atol=1e-6)
    nz = torch.nonzero(U.grad.abs().cpu() > 1e-5)
    for r in nz[rng.choice(len(nz), 16, replace=False)]:
        idx = tuple(int(v) for v in r)
        Up, Um = U64.clone(), U64.clone()
        Up[idx] += 1e-4
        Um[idx] -= 1e-4
        fd = (L(Up, x64) - L(Um, x64)) / 2e-4
        np.testing.assert_allclose(float(U.grad[idx]), fd, rtol=2e-3, atol=1e-7)


@pytest.mark.parametrize("storage", ["float32", "bfloat16"])
def test_adjoint_at_scale_against_fp64_finite_differences(gpu, storage):
    """The adjoint workload's field (random Fourier 128^3 x 8 snapshots at
    t_j = j/7, seed 2), 2^16 particles (seed 7), targets (seed 8), RK4
    dt = 2e-3, periodic box, checkpoints every K = 32 steps, 40 steps from
    t = 0.1: two checkpoint segments, a snapshot change at t = 1/7 (step 21)
    inside the first, 128 blocks of particles spread over a grid far larger
    than one block's LDS box (so the reverse kernel flushes to global
    atomics).  32 entries of dL/dx0 and 32 of dL/dU are checked against
    float64 central differences (eps 1e-6 for x0, 1e-4 for U) of the float64
    map on the stored snapshot values.  Entries are drawn (seed 9) among
    particles, and nodes of particles, whose float64 stage points all stay
    1e-3 cells away from cell faces (so the perturbed tracks see the same
    cells); each difference re-tracks only the particles the entry reaches.
    Tolerance: 2e-3 relative plus 1e-4 of the largest sampled entry."""
    dev = torch.device("cuda")
    n_part, steps, K, dt, t0 = 1 << 16, 40, 32, 2e-3, 0.1
    data, times, g = W.random_fourier(n=128, n_snap=8, t_end=1.0, seed=2, device=dev)
    x0 = np.random.default_rng(7).random((n_part, 3)).astype(np.float32)
    target = np.random.default_rng(8).random((n_part, 3)).astype(np.float32)
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    U = data.clone().requires_grad_(True)
    X = torch.tensor(x0, device=dev, requires_grad=True)
    xT = ADJ.track_rk4(U, times, g, g, g, X, (t0, t0 + steps * dt), steps + 1, dt=dt, boundary=bc,
                       storage=storage, checkpoint_every=K)
    ((xT - torch.tensor(target, device=dev)) ** 2).sum(1).mean().backward()
    gx = X.grad.cpu().numpy()
    gU = U.grad
    Us = data.to(torch.bfloat16).to(torch.float32) if storage == "bfloat16" else data
    U64 = Us.double().cpu().numpy()
    del U, data, Us
    meta = OF.GridMeta(g, g, g)
    tg = np.linspace(t0, t0 + steps * dt, steps + 1, dtype=np.float32)
    st, dt32 = tg[:-1], float(tg[1] - tg[0])
    kinds = ["periodic"] * 3
    x64 = x0.astype(np.float64)
    t64 = target.astype(np.float64)
    _, rec = A.track_np(U64, times, meta, x64, st, dt32, kinds, record=True)
    assert {r[1] for r in rec} == {(0, 1), (1, 2)}
    margin = np.min(np.stack([r[3] for r in rec]), axis=0)
    ok = np.nonzero(margin >= 1e-3)[0]
    assert ok.size > 1000

    def loss_of(Uv, parts, xs):
        xT = A.track_np(Uv, times, meta, xs, st, dt32, kinds)
        return float(((xT - t64[parts]) ** 2).sum() / n_part)

    rng = np.random.default_rng(9)
    got, fd = [], []
    for i in rng.choice(ok, 32, replace=False):
        a = int(rng.integers(0, 3))
        xp, xm = x64[[i]].copy(), x64[[i]].copy()
        xp[0, a] += 1e-6
        xm[0, a] -= 1e-6
        got.append(float(gx[i, a]))
        fd.append((loss_of(U64, [i], xp) - loss_of(U64, [i], xm)) / 2e-6)
    got, fd = np.array(got), np.array(fd)
    np.testing.assert_allclose(got, fd, rtol=2e-3, atol=1e-4 * np.abs(fd).max())

    # nodes of the eligible particles' stage cells; the particles an entry
    # reaches are those with a stage point in one of the node's 8 cells
    # whose blend pair weights that snapshot: a sorted (cell, particle) key
    # list per snapshot answers that
    n = 128
    cells = np.stack([r[0] for r in rec])                 # (stages, N, 3)
    pairs = [r[1] for r in rec]
    wts = [r[2] for r in rec]
    lin = (cells[..., 0] * n + cells[..., 1]) * n + cells[..., 2]
    keys = {}
    for snap in {q for pr in pairs for q in pr}:
        use = [s for s in range(len(rec)) if snap in pairs[s] and wts[s][pairs[s].index(snap)] != 0.0]
        keys[snap] = np.unique((lin[use] * n_part + np.arange(n_part)[None, :]).ravel())
    corner = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)])
    got, fd = [], []
    tries = 0
    while len(got) < 32:
        tries += 1
        assert tries < 5000
        i = int(rng.choice(ok))
        s_i = int(rng.integers(0, len(rec)))
        slot = int(rng.integers(0, 2))
        snap, wsnap = pairs[s_i][slot], wts[s_i][slot]
        if wsnap == 0.0:
            continue
        node = (cells[s_i, i] + rng.integers(0, 2, 3)) % n
        adj = (node[None, :] - corner) % n
        cl = (adj[:, 0] * n + adj[:, 1]) * n + adj[:, 2]
        kk = keys[snap]
        lo_, hi_ = np.searchsorted(kk, cl * n_part), np.searchsorted(kk, (cl + 1) * n_part)
        parts = np.unique(np.concatenate([kk[l:h] % n_part for l, h in zip(lo_, hi_)]))
        assert i in parts
        if (margin[parts] < 1e-3).any():
            continue
        c = int(rng.integers(0, 3))
        idx = (snap, int(node[0]), int(node[1]), int(node[2]), c)
        u = U64[idx]
        U64[idx] = u + 1e-4
        lp = loss_of(U64, parts, x64[parts])
        U64[idx] = u - 1e-4
        lm = loss_of(U64, parts, x64[parts])
        U64[idx] = u
        got.append(float(gU[idx]))
        fd.append((lp - lm) / 2e-4)
    got, fd = np.array(got), np.array(fd)
    np.testing.assert_allclose(got, fd, rtol=2e-3, atol=1e-4 * np.abs(fd).max())
