"""Pins the float64 autograd restatement used to check the adjoint kernel
(oracle/adjoint_ref.py) with float64 central finite differences on sampled
entries of dL/dx0 and dL/dU (cell-face-avoiding positions)."""
import numpy as np
import torch

from jaxtrace_amd import workloads as W
from oracle import adjoint_ref as A
from oracle import fields as OF


def _problem(seed=0, n=12, T=4, N=64, steps=6):
    data, times, g = W.random_fourier(n=n, n_snap=T, t_end=1.0, n_modes=8, kmax=3, seed=seed)
    meta = OF.GridMeta(g, g, g)
    rng = np.random.default_rng(seed + 1)
    x0 = rng.uniform(0.05, 0.95, (N, 3))
    # keep starting points away from cell faces
    x0 = (np.floor(x0 * n) + rng.uniform(0.2, 0.8, x0.shape)) / n
    target = rng.uniform(0, 1, (N, 3))
    tg = np.linspace(0.0, 0.1, steps + 1, dtype=np.float32)
    return data.double(), times, meta, torch.tensor(x0), torch.tensor(target), tg[:-1], float(tg[1] - tg[0])


def _loss(U, times, meta, x0, target, st, dt, kinds):
    xT = A.track(U, times, meta, x0, st, dt, kinds)
    return ((xT - target) ** 2).sum(1).mean()


def test_fd_pins_autograd_restatement():
    U, times, meta, x0, target, st, dt = _problem()
    kinds = ["periodic", "reflective", "reflective"]
    Ur = U.clone().requires_grad_(True)
    xr = x0.clone().requires_grad_(True)
    _loss(Ur, times, meta, xr, target, st, dt, kinds).backward()
    rng = np.random.default_rng(5)
    for _ in range(8):
        i, a = rng.integers(0, x0.shape[0]), rng.integers(0, 3)
        e = 1e-6
        xp, xm = x0.clone(), x0.clone()
        xp[i, a] += e
        xm[i, a] -= e
        fd = (_loss(U, times, meta, xp, target, st, dt, kinds) - _loss(U, times, meta, xm, target, st, dt, kinds)) / (2 * e)
        np.testing.assert_allclose(float(xr.grad[i, a]), float(fd), rtol=1e-5, atol=1e-9)
    nz = torch.nonzero(Ur.grad.abs() > 1e-6)
    for r in nz[rng.choice(len(nz), 8, replace=False)]:
        idx = tuple(int(v) for v in r)
        e = 1e-4
        Up, Um = U.clone(), U.clone()
        Up[idx] += e
        Um[idx] -= e
        fd = (_loss(Up, times, meta, x0, target, st, dt, kinds) - _loss(Um, times, meta, x0, target, st, dt, kinds)) / (2 * e)
        np.testing.assert_allclose(float(Ur.grad[idx]), float(fd), rtol=1e-4, atol=1e-10)


def test_corner_tracker_equals_autograd_restatement():
    """track_np (float64 NumPy, corners only) is the same map as the torch
    restatement: equal to rounding over a run that crosses a snapshot change
    (t = 1/3) with mixed walls, and its recorded stage cells are the floor of
    the stage points."""
    data, times, g = W.random_fourier(n=10, n_snap=4, t_end=1.0, n_modes=8, kmax=3, seed=4)
    meta = OF.GridMeta(g, g, g)
    rng = np.random.default_rng(6)
    x0 = rng.uniform(0, 1, (200, 3))
    kinds = ["periodic", "reflective", "periodic"]
    tg = np.linspace(0.3, 0.38, 41, dtype=np.float32)
    dt = float(tg[1] - tg[0])
    want = A.track(data.double(), times, meta, torch.tensor(x0), tg[:-1], dt, kinds).numpy()
    got, rec = A.track_np(data.double().numpy(), times, meta, x0, tg[:-1], dt, kinds, record=True)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    assert len(rec) == 4 * 40
    assert {r[1] for r in rec} == {(0, 1), (1, 2)}
    cells0, _, _, margin0 = rec[0]
    np.testing.assert_array_equal(cells0, np.floor(x0 * 10).astype(np.int64) % 10)
    assert (margin0 >= 0).all() and (margin0 <= 0.5).all()
