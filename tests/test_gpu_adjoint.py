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
