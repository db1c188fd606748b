"""Synthetic inputs for the benchmark and tests: seeded, generated in float64
and rounded to float32.

* ``abc_flow``            32^3 x 4 time-modulated ABC flow on a period-1 lattice
* ``random_fourier``      divergence-free random Fourier field on an n^3 lattice
* ``uniform_particles``   NumPy ``default_rng(seed).random((n,3))`` in [0,1)^3

Field generation runs in torch float64 on whichever device is given (CPU for
small test grids, the GPU for the 256^3 x 16 workload).
"""
from __future__ import annotations

import math

import numpy as np
import torch

F32 = np.float32


def lattice(n: int) -> np.ndarray:
    """Period-1 lattice i/n without the duplicate endpoint, as f32."""
    return (np.arange(n, dtype=np.float64) / n).astype(F32)


def abc_flow(n: int = 32, times=(0.0, 1.0 / 3.0, 2.0 / 3.0, 1.0),
             A: float = math.sqrt(3.0), B: float = math.sqrt(2.0), C: float = 1.0):
    """u = (A sin2pi z + C cos2pi y, B sin2pi x + A cos2pi z, C sin2pi y + B cos2pi x)
    * (1 + 0.1 sin 2pi t).  Returns (data (T,n,n,n,3) f32, times f32, grid f32)."""
    g64 = np.arange(n, dtype=np.float64) / n
    X, Y, Z = np.meshgrid(g64, g64, g64, indexing="ij")
    tp = 2.0 * np.pi
    u = A * np.sin(tp * Z) + C * np.cos(tp * Y)
    v = B * np.sin(tp * X) + A * np.cos(tp * Z)
    w = C * np.sin(tp * Y) + B * np.cos(tp * X)
    base = np.stack([u, v, w], axis=-1)
    t64 = np.asarray(times, dtype=np.float64)
    data = np.stack([base * (1.0 + 0.1 * np.sin(tp * t)) for t in t64], axis=0).astype(F32)
    return data, t64.astype(F32), g64.astype(F32)


def fourier_modes(n_modes: int = 64, kmax: int = 8, seed: int = 2):
    """Seeded mode table: integer wavevectors with 0 < |k| <= kmax (uniform),
    unit amplitudes a = unit(k x r) (divergence-free), phases and frequencies
    uniform in [0, 2pi)."""
    rng = np.random.default_rng(seed)
    r = np.arange(-kmax, kmax + 1)
    K = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    k2 = (K * K).sum(1)
    cand = K[(k2 > 0) & (k2 <= kmax * kmax)]
    k = cand[rng.integers(0, len(cand), n_modes)].astype(np.float64)
    rr = rng.standard_normal((n_modes, 3))
    a = np.cross(k, rr)
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    phi = rng.uniform(0.0, 2.0 * np.pi, n_modes)
    omega = rng.uniform(0.0, 2.0 * np.pi, n_modes)
    return k, a, phi, omega


def random_fourier(n: int = 256, n_snap: int = 16, t_end: float = 2.0, n_modes: int = 64,
                   kmax: int = 8, seed: int = 2, device="cpu", out_dtype=torch.float32):
    """u(x,t) = sum_m a_m cos(2pi k_m.x + phi_m + omega_m t), rescaled to unit RMS
    speed over all nodes and snapshots.  Snapshots at t_j = t_end*j/(n_snap-1).
    Returns (data (T,n,n,n,3) torch on ``device``, times f32, grid f32)."""
    k, a, phi, omega = fourier_modes(n_modes, kmax, seed)
    t64 = t_end * np.arange(n_snap, dtype=np.float64) / max(n_snap - 1, 1)
    g = torch.arange(n, dtype=torch.float64, device=device) / n
    tp = 2.0 * math.pi
    out = torch.empty((n_snap, n, n, n, 3), dtype=out_dtype, device=device)
    sumsq = 0.0
    raw = []
    for j, t in enumerate(t64):
        u = torch.zeros((n, n, n, 3), dtype=torch.float64, device=device)
        for m in range(n_modes):
            th = (tp * k[m, 0] * g)[:, None, None] + (tp * k[m, 1] * g)[None, :, None] \
                + (tp * k[m, 2] * g)[None, None, :] + (phi[m] + omega[m] * t)
            c = torch.cos(th)
            u += c[..., None] * torch.as_tensor(a[m], dtype=torch.float64, device=device)
        sumsq += float((u * u).sum())
        raw.append(u)
    rms = math.sqrt(sumsq / (n_snap * n * n * n))
    for j, u in enumerate(raw):
        out[j] = (u / rms).to(out_dtype)
    return out, t64.astype(F32), lattice(n)


def uniform_particles(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).random((n, 3)).astype(F32)


def gaussian_mixture(n: int, n_comp: int = 64, seed: int = 6) -> np.ndarray:
    """n points from an isotropic Gaussian mixture in the unit cube: Dirichlet(1)
    weights, means U[0.1,0.9]^3, sigma U[0.01,0.05]; points outside [0,1]^3
    are redrawn from their component."""
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(n_comp))
    mu = rng.uniform(0.1, 0.9, (n_comp, 3))
    sig = rng.uniform(0.01, 0.05, n_comp)
    comp = rng.choice(n_comp, size=n, p=w)
    out = np.empty((n, 3), np.float64)
    todo = np.arange(n)
    while todo.size:
        c = comp[todo]
        x = mu[c] + sig[c, None] * rng.standard_normal((todo.size, 3))
        ok = ((x >= 0) & (x <= 1)).all(1)
        out[todo[ok]] = x[ok]
        todo = todo[~ok]
    return out.astype(F32)


def kde_grid_axes(resolution: int = 256):
    """f32 linspace axes of KDEEstimator._make_grid_3d over bounds (0,1)^3."""
    a = np.linspace(0.0, 1.0, resolution, dtype=np.float32)
    return a, a, a


def kuhn_mesh(n: int = 128, jitter: float = 0.2, seed: int = 4):
    """Unit cube split into n^3 hexes, each into the 6 Kuhn (Freudenthal)
    tetrahedra around its (0,0,0)-(1,1,1) diagonal (a conforming mesh).
    Interior nodes are moved by U(-jitter*h, jitter*h) per coordinate.
    Returns (points ((n+1)^3,3) f32, connectivity (6 n^3, 4) int32)."""
    m = n + 1
    g = np.arange(m, dtype=np.float64) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        I, J, K = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
        interior = ((I > 0) & (I < n) & (J > 0) & (J < n) & (K > 0) & (K < n)).ravel()
        d = rng.uniform(-jitter / n, jitter / n, (int(interior.sum()), 3))
        pts[interior] += d
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = ((i * m + j) * m + k).ravel()
    step = np.array([m * m, m, 1])
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        a = base + step[perm[0]]
        b = a + step[perm[1]]
        tets.append(np.stack([base, a, b, base + step.sum()], 1))
    conn = np.stack(tets, 1).reshape(-1, 4).astype(np.int32)
    return pts.astype(F32), conn


def poiseuille_vortex(points, times):
    """u = (1 - r^2, -0.2 s(t) (z-1/2)/(1/2), 0.2 s(t) (y-1/2)/(1/2)) with
    r = |(y-1/2, z-1/2)|/(1/2), s(t) = 1 + 0.1 sin 2 pi t.  (T, N, 3) f32."""
    p = np.asarray(points, np.float64)
    y, z = (p[:, 1] - 0.5) / 0.5, (p[:, 2] - 0.5) / 0.5
    r2 = y * y + z * z
    out = []
    for t in np.asarray(times, np.float64):
        s = 1.0 + 0.1 * np.sin(2 * np.pi * t)
        out.append(np.stack([1.0 - r2, -0.2 * s * z, 0.2 * s * y], 1))
    return np.stack(out).astype(F32)
