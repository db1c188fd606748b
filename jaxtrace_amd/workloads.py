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
