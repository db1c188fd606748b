"""Float64 differentiable restatement of structured RK4 tracking (torch autograd).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The same algorithm as oracle/fields.py + oracle/integrators.py (temporal
blend time_series.py:155-245 in 'interp'/'copy' form, trilinear
structured.py:167-250 with periodic or clamp bounds, rk4.py:84-106, per-axis
periodic/reflective/clamp boundaries boundary.py:101-172), written with torch
float64 ops so torch.autograd gives the exact gradient of the piecewise-smooth
map.  Its own derivative is pinned in the tests by float64 central finite
differences on sampled entries.
"""
from __future__ import annotations

import numpy as np
import torch

from .fields import blend_params


def _axis(x, o, d, n, periodic):
    f = (x - o) / d
    if periodic:
        f = torch.remainder(f, n)
    else:
        f = torch.clamp(f, 0.0, n - 1)
    i0 = torch.floor(f).detach()
    w = f - i0
    i0 = i0.long()
    i1 = torch.remainder(i0 + 1, n) if periodic else torch.clamp(i0 + 1, max=n - 1)
    return i0.clamp(0, n - 1), i1, w


def sample(U, times, meta, x, t, periodic=True):
    """U (T,nx,ny,nz,3) float64 tensor; x (N,3) float64; t python float."""
    mode, i0, i1, a = blend_params(np.asarray(times, np.float32), np.float32(t))
    a = float(a)
    if mode == "interp":
        V = (1.0 - a) * U[i0] + a * U[i1]
    elif mode == "copy":
        V = U[i0]
    else:
        raise NotImplementedError(mode)
    nx, ny, nz = meta.n
    ix0, ix1, wx = _axis(x[:, 0], float(meta.lo[0]), float(meta.d[0]), nx, periodic)
    iy0, iy1, wy = _axis(x[:, 1], float(meta.lo[1]), float(meta.d[1]), ny, periodic)
    iz0, iz1, wz = _axis(x[:, 2], float(meta.lo[2]), float(meta.d[2]), nz, periodic)
    out = 0.0
    for I, fx in ((ix0, 1 - wx), (ix1, wx)):
        for J, fy in ((iy0, 1 - wy), (iy1, wy)):
            for K, fz in ((iz0, 1 - wz), (iz1, wz)):
                out = out + (fx * fy * fz)[:, None] * V[I, J, K]
    return out


def bc(x, lo, hi, kinds):
    cols = []
    for a, k in enumerate(kinds):
        c = x[:, a]
        if k == "periodic":
            c = lo[a] + torch.remainder(c - lo[a], hi[a] - lo[a])
        elif k == "reflective":
            w = hi[a] - lo[a]
            y = torch.remainder(c - lo[a], 2 * w)
            c = lo[a] + torch.where(y <= w, y, 2 * w - y)
        elif k == "clamping":
            c = torch.clamp(c, lo[a], hi[a])
        cols.append(c)
    return torch.stack(cols, 1)


def track(U, times, meta, x0, step_times, dt, kinds, lo=(0, 0, 0), hi=(1, 1, 1), periodic=True):
    """Final positions after len(step_times) RK4 steps (float64)."""
    x = x0
    dt = float(dt)
    for t in step_times:
        t = float(t)
        th, tf = float(np.float32(t) + np.float32(0.5) * np.float32(dt)), float(np.float32(t) + np.float32(dt))
        k1 = sample(U, times, meta, x, t, periodic)
        k2 = sample(U, times, meta, x + 0.5 * dt * k1, th, periodic)
        k3 = sample(U, times, meta, x + 0.5 * dt * k2, th, periodic)
        k4 = sample(U, times, meta, x + dt * k3, tf, periodic)
        x = x + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        x = bc(x, lo, hi, kinds)
    return x


def _stage_times(t, dt):
    """rk4.py:88-92 stage times in float32: t, t + 0.5 dt, t + dt."""
    t32, d32 = np.float32(t), np.float32(dt)
    return float(t32), float(t32 + np.float32(0.5) * d32), float(t32 + d32)


def track_np(U, times, meta, x0, step_times, dt, kinds, lo=(0, 0, 0), hi=(1, 1, 1), record=False):
    """``track`` in float64 NumPy for a periodic sampler, blending only the
    eight gathered corners of each sample (the same values: per-node blend
    then trilinear), so that large grids and many particles stay cheap.

    U: (T, nx, ny, nz, 3) array (float32 snapshots are read exactly, the
    arithmetic is float64).  Returns the final positions and,
    with record=True, a list with one entry per stage in run order:
    (lower cell corner (N,3) int64, snapshot pair (i0, i1), pair weights
    (1-a, a), smallest distance of the stage point to a cell face in cells)."""
    x = np.array(x0, dtype=np.float64)
    n = np.array(meta.n)
    o = np.array(meta.lo, np.float64)
    d = np.array(meta.d, np.float64)
    rec = []

    def sample(p, t):
        mode, s0, s1, a = blend_params(np.asarray(times, np.float32), np.float32(t))
        a = float(a)
        if mode == "copy":
            s1, a = s0, 0.0
        elif mode != "interp":
            raise NotImplementedError(mode)
        f = np.remainder((p - o) / d, n)
        i0 = np.floor(f).astype(np.int64)
        w = f - i0
        i0 = np.minimum(i0, n - 1)
        i1 = (i0 + 1) % n
        if record:
            rec.append((i0.copy(), (s0, s1), (1.0 - a, a), np.minimum(w, 1.0 - w).min(axis=1)))
        out = np.zeros_like(p)
        for I, fx in ((i0[:, 0], 1 - w[:, 0]), (i1[:, 0], w[:, 0])):
            for J, fy in ((i0[:, 1], 1 - w[:, 1]), (i1[:, 1], w[:, 1])):
                for K, fz in ((i0[:, 2], 1 - w[:, 2]), (i1[:, 2], w[:, 2])):
                    V = (1.0 - a) * U[s0, I, J, K].astype(np.float64) + a * U[s1, I, J, K].astype(np.float64)
                    out += (fx * fy * fz)[:, None] * V
        return out

    dt = float(dt)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    for t in step_times:
        t0, th, tf = _stage_times(float(t), dt)
        k1 = sample(x, t0)
        k2 = sample(x + 0.5 * dt * k1, th)
        k3 = sample(x + 0.5 * dt * k2, th)
        k4 = sample(x + dt * k3, tf)
        x = x + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        for a, k in enumerate(kinds):
            w = hi[a] - lo[a]
            if k == "periodic":
                x[:, a] = lo[a] + np.remainder(x[:, a] - lo[a], w)
            elif k == "reflective":
                y = np.remainder(x[:, a] - lo[a], 2 * w)
                x[:, a] = lo[a] + np.where(y <= w, y, 2 * w - y)
            elif k == "clamping":
                x[:, a] = np.clip(x[:, a], lo[a], hi[a])
    return (x, rec) if record else x
