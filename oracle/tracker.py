"""Oracle: the tracking driver (time grid, step loop, recording).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows jaxtrace/tracking/tracker.py:72-87 (position shape/cast), :342-351
(``np.linspace(t0, t1, T, float32)`` time grid and default dt), :228-246 (T-1
steps from ``times[:-1]``, boundary after every full step, x0 stacked first).
"""
from __future__ import annotations

import numpy as np

from .integrators import STEPPERS

F32 = np.float32


def ensure_positions(p):
    """tracker.py:72-87."""
    pos = np.asarray(p, dtype=F32)
    if pos.ndim == 1:
        pos = pos.reshape(1, -1)
    if pos.ndim != 2:
        raise ValueError("Positions must be 2D")
    if pos.shape[1] == 2:
        pos = np.concatenate([pos, np.zeros((pos.shape[0], 1), F32)], axis=1)
    elif pos.shape[1] != 3:
        raise ValueError("Positions must have 2 or 3 columns")
    return pos


def time_grid(time_span, n_timesteps, dt=None):
    """tracker.py:346-351."""
    t0, t1 = time_span
    times = np.linspace(float(t0), float(t1), int(n_timesteps), dtype=F32)
    if dt is None:
        dt = float(times[1] - times[0]) if n_timesteps > 1 else 0.0
    else:
        dt = float(dt)
    return times, dt


def track(x0, time_span, n_timesteps, field_fn, bc_fn, integrator="rk4", dt=None,
          record_every=1, dtype=F32):
    """Return (times, positions) with positions of shape (1 + steps//record_every, N, 3)."""
    x = ensure_positions(x0).astype(dtype)
    times, dt = time_grid(time_span, n_timesteps, dt)
    step = STEPPERS[integrator]
    out = [x.copy()]
    for i in range(1, len(times)):
        x = step(x, times[i - 1], dt, field_fn, dtype=dtype)
        x = bc_fn(x).astype(dtype)
        if i % record_every == 0:
            out.append(x.copy())
    return times, np.stack(out, axis=0)
