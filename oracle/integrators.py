"""Oracle: Euler / RK2-midpoint / RK4 steps in float32, reference op order.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows
  * jaxtrace/integrators/rk4.py:27-31 (f32 casts), :34-47 (combine), :62-106
  * jaxtrace/integrators/rk2.py:36-41, :154-180 (``t_mid`` is a float64 sum)
  * jaxtrace/integrators/euler.py:36-41, :145-157

``field_fn(x, t)`` receives f32 positions and an f32 (or, for RK2's midpoint,
float64) time.  Pinned in tests by: exact translation under a constant field,
the RK4 amplification polynomial under u = A x, exactness for fields that are
polynomial of degree <= 3 in t, and the float64 twin's convergence order.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def rk4_step(x, t, dt, field_fn, dtype=F32):
    """rk4.py:82-106 (scalar dt path)."""
    x = np.asarray(x, dtype=dtype)
    t = dtype(t)
    dt = dtype(dt)
    k1 = np.asarray(field_fn(x, t), dtype=dtype)
    dt_half = dtype(0.5) * dt
    t_half = dtype(t + dt_half)
    t_full = dtype(t + dt)
    x2 = x + dt_half * k1
    k2 = np.asarray(field_fn(x2, t_half), dtype=dtype)
    x3 = x + dt_half * k2
    k3 = np.asarray(field_fn(x3, t_half), dtype=dtype)
    x4 = x + dt * k3
    k4 = np.asarray(field_fn(x4, t_full), dtype=dtype)
    c = dtype(dt / dtype(6.0))
    return (x + c * (((k1 + dtype(2.0) * k2) + dtype(2.0) * k3) + k4)).astype(dtype)


def rk2_step(x, t, dt, field_fn, dtype=F32):
    """rk2.py:154-180: x + dt * f(x + dt/2 f(x,t), float(t) + float(dt/2))."""
    x = np.asarray(x, dtype=dtype)
    t = dtype(t)
    dt = dtype(dt)
    v1 = np.asarray(field_fn(x, t), dtype=dtype)
    dt_half = dt * dtype(0.5)
    x_mid = x + dt_half * v1
    t_mid = float(t) + float(dt_half)          # Python float64 sum (:165)
    v2 = np.asarray(field_fn(x_mid, dtype(t_mid)), dtype=dtype)
    return (x + dt * v2).astype(dtype)


def euler_step(x, t, dt, field_fn, dtype=F32):
    """euler.py:145-157."""
    x = np.asarray(x, dtype=dtype)
    t = dtype(t)
    dt = dtype(dt)
    v = np.asarray(field_fn(x, t), dtype=dtype)
    return (x + dt * v).astype(dtype)


STEPPERS = {"euler": euler_step, "rk2": rk2_step, "rk4": rk4_step}
