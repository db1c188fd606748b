"""Oracle: boundary conditions in float32.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows
  * jaxtrace/tracking/boundary.py:39-86   (``_ensure_bounds_shape``)
  * :101-123 ``apply_periodic``, :126-143 ``clamp_to_bounds``,
    :146-172 ``reflect_boundary``, :175-210 ``absorbing_boundary``
  * :377-438 ``mixed_boundary`` (per-axis application, x then y then z)
  * :1444-1601 ``continuous_inlet_boundary_factory`` and
    :1249-1318 ``_replace_exited_particles_at_grid_positions``

NumPy-2 scalar promotion is used for the inlet's host arithmetic (a Python
float times an f32 scalar stays f32), matching the NumPy this runs on.
Pinned in tests by: output inside [lo, hi] (periodic may return hi when
x - lo = -eps), idempotence on interior points, closed-form wrap/reflection
of hand-picked values, and the inlet snap equal to argmin over linspace.
"""
from __future__ import annotations

import numpy as np

from .fields import jmod, clip

F32 = np.float32


def ensure_bounds(bounds):
    """boundary.py:39-86."""
    b = np.asarray(bounds, dtype=F32)
    if b.ndim == 1:
        if len(b) == 4:
            b = np.array([[b[0], b[2], 0.0], [b[1], b[3], 0.0]], dtype=F32)
        elif len(b) == 6:
            b = np.array([[b[0], b[2], b[4]], [b[1], b[3], b[5]]], dtype=F32)
        else:
            raise ValueError("1D bounds must have 4 or 6 elements")
    elif b.ndim == 2:
        if b.shape == (2, 2):
            b = np.array([[b[0, 0], b[1, 0], 0.0], [b[0, 1], b[1, 1], 0.0]], dtype=F32)
        elif b.shape == (2, 3):
            pass
        elif b.shape == (3, 2):
            b = b.T
        else:
            raise ValueError("bad bounds shape")
    else:
        raise ValueError("bad bounds ndim")
    if not np.all(b[0] < b[1]):
        raise ValueError("Invalid bounds")
    return b.astype(F32)


def periodic(x, lo, hi):
    """boundary.py:117-123 for one coordinate array."""
    w = np.maximum(F32(hi) - F32(lo), F32(0))
    w_safe = w if w > 0 else F32(1)
    return (F32(lo) + jmod((np.asarray(x, F32) - F32(lo)).astype(F32), w_safe)).astype(F32)


def reflective(x, lo, hi):
    """boundary.py:165-172."""
    lo, hi = F32(lo), F32(hi)
    w = np.maximum(hi - lo, F32(0))
    period = F32(2.0) * (w if w > 0 else F32(1))
    y = jmod((np.asarray(x, F32) - lo).astype(F32), period)
    y_ref = np.where(y <= w, y, F32(2.0) * w - y).astype(F32)
    return (lo + y_ref).astype(F32)


def clamp(x, lo, hi):
    """boundary.py:143."""
    return clip(np.asarray(x, F32), F32(lo), F32(hi))


def absorb_axis(x, lo, hi, outside=np.nan):
    """Per-axis absorbing as ``mixed_boundary`` applies it (:398-436): only
    the offending coordinate becomes ``outside``."""
    x = np.asarray(x, F32)
    out = (x < F32(lo)) | (x > F32(hi))
    return np.where(out, F32(outside), x).astype(F32)


def absorb_rows(x, bounds, outside=np.nan):
    """boundary.py:190-210: any coordinate outside -> whole row ``outside``."""
    x = np.asarray(x, F32)
    b = ensure_bounds(bounds)
    m = np.any(x < b[0], axis=1) | np.any(x > b[1], axis=1)
    return np.where(m[:, None], F32(outside), x).astype(F32)


AXIS_FN = {"periodic": periodic, "reflective": reflective, "clamping": clamp,
           "absorbing": absorb_axis}


def per_axis(x, bounds, kinds):
    """Standalone factories (all axes one kind) and ``mixed_boundary``."""
    b = ensure_bounds(bounds)
    out = np.array(x, dtype=F32, copy=True)
    for i, k in enumerate(kinds):
        if k is None or k == "none":
            continue
        out[:, i] = AXIS_FN[k](out[:, i], b[0, i], b[1, i])
    return out


def continuous_inlet(x, bounds, flow_axis="x", flow_direction="positive",
                     inlet_distribution="grid", concentrations=None):
    """boundary.py:1515-1599 (grid replacement :1249-1318, random
    replacement :1552-1571)."""
    b = ensure_bounds(bounds)
    a = {"x": 0, "y": 1, "z": 2}[flow_axis]
    res = np.array(x, dtype=F32, copy=True)
    amin, amax = b[0, a], b[1, a]
    if flow_direction == "positive":
        exit_mask = res[:, a] > amax
        inlet = amin
    else:
        exit_mask = res[:, a] < amin
        inlet = amax
    if np.any(exit_mask):
        if inlet_distribution == "grid" and concentrations is not None:
            others = [i for i in range(3) if i != a]
            names = "xyz"
            coords = [np.linspace(b[0, i], b[1, i], max(1, int(concentrations[names[i]])),
                                  dtype=F32) for i in others]
            for p in np.where(exit_mask)[0]:
                g = [coords[j][np.argmin(np.abs(coords[j] - res[p, ax]))]
                     for j, ax in enumerate(others)]
                res[p, a] = inlet
                res[p, others[0]] = g[0]
                res[p, others[1]] = g[1]
        elif inlet_distribution in ("uniform", "gaussian"):
            # :1552-1571: draws from the global np.random state, one off-axis
            # at a time in axis order, n_exited values each
            n_ex = int(np.sum(exit_mask))
            newp = np.zeros((n_ex, 3), dtype=F32)
            newp[:, a] = inlet
            for i in range(3):
                if i != a:
                    dim_min, dim_max = b[0, i], b[1, i]
                    if inlet_distribution == "uniform":
                        newp[:, i] = np.random.uniform(dim_min, dim_max, n_ex).astype(F32)
                    else:
                        center = (dim_min + dim_max) / 2
                        sigma = (dim_max - dim_min) / 6
                        newp[:, i] = np.clip(np.random.normal(center, sigma, n_ex), dim_min, dim_max).astype(F32)
            res[exit_mask] = newp
        else:
            # non-grid fallback with an unknown distribution: zeros off-axis
            newp = np.zeros((int(exit_mask.sum()), 3), dtype=F32)
            newp[:, a] = inlet
            res[exit_mask] = newp
    if flow_direction == "positive":
        bm = res[:, a] < amin
        res[bm, a] = amin + 0.001 * (amax - amin)
    else:
        bm = res[:, a] > amax
        res[bm, a] = amax - 0.001 * (amax - amin)
    for i in range(3):
        if i != a:
            low = res[:, i] < b[0, i]
            high = res[:, i] > b[1, i]
            res[low, i] = 2 * b[0, i] - res[low, i]
            res[high, i] = 2 * b[1, i] - res[high, i]
    return res.astype(F32)
