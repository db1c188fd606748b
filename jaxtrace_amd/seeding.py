"""Seed generation on the device: ``uniform_grid_seeds`` and ``random_seeds``.

Mirrors jaxtrace/tracking/seeding.py (``_validate_bounds`` :88-137,
``random_seeds`` :140-182, ``uniform_grid_seeds`` :185-247).  The three float32 coordinate arrays are
built with np.linspace exactly as the reference builds them (a few hundred
values); the (N, 3) meshgrid expansion, the part that grows with N, is the
``jt_grid_seeds`` HIP kernel writing straight into device memory, so a
2^24-seed grid never passes through host memory on its way to
``track_particles``.
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple, Union

import numpy as np
import torch

from . import _lib as L

F32 = np.float32


def _validate_bounds(bounds) -> np.ndarray:
    """seeding.py:88-137: (2, 3) float32 [[min], [max]] from the accepted forms."""
    b = np.asarray(bounds, dtype=F32)
    if b.ndim == 1:
        if len(b) == 4:
            b = np.array([[b[0], b[2], 0.0], [b[1], b[3], 0.0]], dtype=F32)
        elif len(b) == 6:
            b = np.array([[b[0], b[2], b[4]], [b[1], b[3], b[5]]], dtype=F32)
        else:
            raise ValueError(f"1D bounds must have 4 or 6 elements, got {len(b)}")
    elif b.ndim == 2:
        if b.shape == (2, 2):
            b = np.array([[b[0, 0], b[1, 0], 0.0], [b[0, 1], b[1, 1], 0.0]], dtype=F32)
        elif b.shape == (3, 2):
            b = b.T
        elif b.shape != (2, 3):
            raise ValueError(f"2D bounds must have shape (2,2), (2,3), or (3,2), got {b.shape}")
    else:
        raise ValueError(f"Bounds must be 1D or 2D array, got {b.ndim}D")
    if not np.all(b[0] < b[1]):
        raise ValueError(f"Invalid bounds: min {b[0]} >= max {b[1]}")
    return b


def grid_axes(resolution: Union[int, Tuple[int, ...]], bounds,
              include_boundaries: bool = True) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The reference's three coordinate arrays (seeding.py:209-236)."""
    b = _validate_bounds(bounds)
    if isinstance(resolution, (int, np.integer)):
        nx = ny = nz = int(resolution)
    elif len(resolution) == 2:
        (nx, ny), nz = resolution, 1
    elif len(resolution) == 3:
        nx, ny, nz = resolution
    else:
        raise ValueError(f"Resolution must be int or tuple of 2-3 ints, got {resolution}")
    (x_min, y_min, z_min), (x_max, y_max, z_max) = b[0], b[1]
    if include_boundaries:
        xc = np.linspace(x_min, x_max, nx, dtype=F32)
        yc = np.linspace(y_min, y_max, ny, dtype=F32)
        zc = np.linspace(z_min, z_max, nz, dtype=F32) if nz > 1 else np.array([z_min], dtype=F32)
    else:
        dx = (x_max - x_min) / (nx + 1)
        dy = (y_max - y_min) / (ny + 1)
        dz = (z_max - z_min) / (nz + 1) if nz > 1 else 0.0
        xc = np.linspace(x_min + dx, x_max - dx, nx, dtype=F32)
        yc = np.linspace(y_min + dy, y_max - dy, ny, dtype=F32)
        zc = np.linspace(z_min + dz, z_max - dz, nz, dtype=F32) if nz > 1 else np.array([z_min], dtype=F32)
    return xc, yc, zc


def uniform_grid_seeds(resolution, bounds, include_boundaries: bool = True, device=None,
                       return_device: bool = False):
    """seeding.py:185-247 with the expansion on the GPU.

    Returns the (N, 3) float32 seeds, as a NumPy array like the reference, or
    as a torch tensor on ``device`` (default cuda:0) with ``return_device``."""
    xc, yc, zc = grid_axes(resolution, bounds, include_boundaries)
    lib = L.load()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    ax = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (xc, yc, zc)]
    out = torch.empty((len(xc) * len(yc) * len(zc), 3), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        L.check(lib.jt_grid_seeds(L.ptr(ax[0]), len(xc), L.ptr(ax[1]), len(yc), L.ptr(ax[2]), len(zc),
                                  L.ptr(out), L.stream_handle()), "jt_grid_seeds")
    if return_device:
        return out
    return out.cpu().numpy()


def random_seeds(n: int, bounds, rng_seed: int = 0, dtype=np.float32, device=None, return_device: bool = False):
    """seeding.py:140-182 on the GPU: n positions lo + u * (hi - lo) with u =
    jax.random.uniform(PRNGKey(rng_seed), (n, 3)) computed by the
    jt_random_seeds kernel (threefry2x32, JAX's non-partitionable layout).
    Returns (n, 3) float32, NumPy like the reference or a torch tensor on
    ``device`` with ``return_device``."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if n <= 0:
        z = torch.zeros((0, 3), dtype=torch.float32, device=dev)
        return z if return_device else np.zeros((0, 3), dtype=F32)
    b = _validate_bounds(bounds)
    lib = L.load()
    out = torch.empty((int(n), 3), dtype=torch.float32, device=dev)
    lo = (C.c_float * 3)(*[float(v) for v in b[0]])
    hi = (C.c_float * 3)(*[float(v) for v in b[1]])
    with torch.cuda.device(dev):
        L.check(lib.jt_random_seeds(int(n), int(rng_seed) & 0xFFFFFFFFFFFFFFFF, C.cast(lo, C.c_void_p),
                                    C.cast(hi, C.c_void_p), L.ptr(out), L.stream_handle()), "jt_random_seeds")
    if return_device:
        return out
    return out.cpu().numpy()
