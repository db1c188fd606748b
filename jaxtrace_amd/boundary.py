"""Boundary conditions (jaxtrace/tracking/boundary.py) as device descriptors.

Each factory returns a ``BoundaryCondition``: a callable with the reference's
closure name (``_periodic_bc`` ...) that applies the condition on the GPU
(``jt_apply_boundary``), and whose descriptor the tracking kernels fuse into
every step.  Factories, argument meaning and errors follow the reference:

* periodic_boundary / reflective_boundary / clamping_boundary /
  absorbing_boundary_factory   -- boundary.py:215-337
* mixed_boundary               -- boundary.py:377-438 (per axis, x then y then z)
* continuous_inlet_boundary_factory -- boundary.py:1444-1601 with the
  grid-preserving replacement of :1249-1318
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional, Union

import numpy as np
import torch

from . import _lib as L

F32 = np.float32


def _ensure_bounds_shape(bounds) -> np.ndarray:
    """boundary.py:39-86."""
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
        elif b.shape == (2, 3):
            pass
        elif b.shape == (3, 2):
            b = b.T
        else:
            raise ValueError(f"2D bounds must have shape (2,2), (2,3), or (3,2), got {b.shape}")
    else:
        raise ValueError(f"Bounds must be 1D or 2D array, got {b.ndim}D")
    if not np.all(b[0] < b[1]):
        raise ValueError(f"Invalid bounds: min {b[0]} >= max {b[1]}")
    return np.ascontiguousarray(b.astype(F32))


class BoundaryCondition:
    """A boundary condition the tracking kernels can fuse into each step."""

    def __init__(self, name: str, bounds: np.ndarray, axis_kind=(0, 0, 0), absorb_rows=False,
                 outside_value=np.nan, inlet=None):
        self.__name__ = name
        self.bounds = bounds
        self.axis_kind = tuple(int(k) for k in axis_kind)
        self.absorb_rows = bool(absorb_rows)
        self.outside_value = F32(outside_value)
        self.inlet = inlet           # dict or None
        self._dev_cache = {}

    def descriptor(self, device) -> L.Boundary:
        key = str(device)
        if key in self._dev_cache:
            return self._dev_cache[key][0]
        d = L.Boundary()
        d.axis_kind[:] = list(self.axis_kind)
        d.lo[:] = [F32(v) for v in self.bounds[0]]
        d.hi[:] = [F32(v) for v in self.bounds[1]]
        d.absorb_rows = int(self.absorb_rows)
        d.outside_value = self.outside_value
        keep = []
        if self.inlet is not None:
            inl = self.inlet
            d.inlet = inl["mode"]
            d.inlet_axis = inl["axis"]
            d.inlet_positive = int(inl["positive"])
            d.inlet_backstop = inl["backstop"]
            for j, coords in enumerate(inl["coords"]):
                t = torch.from_numpy(np.ascontiguousarray(coords, dtype=F32)).to(device)
                keep.append(t)
                d.inlet_coords[j] = t.data_ptr()
                d.inlet_ncoords[j] = t.numel()
        self._dev_cache[key] = (d, keep)
        return d

    def __call__(self, positions):
        """Apply on the GPU; returns the same array type it was given."""
        lib = L.load()
        dev = positions.device if isinstance(positions, torch.Tensor) and positions.is_cuda else L.device()
        is_torch = isinstance(positions, torch.Tensor)
        x = (positions.detach().to(device=dev, dtype=torch.float32) if is_torch
             else torch.from_numpy(np.asarray(positions, dtype=F32)).to(dev))
        x = x.reshape(-1, 3).clone().contiguous()
        if self.host_inlet:
            self._apply_host_inlet(x)
        else:
            L.check(lib.jt_apply_boundary(C.byref(self.descriptor(dev)), L.ptr(x), x.shape[0],
                                          L.stream_handle()), "jt_apply_boundary")
        return x if is_torch else x.cpu().numpy()

    @property
    def host_inlet(self) -> bool:
        """A continuous inlet whose replacements are drawn on the host
        ('uniform'/'gaussian', boundary.py:1552-1571)."""
        return self.inlet is not None and self.inlet["mode"] == L.JT_INLET_HOST

    def _draw_inlet(self, n_exited: int) -> np.ndarray:
        """The off-axis coordinates of n_exited replacements, drawn from the
        global np.random state with the reference's calls in its order: all
        of the lower off-axis, then all of the higher (boundary.py:1555-1571)."""
        a = self.inlet["axis"]
        cols = []
        for i in range(3):
            if i == a:
                continue
            dim_min, dim_max = self.bounds[0, i], self.bounds[1, i]
            if self.inlet["distribution"] == "uniform":
                cols.append(np.random.uniform(dim_min, dim_max, n_exited).astype(np.float32))
            else:
                center = (dim_min + dim_max) / 2
                sigma = (dim_max - dim_min) / 6
                cols.append(np.clip(np.random.normal(center, sigma, n_exited), dim_min, dim_max).astype(np.float32))
        return np.concatenate(cols)

    def _apply_host_inlet(self, x: torch.Tensor) -> None:
        """In place on the device positions x (n, 3): rank the particles that
        left through the outlet face, draw their replacements on the host, and
        place them (jt_inlet_exits, np.random, jt_inlet_place)."""
        lib = L.load()
        dev = x.device
        n = x.shape[0]
        d = self.descriptor(dev)
        rank = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int32, device=dev)
        wsb = int(lib.jt_inlet_exits_workspace_size(n))
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        L.check(lib.jt_inlet_exits(C.byref(d), L.ptr(x), n, L.ptr(rank), L.ptr(cnt), L.ptr(ws), wsb,
                                   L.stream_handle()), "jt_inlet_exits")
        m = int(cnt.item())
        draws = torch.from_numpy(self._draw_inlet(m)).to(dev) if m > 0 else None
        L.check(lib.jt_inlet_place(C.byref(d), L.ptr(x), n, L.ptr(rank), L.ptr(draws), m, L.stream_handle()),
                "jt_inlet_place")

    def __repr__(self):
        return f"BoundaryCondition({self.__name__}, kinds={self.axis_kind}, inlet={self.inlet is not None})"


def periodic_boundary(bounds) -> BoundaryCondition:
    """boundary.py:215-240."""
    return BoundaryCondition("_periodic_bc", _ensure_bounds_shape(bounds), (L.JT_BC_PERIODIC,) * 3)


def reflective_boundary(bounds) -> BoundaryCondition:
    """boundary.py:243-268."""
    return BoundaryCondition("_reflective_bc", _ensure_bounds_shape(bounds), (L.JT_BC_REFLECTIVE,) * 3)


def clamping_boundary(bounds) -> BoundaryCondition:
    """boundary.py:271-296."""
    return BoundaryCondition("_clamping_bc", _ensure_bounds_shape(bounds), (L.JT_BC_CLAMP,) * 3)


def absorbing_boundary_factory(bounds, outside_value: float = np.nan) -> BoundaryCondition:
    """boundary.py:299-337: a particle with any coordinate outside gets
    ``outside_value`` in all three coordinates."""
    return BoundaryCondition("_absorbing_bc", _ensure_bounds_shape(bounds), (L.JT_BC_NONE,) * 3,
                             absorb_rows=True, outside_value=outside_value)


_MIXED = {"periodic": L.JT_BC_PERIODIC, "reflective": L.JT_BC_REFLECTIVE,
          "clamping": L.JT_BC_CLAMP, "absorbing": L.JT_BC_ABSORB}


def mixed_boundary(bounds, conditions: Dict[str, str]) -> BoundaryCondition:
    """boundary.py:377-438: per-axis kinds, missing axes default to periodic."""
    b = _ensure_bounds_shape(bounds)
    kinds = []
    for axis in ("x", "y", "z"):
        c = conditions.get(axis, "periodic")
        if c not in _MIXED:
            raise ValueError(f"Unknown condition type: {c}")
        kinds.append(_MIXED[c])
    return BoundaryCondition("_mixed_bc", b, kinds, outside_value=np.nan)


def continuous_inlet_boundary_factory(bounds, flow_axis: str = "x", flow_direction: str = "positive",
                                      inlet_distribution: str = "uniform",
                                      concentrations: Optional[Dict[str, int]] = None) -> BoundaryCondition:
    """boundary.py:1444-1601.  The 'grid' distribution (with concentrations)
    snaps an exited particle's off-axis coordinates to the nearest node of
    ``np.linspace(lo, hi, conc, dtype=float32)`` (:1249-1318)."""
    b = _ensure_bounds_shape(bounds)
    axis_map = {"x": 0, "y": 1, "z": 2}
    if flow_axis not in axis_map:
        raise ValueError(f"flow_axis must be 'x', 'y', or 'z', got {flow_axis}")
    a = axis_map[flow_axis]
    positive = flow_direction == "positive"
    amin, amax = b[0, a], b[1, a]
    # :1580-1587, NumPy scalar arithmetic exactly as the reference evaluates it
    backstop = F32(amin + 0.001 * (amax - amin)) if positive else F32(amax - 0.001 * (amax - amin))
    others = [i for i in range(3) if i != a]
    if inlet_distribution == "grid" and concentrations is not None:
        names = "xyz"
        coords = [np.linspace(b[0, i], b[1, i], max(1, int(concentrations[names[i]])), dtype=np.float32)
                  for i in others]
        mode = L.JT_INLET_GRID
    elif inlet_distribution in ("uniform", "gaussian"):
        # the replacements come from the host np.random state every step
        # (:1552-1571): the tracker steps with an open boundary and applies
        # this one between steps (jt_inlet_exits / np.random / jt_inlet_place)
        coords = [np.zeros(1, F32), np.zeros(1, F32)]
        mode = L.JT_INLET_HOST
    else:
        coords = [np.zeros(1, F32), np.zeros(1, F32)]
        mode = L.JT_INLET_ZERO
    inlet = {"mode": mode, "axis": a, "positive": positive, "backstop": backstop, "coords": coords,
             "distribution": inlet_distribution}
    return BoundaryCondition("_continuous_inlet_bc", b, (L.JT_BC_NONE,) * 3, inlet=inlet)


# ---- array functions of the reference (boundary.py:101-210) -----------------

def apply_periodic(x, bounds):
    return periodic_boundary(bounds)(x)


def clamp_to_bounds(x, bounds):
    return clamping_boundary(bounds)(x)


def reflect_boundary(x, bounds):
    return reflective_boundary(bounds)(x)


def absorbing_boundary(x, bounds, outside_value: float = np.nan):
    return absorbing_boundary_factory(bounds, outside_value)(x)


def _inlet_grid_indices(n_particles: int, total: int) -> np.ndarray:
    """The face-grid indices of boundary.py:1387-1424: n systematic samples
    int(i * total / n) (gaps filled at the largest one if rounding merged two),
    or every grid point then int(i * total / extra) for the extra ones."""
    if n_particles <= total:
        step = total / n_particles
        idx = (np.arange(n_particles, dtype=np.float64) * step).astype(np.int64)
        idx = np.unique(np.clip(idx, 0, total - 1))
        while len(idx) < n_particles and len(idx) < total:
            gaps = np.diff(np.concatenate([[0], idx, [total]]))
            g = int(np.argmax(gaps[:-1]))
            new = idx[0] // 2 if g == 0 else (idx[g - 1] + idx[g]) // 2
            idx = np.unique(np.append(idx, new))
        return idx[:n_particles]
    extra = n_particles - total
    step = total / extra
    return np.concatenate([np.arange(total), (np.arange(extra, dtype=np.float64) * step).astype(np.int64)])


def generate_inlet_grid_particles(n_particles: int, inlet_position: float, axis_idx: int, bounds,
                                  concentrations: Dict[str, int], device=None, return_device: bool = False):
    """boundary.py:1321-1441 (3-D bounds): n_particles on the inlet face at
    inlet_position along axis_idx, on the face grid np.linspace(lo, hi, c) of
    the two other axes (c from concentrations).  The index choice is the
    reference's host arithmetic; the (n, 3) expansion is the
    jt_inlet_grid_particles kernel.  NumPy, or a torch tensor with
    return_device."""
    b = _ensure_bounds_shape(bounds)
    if axis_idx not in (0, 1, 2):
        raise ValueError(f"axis_idx must be 0, 1 or 2, got {axis_idx}")
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    names = ["x", "y", "z"]
    other = [i for i in range(3) if i != axis_idx]
    res = [max(1, int(concentrations[names[i]])) for i in other]
    c0 = np.linspace(b[0, other[0]], b[1, other[0]], res[0], dtype=np.float32)
    c1 = np.linspace(b[0, other[1]], b[1, other[1]], res[1], dtype=np.float32)
    n = int(n_particles)
    out = torch.empty((max(n, 0), 3), dtype=torch.float32, device=dev)
    if n > 0:
        idx = torch.from_numpy(_inlet_grid_indices(n, res[0] * res[1]).astype(np.int32)).to(dev)
        c0d, c1d = (torch.from_numpy(a).to(dev) for a in (c0, c1))
        lib = L.load()
        with torch.cuda.device(dev):
            L.check(lib.jt_inlet_grid_particles(L.ptr(idx), n, int(axis_idx), C.c_float(np.float32(inlet_position)),
                                                L.ptr(c0d), res[0], L.ptr(c1d), res[1], L.ptr(out),
                                                L.stream_handle()), "jt_inlet_grid_particles")
    if return_device:
        return out
    return out.cpu().numpy()
