"""Differentiable RK4 tracking through a structured time-series field.

``track_rk4(U, times, grid_x, grid_y, grid_z, x0, time_span, n_timesteps, ...)``
returns the final positions x_T as a torch tensor whose autograd graph
reaches the initial positions ``x0`` and the snapshots ``U`` (T,Nx,Ny,Nz,3).
Forward: ``jt_track_grid`` (the same fused kernel as ``track_particles``)
recording positions every ``checkpoint_every`` steps.  Backward:
``jt_track_grid_adjoint``, a hand-written adjoint RK4 kernel that recomputes
each segment from its checkpoint and scatter-adds dL/dU with float atomics.
Snapshots may be stored as bfloat16/float16 for the forward; gradients are
float32 with respect to the stored values.

The reference has no differentiation (no grad/custom_vjp anywhere); this is
the build-only training-system pattern of the hot path.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib as L
from .boundary import BoundaryCondition
from .fields import StructuredTimeSeriesField

F32 = np.float32


class _Spec:
    pass


class _GridTrackRK4(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, U, spec):
        lib = L.load()
        dev = x0.device
        field = StructuredTimeSeriesField(U.detach(), spec.times, spec.gx, spec.gy, spec.gz,
                                          bounds_handling=spec.bounds, storage=spec.storage,
                                          layout=spec.layout, device=dev)
        desc = field.descriptor()
        bc = spec.bc.descriptor(dev)
        n = x0.shape[0]
        K = spec.K
        n_ck = spec.n_steps // K + 1
        ckpt = torch.empty((n_ck, n, 3), dtype=torch.float32, device=dev)
        state = x0.detach().to(torch.float32).contiguous().clone()
        o = L.TrackOpts()
        o.sort_every = spec.sort_every if n >= 65536 else 0
        ws_bytes = int(lib.jt_track_grid_workspace_size(n, spec.n_steps, C.byref(o)))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        L.check(lib.jt_track_grid(C.byref(desc), C.byref(bc), L.JT_RK4, L.ptr(state), n, L.ptr(spec.step_times),
                                  C.c_float(spec.dt), spec.n_steps, K, L.ptr(ckpt), C.byref(o), L.ptr(ws), ws_bytes,
                                  L.stream_handle()), "jt_track_grid")
        ctx.field = field
        ctx.spec = spec
        ctx.save_for_backward(ckpt)
        ctx.U_shape = tuple(U.shape)
        return state

    @staticmethod
    def backward(ctx, g_xT):
        lib = L.load()
        (ckpt,) = ctx.saved_tensors
        spec, field = ctx.spec, ctx.field
        dev = ckpt.device
        n = ckpt.shape[1]
        g = g_xT.detach().to(torch.float32).contiguous()
        gx0 = torch.empty((n, 3), dtype=torch.float32, device=dev)
        gU = torch.zeros(ctx.U_shape, dtype=torch.float32, device=dev)
        ws_bytes = int(lib.jt_track_grid_adjoint_workspace_size(n, spec.n_steps, spec.K))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        hst = np.ascontiguousarray(spec.step_times_host, dtype=F32)
        hft = np.ascontiguousarray(field.times, dtype=F32)
        L.check(lib.jt_track_grid_adjoint(C.byref(field.descriptor()), C.byref(spec.bc.descriptor(dev)), L.ptr(ckpt),
                                          n, L.ptr(spec.step_times), hst.ctypes.data, hft.ctypes.data,
                                          C.c_float(spec.dt), spec.n_steps, spec.K,
                                          L.ptr(g), L.ptr(gx0), L.ptr(gU), L.ptr(ws), ws_bytes, L.stream_handle()),
                "jt_track_grid_adjoint")
        return gx0, gU, None


def track_rk4(U: torch.Tensor, times, grid_x, grid_y, grid_z, x0: torch.Tensor, time_span, n_timesteps: int,
              dt: Optional[float] = None, boundary: Optional[BoundaryCondition] = None,
              bounds_handling: str = "periodic", storage: str = "float32", layout: str = "halo",
              checkpoint_every: int = 32, sort_every: int = 32) -> torch.Tensor:
    """Final positions after RK4 tracking, differentiable w.r.t. ``x0`` and ``U``.

    Time grid and dt follow track_particles (tracker.py:346-351)."""
    if U.ndim != 5 or U.shape[-1] != 3:
        raise ValueError("U must be (T, Nx, Ny, Nz, 3)")
    if x0.ndim != 2 or x0.shape[1] != 3:
        raise ValueError("x0 must be (N, 3)")
    if checkpoint_every < 1:
        raise ValueError("checkpoint_every must be >= 1")
    spec = _Spec()
    t0, t1 = time_span
    tg = np.linspace(float(t0), float(t1), int(n_timesteps), dtype=np.float32)
    spec.dt = F32(float(tg[1] - tg[0]) if (dt is None and n_timesteps > 1) else (0.0 if dt is None else dt))
    spec.n_steps = max(int(n_timesteps) - 1, 0)
    spec.step_times_host = tg[:-1].copy() if spec.n_steps else np.zeros(1, F32)
    spec.step_times = torch.from_numpy(spec.step_times_host).to(x0.device)
    spec.times = np.asarray(times, F32)
    spec.gx, spec.gy, spec.gz = grid_x, grid_y, grid_z
    spec.bounds = bounds_handling
    spec.storage, spec.layout = storage, layout
    spec.K = int(checkpoint_every)
    spec.sort_every = int(sort_every)
    spec.bc = boundary if boundary is not None else BoundaryCondition("_open_bc", np.array([[0, 0, 0], [1, 1, 1]], F32))
    return _GridTrackRK4.apply(x0, U, spec)
