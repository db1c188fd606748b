"""Particle tracking driver: ``create_tracker`` / ``ParticleTracker.track_particles``.

Mirrors jaxtrace/tracking/tracker.py (``TrackerOptions`` :107-153,
``track_particles`` :320-358, ``create_tracker`` :810-855) and
jaxtrace/tracking/particles.py (``Trajectory`` :82-131).  The whole time loop
runs in one C-ABI call (``jt_track_grid``) whose kernels keep each particle's
state in registers across stages and steps, with the boundary fused into the
step.  Fields and boundaries must be this package's device-backed kinds;
arbitrary Python callables are rejected with ``TypeError`` (there is no
second, host-side code path).
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, Optional, Tuple

import numpy as np
import torch

from . import _lib as L
from .boundary import BoundaryCondition
from .fields import OctreeFEMTimeSeriesFieldOptimized, TimeSeriesField, _GridSeries, _ensure_positions

F32 = np.float32
_INTEGRATORS = {"euler": L.JT_EULER, "rk2": L.JT_RK2, "rk4": L.JT_RK4}


@dataclass
class TrackerOptions:
    """tracker.py:107-153 plus device-side extensions.

    Reference fields are accepted with their meaning where it applies to a
    single fused GPU program (the JIT/scan switches have nothing to select).
    Extensions:
      record_every       keep x0 and every k-th step (compacted output)
      record_final_only  keep only x0 and the final positions
      return_device      Trajectory arrays stay torch tensors on the GPU
      steps_per_launch   steps fused into one kernel launch (0 = library default)
      sort_every         re-sort particles by grid cell every k steps (0 = off;
                         None = from 65536 particles, every 50 steps on
                         structured fields and every 10 on octree fields)
      output_path        stream the recorded frames into this .npy file,
                         (frames, N, 3) float32, while tracking runs; the
                         Trajectory's positions are then a read-only memmap of
                         it (velocities, when recorded, go to
                         '<stem>_velocities.npy').  Host memory holds two
                         chunks, not the whole trajectory.
      frames_per_chunk   recorded frames per streamed chunk (0 = about 1 GiB)
    """
    max_memory_gb: float = 8.0
    batch_size: Optional[int] = None
    oom_recovery: bool = True
    use_jax_jit: bool = True
    static_compilation: bool = True
    use_scan_jit: bool = True
    device_put_inputs: bool = True
    record_on_device: bool = True
    max_batch_size: int = 100_000
    min_batch_size: int = 100
    record_velocities: bool = False
    recording_interval: int = 1
    progress_callback: Optional[Callable[[float], None]] = None
    progress_style: str = "auto"
    progress_desc: str = "Tracking"
    progress_update_every: float = 0.05
    adaptive_dt: bool = False
    error_tolerance: float = 1e-6
    # extensions
    record_every: int = 1
    record_final_only: bool = False
    return_device: bool = False
    steps_per_launch: int = 0
    sort_every: Optional[int] = None
    output_path: Optional[str] = None
    frames_per_chunk: int = 0


class Trajectory:
    """particles.py:82-131: positions (T,N,3) f32, times (T,), velocities, metadata."""

    def __init__(self, positions, times, velocities=None, metadata: Optional[Dict[str, Any]] = None):
        self.positions = positions
        self.times = np.asarray(times, dtype=F32)
        self.velocities = velocities
        self.metadata = dict(metadata or {})
        shp = tuple(positions.shape)
        if len(shp) != 3 or shp[2] != 3:
            raise ValueError(f"Positions must have shape (T,N,3), got {shp}")
        self.T, self.N, self.D = shp
        if self.times.shape != (self.T,):
            raise ValueError(f"Times shape {self.times.shape} doesn't match trajectory length {self.T}")
        if velocities is not None and tuple(velocities.shape) != shp:
            raise ValueError("Velocities shape doesn't match positions shape")
        self.metadata.setdefault("format_version", "1.0")
        self.metadata.setdefault("dtype", "float32")
        self.metadata.setdefault("coordinate_system", "cartesian")

    def __len__(self):
        return self.T

    @property
    def num_particles(self):
        return self.N

    @property
    def num_timesteps(self):
        return self.T

    @property
    def duration(self):
        return float(self.times[-1] - self.times[0]) if self.T > 1 else 0.0

    def get_positions_at_time(self, t_idx: int):
        return self.positions[t_idx]

    def get_final_positions(self):
        return self.positions[-1]


@dataclass
class ParticleTracker:
    """tracker.py:155-535, driving the fused HIP kernels."""
    integrator_name: str
    field: Any
    boundary: BoundaryCondition
    options: TrackerOptions = field(default_factory=TrackerOptions)

    def __post_init__(self):
        if self.integrator_name not in _INTEGRATORS:
            raise ValueError(f"Unknown integrator: {self.integrator_name}. "
                             f"Available: {list(_INTEGRATORS)}")

    @property
    def integrator(self):
        return self.integrator_name

    def _opts(self, n: int):
        o = L.TrackOpts()
        o.steps_per_launch = int(self.options.steps_per_launch)
        se = self.options.sort_every
        if se is None:
            kind = getattr(self.field, "kind", "grid")
            se = 0 if n < 65536 else {"grid": 50, "octree": 10}.get(kind, 0)
        o.sort_every = int(se)
        o.block = 0
        return o

    def track_particles(self, initial_positions, time_span: Tuple[float, float], n_timesteps: int,
                        dt: Optional[float] = None, **kwargs) -> Trajectory:
        """tracker.py:320-358."""
        lib = L.load()
        fld = self.field
        dev = fld.device
        x, _ = _ensure_positions(initial_positions, dev)
        n = x.shape[0]
        t_start, t_end = time_span
        times = np.linspace(float(t_start), float(t_end), int(n_timesteps), dtype=np.float32)
        if dt is None:
            dt = float(times[1] - times[0]) if n_timesteps > 1 else 0.0
        else:
            dt = float(dt)
        n_steps = max(int(n_timesteps) - 1, 0)
        opt = self.options
        rec = n_steps if (opt.record_final_only and n_steps > 0) else max(int(opt.record_every), 1)
        n_rec = 1 + n_steps // rec
        rec_idx = np.arange(n_rec) * rec
        out_times = times[rec_idx] if n_steps > 0 else times[:1]

        step_times = torch.from_numpy(times[:-1].copy() if n_steps > 0 else np.zeros(1, F32)).to(dev)
        meta = {"integrator": self.integrator_name, "batch_processing": "single_batch",
                "n_particles": n, "n_timesteps": int(n_timesteps), "dt": dt,
                "record_every": rec, "backend": "hip-gfx950", "field_storage": fld.storage}
        state = x.clone()
        t0 = time.perf_counter()
        if opt.output_path is not None:
            if opt.return_device:
                raise ValueError("output_path streams frames to a file; it cannot be combined with return_device")
            traj, velocities = self._track_streamed(state, n, step_times, dt, n_steps, rec, out_times)
            meta["output_path"] = str(opt.output_path)
        else:
            launch = self._launcher(state, n, dt, n_steps)
            traj = torch.empty((n_rec, n, 3), dtype=torch.float32, device=dev)
            launch(step_times, n_steps, rec, traj)
            velocities = None
            if opt.record_velocities:
                velocities = torch.empty_like(traj)
                for r in range(n_rec):
                    velocities[r] = fld._sample(traj[r], float(out_times[r]))
            if not opt.return_device:
                traj = traj.cpu().numpy()
                if velocities is not None:
                    velocities = velocities.cpu().numpy()
        if opt.progress_callback is not None:
            opt.progress_callback(1.0)
        meta["wall_s"] = time.perf_counter() - t0
        return Trajectory(traj, out_times, velocities, meta)

    def _launcher(self, state, n, dt, max_steps):
        """The tracking call for this field and boundary, with its workspace
        sized for up to max_steps steps: launch(step_times, n_steps, rec, traj)
        advances state by n_steps steps from step_times[0] and writes state
        before the first step and after every rec-th step into traj."""
        lib = L.load()
        fld = self.field
        dev = fld.device
        desc = fld.descriptor()
        bc = self.boundary.descriptor(dev)
        o = self._opts(n)
        kind = getattr(fld, "kind", "grid")
        if kind == "octree":
            ws_bytes = int(lib.jt_track_octree_workspace_size(n, max_steps, C.byref(o)))
            entry, name = lib.jt_track_octree, "jt_track_octree"
        elif kind == "nodes":
            ws_bytes = int(lib.jt_track_nodes_workspace_size(n, max_steps))
            entry, name = lib.jt_track_nodes, "jt_track_nodes"
        else:
            ws_bytes = int(lib.jt_track_grid_workspace_size(n, max_steps, C.byref(o)))
            entry, name = lib.jt_track_grid, "jt_track_grid"
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)

        def launch(step_times, n_steps, rec, traj):
            if self.boundary.host_inlet:
                self._track_host_inlet(entry, name, desc, state, n, step_times, dt, n_steps, rec, traj)
            else:
                L.check(entry(C.byref(desc), C.byref(bc), _INTEGRATORS[self.integrator_name],
                              L.ptr(state), n, L.ptr(step_times), C.c_float(F32(dt)), n_steps,
                              rec, L.ptr(traj), C.byref(o), L.ptr(ws), ws_bytes,
                              L.stream_handle()), name)
        return launch

    def _track_streamed(self, state, n, step_times, dt, n_steps, rec, out_times):
        """Chunked tracking with the recorded frames streamed to .npy files.

        Chunk k advances the state by (frames in the chunk) x rec steps into
        device buffer k % 2 (about 1 GiB each by default) on the current
        stream.  While chunk k + 1 computes, a second stream (default
        priority) copies chunk k to the host in 64 MiB pieces through two
        pinned staging slots, and the host writes each landed piece into the
        file at its byte offset while the next piece copies.  Pinned memory
        stays at 128 MiB whatever the particle count.  Frames are the same
        values as one whole-run call: each particle's arithmetic does not
        depend on where launches split."""
        opt = self.options
        fld = self.field
        dev = fld.device
        n_rec = len(out_times)
        path = str(opt.output_path)
        vpath = path[:-4] + "_velocities.npy" if path.endswith(".npy") else path + "_velocities.npy"
        frame_bytes = max(n * 12, 1)
        fpc = int(opt.frames_per_chunk) if opt.frames_per_chunk > 0 else max(1, (1 << 30) // frame_bytes)
        fpc = max(1, min(fpc, max(n_rec - 1, 1)))
        pos_f = _NpyFrames(path, n_rec, n)
        vel_f = _NpyFrames(vpath, n_rec, n) if opt.record_velocities else None
        pos_f.write(0, state.cpu().numpy())
        if vel_f is not None:
            vel_f.write(0, fld._sample(state, float(out_times[0])).cpu().numpy())
        launch = self._launcher(state, n, dt, fpc * rec)
        comp = torch.cuda.current_stream(dev)
        copy = torch.cuda.Stream(dev)
        nb = 2 if n_rec - 1 > fpc else 1
        dbuf = [torch.empty((fpc + 1, n, 3), dtype=torch.float32, device=dev) for _ in range(nb)]
        vdbuf = None
        if vel_f is not None:
            vdbuf = [torch.empty((fpc, n, 3), dtype=torch.float32, device=dev) for _ in range(nb)]
        # two pinned staging slots: one copies on the copy stream while the
        # host writes the other
        piece = _STREAM_PIECE_BYTES // 4
        slots = [torch.empty(piece, dtype=torch.float32, pin_memory=True) for _ in range(2)]
        landed = [torch.cuda.Event() for _ in range(2)]
        computed = [torch.cuda.Event() for _ in range(nb)]
        drained = [torch.cuda.Event() for _ in range(nb)]

        def drain(b, f0, nf):
            """Copy chunk b's frames to the host piece by piece and write them."""
            srcs = [(dbuf[b][1:nf + 1].reshape(-1), pos_f)]
            if vdbuf is not None:
                srcs.append((vdbuf[b][:nf].reshape(-1), vel_f))
            base = (f0 + 1) * n * 3
            j, prev = 0, None
            with torch.cuda.stream(copy):
                copy.wait_event(computed[b])
                for src, out in srcs:
                    for lo in range(0, src.numel(), piece):
                        hi = min(src.numel(), lo + piece)
                        sl = j % 2
                        slots[sl][:hi - lo].copy_(src[lo:hi], non_blocking=True)
                        landed[sl].record(copy)
                        if prev is not None:
                            _write_piece(*prev)
                        prev = (slots[sl], landed[sl], out, base + lo, hi - lo)
                        j += 1
                drained[b].record(copy)
            if prev is not None:
                _write_piece(*prev)

        pending = None
        f0, k = 0, 0
        while f0 < n_rec - 1:
            nf = min(fpc, n_rec - 1 - f0)
            b = k % nb
            if k >= nb:
                comp.wait_event(drained[b])           # device buffer b is free again
            launch(step_times[f0 * rec:(f0 + nf) * rec], nf * rec, rec, dbuf[b][:nf + 1])
            if vdbuf is not None:
                for i in range(nf):
                    vdbuf[b][i] = fld._sample(dbuf[b][i + 1], float(out_times[f0 + 1 + i]))
            computed[b].record(comp)
            if pending is not None:
                drain(*pending)                         # the previous chunk, while this one computes
            pending = (b, f0, nf)
            f0 += nf
            k += 1
        if pending is not None:
            drain(*pending)
        torch.cuda.synchronize(dev)
        pos_f.close()
        positions = np.load(path, mmap_mode="r")
        velocities = None
        if vel_f is not None:
            vel_f.close()
            velocities = np.load(vpath, mmap_mode="r")
        return positions, velocities

    def _track_host_inlet(self, entry, name, desc, state, n, step_times, dt, n_steps, rec, traj):
        """A continuous inlet with host-drawn replacements: one step at a time
        with an open boundary, then the inlet boundary between steps (the
        reference applies it after each step, tracker.py:238-239)."""
        lib = L.load()
        dev = state.device
        open_bc = BoundaryCondition("_open_bc", self.boundary.bounds).descriptor(dev)
        o = L.TrackOpts()
        o.sort_every = 0
        kind = getattr(self.field, "kind", "grid")
        wsf = {"octree": lambda: lib.jt_track_octree_workspace_size(n, 1, C.byref(o)),
               "nodes": lambda: lib.jt_track_nodes_workspace_size(n, 1)}.get(
            kind, lambda: lib.jt_track_grid_workspace_size(n, 1, C.byref(o)))
        ws_bytes = int(wsf())
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        traj[0] = state
        for s in range(n_steps):
            L.check(entry(C.byref(desc), C.byref(open_bc), _INTEGRATORS[self.integrator_name], L.ptr(state), n,
                          L.ptr(step_times[s:s + 1]), C.c_float(F32(dt)), 1, 1, None, C.byref(o), L.ptr(ws),
                          ws_bytes, L.stream_handle()), name)
            self.boundary._apply_host_inlet(state)
            if (s + 1) % rec == 0:
                traj[(s + 1) // rec] = state


def create_tracker(integrator_name: str, field, boundary_condition, **options) -> ParticleTracker:
    """tracker.py:810-855."""
    if integrator_name not in _INTEGRATORS:
        raise ValueError(f"Unknown integrator: {integrator_name}. Available: {list(_INTEGRATORS)}")
    if not isinstance(field, (_GridSeries, OctreeFEMTimeSeriesFieldOptimized, TimeSeriesField)):
        raise TypeError("field must be a device-backed jaxtrace_amd field "
                        "(TimeSeriesField, StructuredGridSampler, StructuredTimeSeriesField, "
                        f"OctreeFEMTimeSeriesFieldOptimized); got {type(field).__name__}")
    if not isinstance(boundary_condition, BoundaryCondition):
        if callable(boundary_condition):
            raise TypeError("boundary_condition must come from a jaxtrace_amd boundary factory "
                            "(periodic/reflective/clamping/absorbing/mixed/continuous_inlet)")
        raise TypeError("boundary_condition must be callable")
    return ParticleTracker(integrator_name, field, boundary_condition, TrackerOptions(**options))


class _NpyFrames:
    """A (frames, n, 3) float32 .npy file written frame range by frame range
    with pwrite at each range's byte offset (no page faults on a mapping)."""

    def __init__(self, path: str, n_frames: int, n: int):
        mm = np.lib.format.open_memmap(path, mode="w+", dtype=np.float32, shape=(n_frames, n, 3))
        self.offset = int(mm.offset)
        del mm
        self.frame_bytes = n * 12
        self.fd = os.open(path, os.O_WRONLY)

    def write(self, first: int, frames: np.ndarray) -> None:
        self.write_values(first * (self.frame_bytes // 4), frames)

    def write_values(self, first_value: int, values: np.ndarray) -> None:
        """Write float32 values starting at flat value index first_value."""
        flat = np.ascontiguousarray(values, dtype=np.float32).reshape(-1).view(np.uint8)
        if flat.size == 0:
            return
        buf = memoryview(flat)
        pos = self.offset + first_value * 4
        while len(buf):
            k = os.pwrite(self.fd, buf, pos)
            buf, pos = buf[k:], pos + k

    def close(self) -> None:
        os.close(self.fd)


# pinned staging per slot for streamed output (two slots)
_STREAM_PIECE_BYTES = 64 << 20


def _write_piece(slot, landed, out, first_value, count):
    landed.synchronize()
    out.write_values(first_value, slot[:count].numpy())
