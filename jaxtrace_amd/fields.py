"""Velocity fields backed by device-resident snapshots and HIP sampling kernels.

Mirrors the reference's field classes (same constructor arguments, same
errors) while holding torch tensors on the ROCm device:

* ``StructuredGridSampler``      -- jaxtrace/fields/structured.py:65-357
* ``StructuredTimeSeriesField``  -- the configs' time-dependent structured
  field: TimeSeriesField's temporal blend (time_series.py:155-245) composed
  with StructuredGridSampler's trilinear interpolation (structured.py:167-250).
  The reference has no single class for it; this is that composition.

Snapshots may be stored in float32 or in bfloat16/float16 (padded to 4
components, 8 bytes per node) to halve the gather traffic; all arithmetic
stays float32.  The default 'halo' layout adds one z-plane so the two
z-corners of a cell are one contiguous load, with 4 components per node (16
bytes in float32: 16-byte aligned loads, measured 6 % faster than 12-byte
nodes); 'halo3' is the same with 3 float32 components; 'packed' keeps the
caller's float32 (T,Nx,Ny,Nz,3) array without a copy.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib as L

F32 = np.float32
_STORAGE = {"float32": (L.JT_DTYPE_F32, torch.float32), "bfloat16": (L.JT_DTYPE_BF16, torch.bfloat16),
            "float16": (L.JT_DTYPE_F16, torch.float16)}


def _as_f32_device(a, dev) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=dev, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=F32))).to(dev)


def _ensure_positions(positions, dev):
    """(N,2|3) -> (N,3) float32 on device (structured.py:49-61, tracker.py:72-87)."""
    is_torch = isinstance(positions, torch.Tensor)
    x = _as_f32_device(positions, dev)
    if x.ndim == 1:
        x = x.reshape(1, -1)
    if x.ndim != 2:
        raise ValueError(f"Positions must be 2D array, got shape {tuple(x.shape)}")
    if x.shape[1] == 2:
        x = torch.cat([x, torch.zeros((x.shape[0], 1), dtype=torch.float32, device=dev)], dim=1)
    elif x.shape[1] != 3:
        raise ValueError(f"Positions must have 2 or 3 columns, got {x.shape[1]}")
    return x.contiguous(), is_torch


class _GridSeries:
    """Shared machinery: a (T, Nx, Ny, Nz, 3) snapshot series on a uniform grid."""

    kind = "grid"

    def _setup(self, data, times, grid_x, grid_y, grid_z, bounds_handling, interpolation,
               extrapolation, storage, layout, device, pair_records=False):
        self.device = torch.device(device) if device is not None else L.device()
        mode = str(bounds_handling).lower()
        if mode not in L.JT_BOUNDS:
            raise ValueError(f"Unsupported bounds_handling: {bounds_handling}")
        interp = str(interpolation).lower()
        extra = str(extrapolation).lower()
        if interp not in ("nearest", "linear", "cubic"):
            raise ValueError(f"Unsupported interpolation: {interpolation}")
        if extra not in L.JT_EXTRAP:
            raise ValueError(f"Unsupported extrapolation: {extrapolation}")
        if storage not in _STORAGE:
            raise ValueError(f"storage must be one of {sorted(_STORAGE)}, got {storage!r}")
        if layout not in ("packed", "padded", "halo", "halo3", "halo4"):
            raise ValueError("layout must be 'packed', 'padded', 'halo', 'halo3' or 'halo4'")
        gx = np.asarray(grid_x.cpu() if isinstance(grid_x, torch.Tensor) else grid_x, dtype=F32)
        gy = np.asarray(grid_y.cpu() if isinstance(grid_y, torch.Tensor) else grid_y, dtype=F32)
        gz = np.asarray(grid_z.cpu() if isinstance(grid_z, torch.Tensor) else grid_z, dtype=F32)
        T, Nx, Ny, Nz, D = data.shape
        if D != 3:
            raise ValueError(f"velocity_data must have 3 components, got {D}")
        for name, g, n in (("grid_x", gx, Nx), ("grid_y", gy, Ny), ("grid_z", gz, Nz)):
            if g.shape != (n,):
                raise ValueError(f"{name} shape {g.shape} doesn't match velocity_data ({n})")
        tt = np.asarray(times.cpu() if isinstance(times, torch.Tensor) else times, dtype=F32).reshape(-1)
        if tt.shape != (T,):
            raise ValueError(f"times shape {tt.shape} != (T,) with T={T}")
        self.Nx, self.Ny, self.Nz, self.T = Nx, Ny, Nz, T
        # structured.py:123-130: dx = float(grid_x[1] - grid_x[0]) in f32
        self.dx = F32(gx[1] - gx[0]) if Nx > 1 else F32(1.0)
        self.dy = F32(gy[1] - gy[0]) if Ny > 1 else F32(1.0)
        self.dz = F32(gz[1] - gz[0]) if Nz > 1 else F32(1.0)
        self.x_min, self.x_max = float(gx[0]), float(gx[-1])
        self.y_min, self.y_max = float(gy[0]), float(gy[-1])
        self.z_min, self.z_max = float(gz[0]), float(gz[-1])
        self.grid_x, self.grid_y, self.grid_z = gx, gy, gz
        self.bounds_handling, self.interpolation, self.extrapolation = mode, interp, extra
        self.storage, self.layout = storage, layout
        dt_code, tdt = _STORAGE[storage]
        # packed: (T,Nx,Ny,Nz,3) f32 as given; padded: 4 components per node;
        # halo/halo3/halo4: one extra z-plane (plane Nz repeats plane 0 for periodic
        # bounds, plane Nz-1 otherwise) so both z-corners of a cell are one load
        # 16-bit nodes are (x, y, 0, z): z in the high half of the second
        # dword over a zero low half, which the kernels read as float32 bits
        ncomp = 3 if (storage == "float32" and layout in ("packed", "halo3")) else 4
        z_halo = 1 if layout in ("halo", "halo3", "halo4") else 0
        self._slots = [0, 1, 2] if storage == "float32" else [0, 1, 3]
        if ncomp == 3 and not z_halo:
            snaps = data.to(device=self.device, dtype=torch.float32).contiguous()
        else:
            snaps = torch.zeros((T, Nx, Ny, Nz + z_halo, ncomp), dtype=tdt, device=self.device)
            src = data.to(device=self.device, dtype=torch.float32)
            for c, slot in enumerate(self._slots):
                snaps[:, :, :, :Nz, slot] = src[..., c].to(tdt)
                if z_halo:
                    edge = 0 if mode == "periodic" else Nz - 1
                    snaps[:, :, :, Nz, slot] = src[:, :, :, edge, c].to(tdt)
            del src
        self.z_halo = z_halo
        self._snaps = snaps
        self._times_dev = torch.from_numpy(tt.copy()).to(self.device)
        self.times = tt
        self.t_min, self.t_max = float(tt[0]), float(tt[-1])
        d = L.GridField()
        d.data = snaps.data_ptr()
        d.times = self._times_dev.data_ptr()
        d.n_snap, d.nx, d.ny, d.nz = T, Nx, Ny, Nz
        d.dtype, d.ncomp = dt_code, ncomp
        d.z_halo = z_halo
        d.origin[:] = [F32(gx[0]), F32(gy[0]), F32(gz[0])]
        d.spacing[:] = [self.dx, self.dy, self.dz]
        d.bounds_mode = L.JT_BOUNDS[mode]
        d.time_interp = L.JT_TIME.get(interp, 0)
        d.time_extrap = L.JT_EXTRAP[extra]
        self._pairs = None
        self._desc = d
        self._build_pair_records(pair_records)

    def _build_pair_records(self, want):
        """Consecutive-pair records (jt_grid_pair_records): for each snapshot
        pair (s, s+1) and node, both snapshots' z-pairs in one 48-byte record,
        so the tracking fast path gathers a cell row with three 16-byte loads
        instead of four.  They cost (T-1) * Nx * Ny * Nz * 48 bytes of device
        memory.  want None: built for float32 'halo' fields with periodic
        bounds (the fast path's fields) when they take at most a quarter of
        the free device memory; True: built for any float32 'halo' field;
        False: not built."""
        applicable = (self.storage == "float32" and self._desc.ncomp == 4 and self.z_halo == 1 and self.T >= 2)
        if want is None:
            nbytes = (self.T - 1) * self.Nx * self.Ny * self.Nz * 48
            want = applicable and self.bounds_handling == "periodic" and \
                nbytes <= torch.cuda.mem_get_info(self.device)[0] // 4
        elif want and not applicable:
            raise ValueError("pair_records needs float32 storage, layout 'halo' or 'halo4' and at least two snapshots")
        if not want:
            self._desc.pair_data = None
            return
        lib = L.load()
        pairs = torch.empty((self.T - 1, self.Nx, self.Ny, self.Nz, 12), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            L.check(lib.jt_grid_pair_records(C.byref(self._desc), L.ptr(pairs), L.stream_handle()),
                    "jt_grid_pair_records")
        self._pairs = pairs
        self._desc.pair_data = pairs.data_ptr()

    # -- kernel descriptor ---------------------------------------------------
    def descriptor(self) -> L.GridField:
        if self.interpolation == "cubic":
            raise NotImplementedError("cubic interpolation not supported with JAX")
        return self._desc

    @property
    def snapshot_bytes_per_node(self) -> int:
        return self._snaps.element_size() * 3

    def _sample(self, positions, t):
        lib = L.load()
        x, is_torch = _ensure_positions(positions, self.device)
        out = torch.empty_like(x)
        L.check(lib.jt_grid_sample(C.byref(self.descriptor()), L.ptr(x), x.shape[0], C.c_float(F32(t)),
                                   L.ptr(out), L.stream_handle()), "jt_grid_sample")
        if is_torch:
            return out
        return out.cpu().numpy()

    def get_spatial_bounds(self):
        return (np.array([self.x_min, self.y_min, self.z_min], F32),
                np.array([self.x_max, self.y_max, self.z_max], F32))


class StructuredGridSampler(_GridSeries):
    """structured.py:65-357: time-independent trilinear sampler.

    ``sample(positions)`` returns trilinear values with 'clamp' | 'zero' |
    'nan' | 'periodic' bounds handling.  Stored as a one-snapshot series
    whose temporal blend is an exact copy."""

    def __init__(self, velocity_data, grid_x, grid_y, grid_z, bounds_handling: str = "clamp",
                 grid_meta=None, storage: str = "float32", layout: str = "halo", device=None):
        v = velocity_data if isinstance(velocity_data, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(np.asarray(velocity_data, dtype=F32)))
        if v.ndim != 4:
            raise ValueError(f"velocity_data must be 4D (Nx,Ny,Nz,3), got {tuple(v.shape)}")
        self._setup(v[None], np.zeros(1, F32), grid_x, grid_y, grid_z, bounds_handling, "linear",
                    "constant", storage, layout, device)
        self.grid_meta = grid_meta

    def sample(self, positions):
        return self._sample(positions, 0.0)


class StructuredTimeSeriesField(_GridSeries):
    """Time-dependent structured field: temporal blend (time_series.py:155-245)
    of (T, Nx, Ny, Nz, 3) snapshots composed with trilinear interpolation in
    space (structured.py:167-250).

    ``sample_at_positions(positions, t)`` is the reference composition
    ``_trilinear_jax(_sample_all_nodes_at_time_jax(t))``; the kernel blends the
    8 gathered corners with the identical per-node arithmetic."""

    def __init__(self, data, times, grid_x, grid_y, grid_z, bounds_handling: str = "clamp",
                 interpolation: str = "linear", extrapolation: str = "constant",
                 storage: str = "float32", layout: str = "halo", device=None,
                 _source_info: Optional[dict] = None, pair_records: Optional[bool] = None):
        v = data if isinstance(data, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(np.asarray(data, dtype=F32)))
        if v.ndim == 3:
            # TimeSeriesField layout (T, N, 3) with N = Nx*Ny*Nz in C order
            n = (len(grid_x), len(grid_y), len(grid_z))
            if v.shape[1] != n[0] * n[1] * n[2]:
                raise ValueError(f"data has {v.shape[1]} nodes, grid has {n[0] * n[1] * n[2]}")
            v = v.reshape(v.shape[0], *n, 3)
        if v.ndim != 5:
            raise ValueError(f"data must be (T,Nx,Ny,Nz,3) or (T,N,3), got {tuple(v.shape)}")
        self._setup(v, times, grid_x, grid_y, grid_z, bounds_handling, interpolation,
                    extrapolation, storage, layout, device, pair_records)
        self._source_info = _source_info or {}

    def sample_at_positions(self, positions, t: float):
        return self._sample(positions, t)

    def __call__(self, a, b):
        if np.ndim(a) == 0:
            return self.sample_at_positions(b, float(a))
        return self.sample_at_positions(a, float(b))

    def get_time_bounds(self):
        return (self.t_min, self.t_max)

    def with_storage(self, storage: str, layout: str = "halo",
                     pair_records: Optional[bool] = None) -> "StructuredTimeSeriesField":
        """Same field, snapshots re-stored as ``storage`` (RNE from float32)."""
        src = self._snaps[:, :, :, :self.Nz, self._slots].to(torch.float32)
        return StructuredTimeSeriesField(src, self.times, self.grid_x, self.grid_y, self.grid_z,
                                         self.bounds_handling, self.interpolation,
                                         self.extrapolation, storage, layout, self.device,
                                         pair_records=pair_records)


class OctreeFEMTimeSeriesFieldOptimized:
    """octree_fem_time_series_optimized.py:17-127: nodal snapshots on a
    tetrahedral mesh, located through the element octree of
    build_octree_mesh_optimized (built on the GPU by default, build_on='host'
    for the library's C++ host build; arrays identical to the reference's
    either way) and interpolated with P1 shape functions."""

    kind = "octree"

    def __init__(self, data, times, positions, connectivity, interpolation: str = "linear",
                 extrapolation: str = "constant", _source_info: Optional[dict] = None,
                 max_elements_per_leaf: int = 32, max_depth: int = 12, device=None, build_on: str = "device",
                 element_records: bool = True):
        lib = L.load()
        self.device = torch.device(device) if device is not None else L.device()
        data = np.ascontiguousarray(np.asarray(data.cpu() if isinstance(data, torch.Tensor) else data, F32))
        times = np.asarray(times.cpu() if isinstance(times, torch.Tensor) else times, F32).reshape(-1)
        pts = np.ascontiguousarray(np.asarray(positions.cpu() if isinstance(positions, torch.Tensor)
                                              else positions, F32))
        conn = np.ascontiguousarray(np.asarray(connectivity.cpu() if isinstance(connectivity, torch.Tensor)
                                               else connectivity, np.int32))
        if data.ndim != 3 or data.shape[2] != 3:                               # time_series.py:120-128
            raise ValueError(f"data must have 3 dims (T,N,3), got {data.shape}")
        T, N, _ = data.shape
        if times.shape != (T,):
            raise ValueError(f"times shape {times.shape} != (T,) with T={T}")
        if pts.shape != (N, 3):
            raise ValueError(f"positions shape {pts.shape} != (N,3) with N={N}")
        if conn.ndim != 2 or conn.shape[1] != 4:
            raise ValueError("connectivity must be (M,4)")
        interp, extra = interpolation.lower(), extrapolation.lower()
        if interp not in ("nearest", "linear", "cubic"):
            raise ValueError(f"Unsupported interpolation: {interpolation}")
        if extra not in L.JT_EXTRAP:
            raise ValueError(f"Unsupported extrapolation: {extrapolation}")
        self.interpolation, self.extrapolation = interp, extra
        self.data, self.times, self.positions, self.connectivity = data, times, pts, conn
        self.T, self.N = T, N
        self.t_min, self.t_max = float(times[0]), float(times[-1])
        self._source_info = _source_info or {}
        self.storage = "float32"
        self.max_depth = int(max_depth)
        dev = self.device
        M = conn.shape[0]
        if build_on == "device":
            self._dev = self._build_on_device(pts, conn, int(max_elements_per_leaf), int(max_depth))
            self._octree_host = None
        elif build_on == "host":
            tree = self._build_on_host(pts, conn, int(max_elements_per_leaf), int(max_depth))
            self._octree_host = tree
            self._dev = {k: torch.from_numpy(v).to(dev) for k, v in tree.items()}
        else:
            raise ValueError(f"build_on must be 'device' or 'host', got {build_on!r}")
        n_nodes, max_elems = self._dev["nodes_elements"].shape
        self._dev.pop("element_centroids", None)
        self._dev["points"] = torch.from_numpy(pts).to(dev)
        self._dev["conn"] = torch.from_numpy(conn).to(dev)
        self._dev["data"] = torch.from_numpy(data).to(dev)
        self._dev["times"] = torch.from_numpy(times.copy()).to(dev)
        d = L.OctreeField()
        for k in ("points", "conn", "element_bounds", "nodes_min", "nodes_max", "nodes_elements",
                  "nodes_elem_counts", "nodes_children", "nodes_is_leaf", "data", "times"):
            setattr(d, k, self._dev[k].data_ptr())
        d.n_points, d.n_elems, d.n_nodes = N, M, n_nodes
        d.max_elems, d.max_depth = max_elems, int(max_depth)
        d.n_snap = T
        d.time_interp = L.JT_TIME.get(interp, 0)
        d.time_extrap = L.JT_EXTRAP[extra]
        self._desc = d
        # per-element records (box, vertices, point ids; 96 B per element)
        # for the tracking kernels' leaf scan
        self._erec = None
        if element_records:
            lib = L.load()
            self._erec = torch.empty((M, 24), dtype=torch.float32, device=dev)
            with torch.cuda.device(dev):
                L.check(lib.jt_octree_element_records(C.byref(d), L.ptr(self._erec), L.stream_handle()),
                        "jt_octree_element_records")
            d.element_records = self._erec.data_ptr()

    def _build_on_host(self, pts, conn, max_per_leaf, max_depth):
        """jt_octree_build: the C++ host build, exported as NumPy arrays."""
        lib = L.load()
        h = C.c_void_p()
        L.check(lib.jt_octree_build(pts.ctypes.data, pts.shape[0], conn.ctypes.data, conn.shape[0], max_per_leaf,
                                    max_depth, C.byref(h)), "jt_octree_build")
        try:
            nn, me, dep = C.c_int64(), C.c_int32(), C.c_int32()
            L.check(lib.jt_octree_info(h, C.byref(nn), C.byref(me), C.byref(dep)), "jt_octree_info")
            n_nodes, max_elems, M = int(nn.value), int(me.value), conn.shape[0]
            tree = {
                "nodes_min": np.empty((n_nodes, 3), F32), "nodes_max": np.empty((n_nodes, 3), F32),
                "nodes_elements": np.empty((n_nodes, max_elems), np.int32),
                "nodes_elem_counts": np.empty(n_nodes, np.int32),
                "nodes_children": np.empty((n_nodes, 8), np.int32),
                "nodes_is_leaf": np.empty(n_nodes, np.uint8),
                "element_bounds": np.empty((M, 2, 3), F32), "element_centroids": np.empty((M, 3), F32),
            }
            L.check(lib.jt_octree_export(h, *[tree[k].ctypes.data for k in (
                "nodes_min", "nodes_max", "nodes_elements", "nodes_elem_counts", "nodes_children",
                "nodes_is_leaf", "element_bounds", "element_centroids")]), "jt_octree_export")
        finally:
            lib.jt_octree_free(h)
        self.tree_depth = int(dep.value)
        return tree

    def _build_on_device(self, pts, conn, max_per_leaf, max_depth):
        """jt_octree_build_device: the level-by-level GPU build; the arrays
        stay on the device.  The node capacity starts at 4 M / max_per_leaf
        and grows if the tree needs more."""
        lib = L.load()
        dev = self.device
        M = conn.shape[0]
        pts_d = torch.from_numpy(pts).to(dev)
        conn_d = torch.from_numpy(conn).to(dev)
        eb = torch.empty((M, 2, 3), dtype=torch.float32, device=dev)
        cent = torch.empty((M, 3), dtype=torch.float32, device=dev)
        cap = max(1024, 4 * M // max(max_per_leaf, 1) + 64)
        nn, me, dep = C.c_int64(), C.c_int32(), C.c_int32()
        while True:
            wsb = int(lib.jt_octree_build_device_workspace_size(M, cap))
            if wsb == 0:
                raise MemoryError(f"octree node capacity {cap} out of range")
            ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
            rc = lib.jt_octree_build_device(L.ptr(pts_d), pts.shape[0], L.ptr(conn_d), M, max_per_leaf, max_depth,
                                            cap, L.ptr(eb), L.ptr(cent), L.ptr(ws), wsb, C.byref(nn), C.byref(me),
                                            C.byref(dep), L.stream_handle())
            if rc == L.JT_ENOMEM and b"max_nodes" in lib.jt_last_error():
                cap *= 4
                continue
            L.check(rc, "jt_octree_build_device")
            break
        n_nodes, max_elems = int(nn.value), int(me.value)
        out = {"nodes_min": torch.empty((n_nodes, 3), dtype=torch.float32, device=dev),
               "nodes_max": torch.empty((n_nodes, 3), dtype=torch.float32, device=dev),
               "nodes_elements": torch.empty((n_nodes, max_elems), dtype=torch.int32, device=dev),
               "nodes_elem_counts": torch.empty(n_nodes, dtype=torch.int32, device=dev),
               "nodes_children": torch.empty((n_nodes, 8), dtype=torch.int32, device=dev),
               "nodes_is_leaf": torch.empty(n_nodes, dtype=torch.uint8, device=dev)}
        L.check(lib.jt_octree_export_device(L.ptr(ws), wsb, M, cap, n_nodes, max_elems,
                                            *[L.ptr(out[k]) for k in ("nodes_min", "nodes_max", "nodes_elements",
                                                                      "nodes_elem_counts", "nodes_children",
                                                                      "nodes_is_leaf")], L.stream_handle()),
                "jt_octree_export_device")
        out["element_bounds"] = eb
        out["element_centroids"] = cent
        self._centroids_dev = cent
        self.tree_depth = int(dep.value)
        del ws
        return out

    @property
    def octree(self):
        """The flattened octree arrays (build_octree_mesh_optimized's, :178-196)
        as NumPy arrays (copied from the device on first use)."""
        if self._octree_host is None:
            tree = {k: v.cpu().numpy() for k, v in self._dev.items()
                    if k in ("nodes_min", "nodes_max", "nodes_elements", "nodes_elem_counts", "nodes_children",
                             "nodes_is_leaf", "element_bounds")}
            tree["element_centroids"] = self._centroids_dev.cpu().numpy()
            self._octree_host = tree
        return self._octree_host

    def descriptor(self):
        if self.interpolation == "cubic":
            raise NotImplementedError("cubic interpolation not supported with JAX")
        return self._desc

    def sample_at_positions(self, query_positions, t: float, return_found: bool = False):
        lib = L.load()
        x, is_torch = _ensure_positions(query_positions, self.device)
        out = torch.empty_like(x)
        found = torch.empty(x.shape[0], dtype=torch.int32, device=self.device)
        L.check(lib.jt_octree_sample(C.byref(self.descriptor()), L.ptr(x), x.shape[0], C.c_float(F32(t)),
                                     L.ptr(out), L.ptr(found), L.stream_handle()), "jt_octree_sample")
        if not is_torch:
            out, found = out.cpu().numpy(), found.cpu().numpy()
        return (out, found) if return_found else out

    _sample = sample_at_positions

    def get_time_bounds(self):
        return (self.t_min, self.t_max)


class TimeSeriesField:
    """time_series.py:82-473: (T, N, 3) snapshots at fixed node positions,
    blended in time and sampled at the nearest node (brute force over all
    nodes, as the reference; ties go to the lowest node index)."""

    kind = "nodes"

    def __init__(self, data, times, positions, interpolation: str = "linear", extrapolation: str = "constant",
                 _source_info: Optional[dict] = None, device=None):
        self.device = torch.device(device) if device is not None else L.device()
        d = np.asarray(data.cpu() if isinstance(data, torch.Tensor) else data, F32)
        t = np.asarray(times.cpu() if isinstance(times, torch.Tensor) else times, F32).reshape(-1)
        p = np.asarray(positions.cpu() if isinstance(positions, torch.Tensor) else positions, F32)
        if p.ndim == 2 and p.shape[1] == 2:
            p = np.concatenate([p, np.zeros((p.shape[0], 1), F32)], 1)
        if d.ndim != 3:                                                       # :120-128
            raise ValueError(f"data must have 3 dims (T,N,3), got {d.shape}")
        T, N, D = d.shape
        if D != 3:
            raise ValueError(f"data last dimension must be 3, got {D}")
        if t.shape != (T,):
            raise ValueError(f"times shape {t.shape} != (T,) with T={T}")
        if p.shape != (N, 3):
            raise ValueError(f"positions shape {p.shape} != (N,3) with N={N}")
        interp, extra = interpolation.lower(), extrapolation.lower()
        if interp not in ("nearest", "linear", "cubic"):
            raise ValueError(f"Unsupported interpolation: {interpolation}")
        if extra not in L.JT_EXTRAP:
            raise ValueError(f"Unsupported extrapolation: {extrapolation}")
        self.data, self.times, self.positions = d, t, p
        self.interpolation, self.extrapolation = interp, extra
        self.T, self.N = T, N
        self.t_min, self.t_max = float(t[0]), float(t[-1])
        self._source_info = _source_info or {}
        self.storage = "float32"
        self._dev = {"pos": torch.from_numpy(np.ascontiguousarray(p)).to(self.device),
                     "data": torch.from_numpy(np.ascontiguousarray(d)).to(self.device),
                     "times": torch.from_numpy(t.copy()).to(self.device)}
        desc = L.NodesField()
        desc.node_pos = self._dev["pos"].data_ptr()
        desc.n_nodes = N
        desc.data = self._dev["data"].data_ptr()
        desc.times = self._dev["times"].data_ptr()
        desc.n_snap = T
        desc.time_interp = L.JT_TIME.get(interp, 0)
        desc.time_extrap = L.JT_EXTRAP[extra]
        self._desc = desc

    def descriptor(self):
        if self.interpolation == "cubic":
            raise NotImplementedError("cubic interpolation not supported with JAX")
        return self._desc

    def sample_at_positions(self, query_positions, t: float):
        lib = L.load()
        x, is_torch = _ensure_positions(query_positions, self.device)
        out = torch.empty_like(x)
        L.check(lib.jt_nodes_sample(C.byref(self.descriptor()), L.ptr(x), x.shape[0], C.c_float(F32(t)),
                                    L.ptr(out), None, L.stream_handle()), "jt_nodes_sample")
        return out if is_torch else out.cpu().numpy()

    _sample = sample_at_positions

    def _parse(self, a, b):
        if np.ndim(a) == 0 and np.ndim(b) != 0:
            return float(a), b
        if np.ndim(b) == 0 and np.ndim(a) != 0:
            return float(b), a
        raise TypeError("Cannot determine (t, positions) from arguments")

    def __call__(self, a, b):
        t, pos = self._parse(a, b)
        return self.sample_at_positions(pos, t)

    evaluate = __call__

    def get_time_bounds(self):
        return (self.t_min, self.t_max)

    def get_spatial_bounds(self):
        return self.positions.min(axis=0), self.positions.max(axis=0)
