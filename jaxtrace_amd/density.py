"""Density estimation on the GPU: ``HashGridNeighbors``, ``SPHDensityEstimator``
and ``KDEEstimator`` with the reference's constructors and evaluation APIs.

* HashGridNeighbors     -- jaxtrace/density/neighbors.py:39-168
* SPHDensityEstimator   -- jaxtrace/density/sph.py:42-380
* KDEEstimator          -- jaxtrace/density/kde.py:43-282 (bandwidth rules
  kernels.py:30-70)

``key_mode`` selects the hash-key packing: 'x32' is what the reference
computes as shipped (JAX without x64: the cell's x index drops out of the
key), 'x64' the packing its source intends.  Host-side preparation (slab
extraction, grids, bandwidth, normalisation) is the reference's own NumPy
code path; neighbour build, queries and kernel sums run in HIP.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Tuple, Union

import numpy as np
import torch

from . import _lib as L

F32 = np.float32
_KEY = {"x32": L.JT_KEY_X32, "x64": L.JT_KEY_X64}


def _to_dev(a, dev, dtype=torch.float32):
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=dev, dtype=dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a))).to(device=dev, dtype=dtype)


class HashGridNeighbors:
    """neighbors.py:39-168: CSR uniform hash grid with fixed-size range queries."""

    def __init__(self, positions, cell_size: float, per_cell_cap: int = 64, key_mode: str = "x32",
                 device=None):
        lib = L.load()
        self.device = torch.device(device) if device is not None else L.device()
        P = _to_dev(positions, self.device)
        if P.ndim != 2 or P.shape[1] not in (2, 3):
            raise ValueError("positions must be (N,2) or (N,3)")
        if key_mode not in _KEY:
            raise ValueError(f"key_mode must be 'x32' or 'x64', got {key_mode!r}")
        self.positions = P
        self.D = int(P.shape[1])
        self.N = int(P.shape[0])
        self.cell_size = float(cell_size)
        self.per_cell_cap = int(per_cell_cap)
        self.key_mode = key_mode
        if self.N < 1:
            raise ValueError("HashGridNeighbors needs at least one point")
        mm = torch.empty(2 * self.D, dtype=torch.float32, device=self.device)
        ws = torch.empty(L.JT_BOUNDS_WS_BYTES, dtype=torch.uint8, device=self.device)
        L.check(lib.jt_points_bounds(L.ptr(P), self.N, self.D, L.ptr(mm), L.ptr(ws), L.JT_BOUNDS_WS_BYTES,
                                     L.stream_handle()),
                "jt_points_bounds")
        mm = mm.cpu().numpy()
        self.origin = mm[:self.D].astype(F32)                              # :68
        pmax = mm[self.D:].astype(F32)
        self.inv = F32(F32(1.0) / F32(max(self.cell_size, 1e-12)))        # :67
        cmax = np.floor((pmax - self.origin) * self.inv).astype(np.int64)  # :70 for the max point
        self.Cmin = np.zeros(self.D, np.int32)
        dims = [int(c) + 1 for c in cmax]
        d = L.HashGridDesc()
        d.dim = self.D
        d.key_mode = _KEY[key_mode]
        d.origin[:] = [float(v) for v in self.origin] + [0.0] * (3 - self.D)
        d.inv = self.inv
        d.dims[:] = dims + [1] * (3 - self.D)
        d.n_points = self.N
        self.n_cells = int(lib.jt_hashgrid_num_cells(C.byref(d)))
        self.pt_indices_sorted = torch.empty(self.N, dtype=torch.int32, device=self.device)
        self.sorted_positions = torch.empty_like(P)
        self.cell_start = torch.empty(self.n_cells, dtype=torch.int32, device=self.device)
        self.cell_end = torch.empty(self.n_cells, dtype=torch.int32, device=self.device)
        d.order = self.pt_indices_sorted.data_ptr()
        d.sorted_pos = self.sorted_positions.data_ptr()
        d.cell_start = self.cell_start.data_ptr()
        d.cell_end = self.cell_end.data_ptr()
        wsb = int(lib.jt_hashgrid_workspace_size(self.N))
        ws = torch.empty(wsb, dtype=torch.uint8, device=self.device)
        L.check(lib.jt_hashgrid_build(C.byref(d), L.ptr(P), L.ptr(ws), wsb, L.stream_handle()),
                "jt_hashgrid_build")
        self._desc = d
        self.dims = tuple(dims)

    def query_many(self, Q, radius: float, max_neighbors: int) -> Tuple:
        """neighbors.py:109-168 -> (idx int32 (M,K), dist f32 (M,K))."""
        lib = L.load()
        is_torch = isinstance(Q, torch.Tensor)
        q = _to_dev(Q, self.device)
        if q.ndim != 2 or q.shape[1] != self.D:
            raise ValueError(f"queries must be (M,{self.D})")
        M, K = int(q.shape[0]), int(max_neighbors)
        idx = torch.empty((M, K), dtype=torch.int32, device=self.device)
        dist = torch.empty((M, K), dtype=torch.float32, device=self.device)
        L.check(lib.jt_hashgrid_query(C.byref(self._desc), L.ptr(q), M, C.c_float(F32(radius)),
                                      self.per_cell_cap, K, L.ptr(idx), L.ptr(dist), L.stream_handle()),
                "jt_hashgrid_query")
        if is_torch:
            return idx, dist
        return idx.cpu().numpy(), dist.cpu().numpy()


class SPHDensityEstimator:
    """sph.py:42-380 (hash-grid path on the GPU)."""

    def __init__(self, positions, dimensions: str = "2d", plane: str = "xy", position: float = 0.0,
                 slab_thickness: float = 0.1, smoothing_length: Optional[float] = None,
                 kernel_type: str = "cubic_spline", adaptive: bool = False, n_neighbors: int = 32,
                 masses=None, max_particles_for_adaptive: int = 5000, neighbor_search: str = "hashgrid",
                 per_cell_cap: int = 64, max_neighbors: int = 512, normalize: bool = True,
                 resolution: Union[int, Tuple[int, int, int]] = 100, bounds=None,
                 key_mode: str = "x32", device=None, tiled_queries: bool = True):
        self.device = torch.device(device) if device is not None else L.device()
        self.dimensions, self.plane, self.position = dimensions, plane, position
        self.slab_thickness, self.kernel_type = slab_thickness, kernel_type
        self.adaptive, self.n_neighbors = adaptive, n_neighbors
        self.max_particles_for_adaptive = max_particles_for_adaptive
        self.neighbor_search, self.per_cell_cap = neighbor_search, int(per_cell_cap)
        self.max_neighbors, self.normalize = int(max_neighbors), normalize
        self.resolution, self.bounds, self.key_mode = resolution, bounds, key_mode
        # x32 keys with uniform masses and smoothing: evaluate queries in tiles
        # that share a (y, z) cell column (jt_sph_eval_tiled); False keeps one
        # wave per query (jt_sph_eval)
        self.tiled_queries = bool(tiled_queries)
        if kernel_type not in L.JT_SPH:
            raise ValueError(f"Unknown kernel type: {kernel_type}")
        if neighbor_search != "hashgrid":
            raise TypeError("neighbor_search must be 'hashgrid' on the GPU path")
        self._P_host = None
        if (isinstance(positions, torch.Tensor) and positions.device.type == "cuda"
                and self.dimensions.lower() != "2d"):
            # device positions stay on the device (the host copy is made on
            # first use of .P)
            if positions.ndim != 2 or positions.shape[1] != 3:
                raise ValueError("3D SPH requires (N,3) positions")
            self._P_dev = positions.detach().to(device=self.device, dtype=torch.float32).contiguous()
            self.D = 3
            self.N = int(self._P_dev.shape[0])
        else:
            P3 = positions.detach().cpu().numpy() if isinstance(positions, torch.Tensor) else np.asarray(positions)
            if self.dimensions.lower() == "2d":                               # :112-117
                if P3.shape[1] != 3:
                    raise ValueError("2D SPH requires (N,3) positions for slicing")
                self._P_host = self._extract_2d(P3)
                self.D = 2
            else:
                if P3.shape[1] != 3:
                    raise ValueError("3D SPH requires (N,3) positions")
                self._P_host = np.asarray(P3, dtype=F32)
                self.D = 3
            self.N = int(self._P_host.shape[0])
            self._P_dev = None
        if self.N < 2:
            raise ValueError("At least two particles are required")
        if masses is None:
            self.masses = None
        else:
            self.masses = np.asarray(masses, dtype=F32)
            if self.masses.shape[0] != self.N:
                raise ValueError("masses must be length N")
        if self._P_dev is None:
            self._P_dev = _to_dev(self._P_host, self.device)
        if smoothing_length is not None:                                      # :137-140
            self.smoothing_lengths = np.full(self.N, F32(float(smoothing_length)), F32)
        else:
            self.smoothing_lengths = self._calculate_smoothing_lengths()
        self._mean_h = float(np.mean(self.smoothing_lengths))                # :142
        self._hash = HashGridNeighbors(self._P_dev, max(self._mean_h, 1e-6), self.per_cell_cap,
                                       key_mode, self.device)                 # :148-152
        self._neighbor_backend = "hashgrid"
        if smoothing_length is not None:                                      # every h_i is that value
            uniform_h = bool(F32(float(smoothing_length)) == F32(self._mean_h))
        else:
            uniform_h = bool(np.all(self.smoothing_lengths == F32(self._mean_h)))
        self._h_dev = None if uniform_h else _to_dev(self.smoothing_lengths, self.device)
        self._m_dev = None if self.masses is None else _to_dev(self.masses, self.device)

    @property
    def P(self) -> np.ndarray:
        """The particle positions on the host, (N, D) float32."""
        if self._P_host is None:
            self._P_host = self._P_dev.cpu().numpy()
        return self._P_host

    def _extract_2d(self, P3):
        """sph.py:164-177."""
        axis_map = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}
        plane = self.plane.lower()
        if plane not in axis_map:
            raise ValueError("plane must be one of 'xy','xz','yz'")
        i, j, k = axis_map[plane]
        mask = np.abs(P3[:, k] - float(self.position)) <= (float(self.slab_thickness) * 0.5)
        pts2 = P3[mask][:, [i, j]]
        if pts2.size == 0:
            pts2 = P3[:, [i, j]]
        return pts2.astype(F32, copy=False)

    def _calculate_smoothing_lengths(self):
        """sph.py:181-212."""
        if self.adaptive and self.N <= int(self.max_particles_for_adaptive):
            k = int(min(self.n_neighbors, max(self.N - 1, 1)))
            P = self._P_dev
            d = torch.cdist(P, P)
            kth = torch.kthvalue(d, k + 1, dim=1).values
            return (F32(1.2) * kth.cpu().numpy().astype(F32)).astype(F32)
        n_per_dim = self.N ** (1.0 / self.D)
        pmin, pmax = self.P.min(axis=0), self.P.max(axis=0)
        domain_size = F32(np.mean(pmax - pmin))
        h = F32(domain_size / F32(n_per_dim) * F32(2.0))
        h = F32(min(max(h, F32(0.01)), F32(domain_size * F32(0.1))))
        return np.full(self.N, h, F32)

    def evaluate(self, eval_points) -> np.ndarray:
        """sph.py:333-380."""
        lib = L.load()
        Q_np = eval_points.detach().cpu().numpy() if isinstance(eval_points, torch.Tensor) else eval_points
        Q_np = np.asarray(Q_np, dtype=F32)
        if Q_np.ndim != 2:
            raise ValueError(f"eval_points must be 2D, got {Q_np.shape}")
        if self.D == 2:
            if Q_np.shape[1] == 3:
                i, j, _ = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}[self.plane.lower()]
                Q_np = Q_np[:, [i, j]]
            elif Q_np.shape[1] != 2:
                raise ValueError(f"2D SPH expects (M,2) or (M,3) points, got {Q_np.shape}")
        elif Q_np.shape[1] != 3:
            raise ValueError(f"3D SPH expects (M,3) points, got {Q_np.shape}")
        Q = _to_dev(np.ascontiguousarray(Q_np), self.device)
        M = int(Q.shape[0])
        rho = torch.empty(M, dtype=torch.float32, device=self.device)
        radius = F32(2.0 * self._mean_h)                                       # :273
        if (self.tiled_queries and self.key_mode == "x32" and self._m_dev is None and self._h_dev is None
                and self.per_cell_cap <= 64):
            desc = C.byref(self._hash._desc)
            ws_bytes = int(lib.jt_sph_tiled_workspace_size(desc, M))
            ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=self.device)
            L.check(lib.jt_sph_eval_tiled(desc, L.ptr(Q), M, C.c_float(radius), self.per_cell_cap,
                                          self.max_neighbors, C.c_float(F32(self._mean_h)),
                                          L.JT_SPH[self.kernel_type], L.ptr(rho), L.ptr(ws), ws_bytes,
                                          L.stream_handle()), "jt_sph_eval_tiled")
        else:
            L.check(lib.jt_sph_eval(C.byref(self._hash._desc), L.ptr(Q), M, C.c_float(radius),
                                    self.per_cell_cap, self.max_neighbors, L.ptr(self._m_dev),
                                    L.ptr(self._h_dev), C.c_float(F32(self._mean_h)),
                                    L.JT_SPH[self.kernel_type], L.ptr(rho), L.stream_handle()),
                    "jt_sph_eval")
        rho = rho.cpu().numpy()
        if self.normalize and M > 0:                                          # :365-378
            if self.D == 2:
                xmin, xmax = float(Q_np[:, 0].min()), float(Q_np[:, 0].max())
                ymin, ymax = float(Q_np[:, 1].min()), float(Q_np[:, 1].max())
                area = max((xmax - xmin) * (ymax - ymin), 1e-12)
                rho = rho / F32(self.N / area)
            else:
                xmin, xmax = float(Q_np[:, 0].min()), float(Q_np[:, 0].max())
                ymin, ymax = float(Q_np[:, 1].min()), float(Q_np[:, 1].max())
                zmin, zmax = float(Q_np[:, 2].min()), float(Q_np[:, 2].max())
                vol = max((xmax - xmin) * (ymax - ymin) * (zmax - zmin), 1e-12)
                rho = rho / F32(self.N / vol)
        return np.asarray(rho, dtype=F32)


def scott_bandwidth(data) -> float:
    """kernels.py:30-47."""
    x = np.asarray(data)
    n, d = x.shape
    sigma = np.std(x, axis=0, ddof=1).mean()
    return sigma * n ** (-1.0 / (d + 4.0))


def silverman_bandwidth(data) -> float:
    """kernels.py:50-70."""
    x = np.asarray(data)
    n, d = x.shape
    sigma = np.std(x, axis=0, ddof=1).mean()
    factor = (4.0 / (d + 2.0)) ** (1.0 / (d + 4.0))
    return factor * sigma * n ** (-1.0 / (d + 4.0))


class KDEEstimator:
    """kde.py:43-282: Gaussian KDE with Scott/Silverman bandwidth.

    ``cutoff_sigmas=None`` evaluates the reference's exhaustive all-pairs
    sum.  A number c makes ``evaluate_3d`` sum, for each grid point, every
    particle within c*h of it (and some farther ones: those in the hash-grid
    cells, of size c*h/2, around the point's block of the grid); every
    dropped term is below (2pi)^(-3/2) exp(-c^2/2), and so is their mean.
    ``evaluate_2d`` and ``evaluate_3d`` without a cut-off sum every pair;
    with ``matrix_cores=True`` (the default) they do so as products of
    per-particle x, y (and z) factors on the bf16 matrix cores (each term to
    about 2^-17 relative), else one thread per grid point over all particles."""

    def __init__(self, positions, bandwidth: Optional[float] = None, bandwidth_rule: str = "scott",
                 normalize: bool = True, plane: str = "xy", slab_position: float = 0.0,
                 slab_thickness: float = 0.1, resolution: Union[int, Tuple[int, int, int]] = 100,
                 bounds=None, cutoff_sigmas: Optional[float] = None, device=None, matrix_cores: bool = True):
        self.device = torch.device(device) if device is not None else L.device()
        P = positions.detach().cpu().numpy() if isinstance(positions, torch.Tensor) else positions
        P = np.asarray(P, dtype=F32)
        if P.ndim != 2 or P.shape[1] not in (2, 3):
            raise ValueError("positions must be (N,2) or (N,3)")
        self.N, self.D = int(P.shape[0]), int(P.shape[1])
        self.P = P
        self.bandwidth, self.bandwidth_rule, self.normalize = bandwidth, bandwidth_rule, normalize
        self.plane, self.slab_position, self.slab_thickness = plane, slab_position, slab_thickness
        self.resolution, self.bounds, self.cutoff_sigmas = resolution, bounds, cutoff_sigmas
        self.matrix_cores = bool(matrix_cores)
        if bandwidth is None:                                                  # :81-97
            if bandwidth_rule == "scott":
                self.h = float(scott_bandwidth(P))
            elif bandwidth_rule == "silverman":
                self.h = float(silverman_bandwidth(P))
            else:
                raise ValueError("bandwidth_rule must be 'scott' or 'silverman' if bandwidth is None")
        else:
            self.h = float(bandwidth)
        self._P_dev = {}

    def _axes_for_plane(self):
        axis_map = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}
        plane = self.plane.lower()
        if plane not in axis_map:
            raise ValueError("plane must be one of 'xy','xz','yz'")
        return axis_map[plane]

    def _slice_2d_points(self):
        """kde.py:112-124."""
        if self.D == 2:
            return self.P
        i, j, k = self._axes_for_plane()
        mask = np.abs(self.P[:, k] - float(self.slab_position)) <= (float(self.slab_thickness) * 0.5)
        pts2 = self.P[mask][:, [i, j]]
        if pts2.size == 0:
            pts2 = self.P[:, [i, j]]
        return pts2.astype(F32, copy=False)

    def _dev_points(self, pts):
        key = (pts.shape, pts.__array_interface__["data"][0])
        if key not in self._P_dev:
            self._P_dev = {key: _to_dev(np.ascontiguousarray(pts), self.device)}
        return self._P_dev[key]

    def _evaluate_grid(self, grid_pts, pts) -> np.ndarray:
        """kde.py:244-270 (exhaustive)."""
        lib = L.load()
        P = self._dev_points(pts)
        Q = _to_dev(np.ascontiguousarray(np.asarray(grid_pts, dtype=F32)), self.device)
        M, d = int(Q.shape[0]), int(Q.shape[1])
        out = torch.empty(M, dtype=torch.float32, device=self.device)
        ws_bytes = int(lib.jt_kde_eval_workspace_size(int(P.shape[0]), M))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=self.device)
        L.check(lib.jt_kde_eval_ws(L.ptr(P), int(P.shape[0]), d, L.ptr(Q), M, C.c_float(F32(self.h)),
                                   L.ptr(out), L.ptr(ws), ws_bytes, L.stream_handle()), "jt_kde_eval_ws")
        return out.cpu().numpy()

    def evaluate(self, points) -> np.ndarray:
        """kde.py:128-184."""
        Q = points.detach().cpu().numpy() if isinstance(points, torch.Tensor) else points
        Q = np.asarray(Q, dtype=F32)
        if Q.ndim != 2:
            raise ValueError(f"points must be 2D, got {Q.shape}")
        if self.D == 2:
            if Q.shape[1] == 3:
                i, j, _ = self._axes_for_plane()
                grid = Q[:, [i, j]]
            elif Q.shape[1] == 2:
                grid = Q
            else:
                raise ValueError(f"Expected points with 2 or 3 columns, got {Q.shape[1]}")
            pts = self.P
        else:
            if Q.shape[1] == 3:
                grid, pts = Q, self.P
            elif Q.shape[1] == 2:
                pts = self._slice_2d_points()
                grid = Q[:, [0, 1]]
            else:
                raise ValueError(f"Expected points with 2 or 3 columns, got {Q.shape[1]}")
        return self._evaluate_grid(grid, pts)

    def _make_grid_2d(self):
        """kde.py:188-197."""
        pts = self.P[:, :2] if self.D == 2 else self._slice_2d_points()
        if self.bounds is None:
            lo, hi = pts.min(axis=0), pts.max(axis=0)
        else:
            lo = np.array([self.bounds[0][0], self.bounds[1][0]], dtype=float)
            hi = np.array([self.bounds[0][1], self.bounds[1][1]], dtype=float)
        res = self.resolution if isinstance(self.resolution, int) else self.resolution[:2]
        nx = ny = int(res)
        x = np.linspace(lo[0], hi[0], nx, dtype=np.float32)
        y = np.linspace(lo[1], hi[1], ny, dtype=np.float32)
        X, Y = np.meshgrid(x, y, indexing="xy")
        return X, Y, pts

    def _axes_3d(self):
        if self.bounds is None:
            lo, hi = self.P.min(axis=0), self.P.max(axis=0)
        else:
            lo = np.array([b[0] for b in self.bounds], dtype=float)
            hi = np.array([b[1] for b in self.bounds], dtype=float)
        res = self.resolution if not isinstance(self.resolution, int) else (self.resolution,) * 3
        nx, ny, nz = map(int, res)
        return (np.linspace(lo[0], hi[0], nx, dtype=np.float32), np.linspace(lo[1], hi[1], ny, dtype=np.float32),
                np.linspace(lo[2], hi[2], nz, dtype=np.float32))

    def density_grid_2d(self, x, y, pts) -> torch.Tensor:
        """All-pairs densities (kde.py:244-270) on the grid x (nx) by y (ny),
        (ny, nx) on the device, on the matrix cores."""
        lib = L.load()
        P = self._dev_points(np.ascontiguousarray(pts, dtype=F32))
        gx, gy = (_to_dev(np.ascontiguousarray(a, dtype=F32), self.device) for a in (x, y))
        nx, ny, n = len(x), len(y), int(P.shape[0])
        out = torch.empty((ny, nx), dtype=torch.float32, device=self.device)
        ws_bytes = int(lib.jt_kde_grid2d_workspace_size(n, nx, ny))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=self.device)
        L.check(lib.jt_kde_grid2d(L.ptr(P), n, L.ptr(gx), nx, L.ptr(gy), ny, C.c_float(F32(self.h)), L.ptr(out),
                                  L.ptr(ws), ws_bytes, L.stream_handle()), "jt_kde_grid2d")
        return out

    def evaluate_2d(self):
        """kde.py:205-221."""
        X, Y, pts = self._make_grid_2d()
        if self.matrix_cores:
            Z = self.density_grid_2d(X[0], Y[:, 0], pts).cpu().numpy()
        else:
            grid = np.stack([X.ravel(), Y.ravel()], axis=1)
            Z = self._evaluate_grid(grid, pts).reshape(X.shape)
        if self.normalize:
            dx = (X.max() - X.min()) / max(X.shape[1] - 1, 1)
            dy = (Y.max() - Y.min()) / max(Y.shape[0] - 1, 1)
            cell_area = float(dx * dy)
            total = float((Z * cell_area).sum())
            if total > 0:
                Z = Z * (self.N / total)
            domain_area = float((X.max() - X.min()) * (Y.max() - Y.min()))
            avg = self.N / max(domain_area, 1e-12)
            Z = Z / max(avg, 1e-12)
        return X, Y, Z

    def density_grid_3d(self):
        """Un-normalised densities on the evaluate_3d grid, (ny, nx, nz) on the device."""
        lib = L.load()
        x, y, z = self._axes_3d()
        nx, ny, nz = len(x), len(y), len(z)
        if self.D != 3:
            raise ValueError("evaluate_3d needs 3-D positions")
        if self.cutoff_sigmas is None:
            if self.matrix_cores:
                P = self._dev_points(self.P)
                gx, gy, gz = (_to_dev(a, self.device) for a in (x, y, z))
                out = torch.empty((ny, nx, nz), dtype=torch.float32, device=self.device)
                ws_bytes = int(lib.jt_kde_grid3d_workspace_size(self.N, nx, ny, nz))
                ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=self.device)
                L.check(lib.jt_kde_grid3d(L.ptr(P), self.N, L.ptr(gx), nx, L.ptr(gy), ny, L.ptr(gz), nz,
                                          C.c_float(F32(self.h)), L.ptr(out), L.ptr(ws), ws_bytes, L.stream_handle()),
                        "jt_kde_grid3d")
                return out
            X, Y, Z = np.meshgrid(x, y, z, indexing="xy")
            grid = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
            return torch.from_numpy(self._evaluate_grid(grid, self.P).reshape(ny, nx, nz))
        R = F32(float(self.cutoff_sigmas) * self.h)
        hg = HashGridNeighbors(self._dev_points(self.P), float(R) / 2, 1, "x64", self.device)
        out = torch.empty((ny, nx, nz), dtype=torch.float32, device=self.device)
        gx, gy, gz = (_to_dev(a, self.device) for a in (x, y, z))
        ws_bytes = int(lib.jt_kde_grid_cutoff_workspace_size(nx, ny, nz))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=self.device)
        L.check(lib.jt_kde_grid_cutoff_ws(C.byref(hg._desc), self.N, L.ptr(gx), nx, L.ptr(gy), ny, L.ptr(gz), nz,
                                          C.c_float(F32(self.h)), C.c_float(R), L.ptr(out), L.ptr(ws), ws_bytes,
                                          L.stream_handle()), "jt_kde_grid_cutoff_ws")
        return out

    def evaluate_3d(self):
        """kde.py:223-240."""
        x, y, z = self._axes_3d()
        X, Y, Z = np.meshgrid(x, y, z, indexing="xy")
        D = self.density_grid_3d()
        D = D.cpu().numpy() if isinstance(D, torch.Tensor) else D
        if self.normalize:
            dx = (X.max() - X.min()) / max(X.shape[1] - 1, 1)
            dy = (Y.max() - Y.min()) / max(Y.shape[0] - 1, 1)
            dz = (Z.max() - Z.min()) / max(Z.shape[2] - 1, 1)
            cell_vol = float(dx * dy * dz)
            total = float((D * cell_vol).sum())
            if total > 0:
                D = D * (self.N / total)
            domain_vol = float((X.max() - X.min()) * (Y.max() - Y.min()) * (Z.max() - Z.min()))
            avg = self.N / max(domain_vol, 1e-12)
            D = D / max(avg, 1e-12)
        return X, Y, Z, D
