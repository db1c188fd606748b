"""Oracle: hash-grid neighbours, SPH density and Gaussian KDE in float32.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows
  * jaxtrace/density/neighbors.py:61-103  (``HashGridNeighbors.__post_init__``)
  * jaxtrace/density/neighbors.py:109-168 (``query_many``)
  * jaxtrace/density/sph.py:107-160, 227-236, 269-279, 333-380
  * jaxtrace/density/kde.py:81-97, 188-270 and kernels.py:30-47, 73-95

Key modes.  The reference packs cell coordinates with ``astype(jnp.int64)``
but never enables ``jax_enable_x64``, so the keys are int32 and
``c << 42`` (a shift >= the bit width) is 0 under XLA's shift semantics:
the as-shipped 3-D key is ``(cy << 21) | cz``.  ``key_mode='x32'``
restates that; ``key_mode='x64'`` restates the intended
``(cx << 42) | (cy << 21) | cz``.  The x32 truncation itself can only be
confirmed by running JAX, which is absent here: parity unpinned for that one
semantic choice (both modes are implemented and tested).

Op order choices that the reference leaves to XLA (parity pinned only to
rounding): the distance is sqrt((dx*dx + dy*dy) + dz*dz); sums and means use
NumPy's order.  Candidate order, caps, dummy slots (index
``pt_indices_sorted[0]``, distance inf) and the stable distance sort follow
the reference exactly; the tests pin them by brute-force enumeration.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


class HashGrid:
    """neighbors.py:39-103 (3-D and 2-D)."""

    def __init__(self, positions, cell_size, per_cell_cap=64, key_mode="x32"):
        P = np.asarray(positions, dtype=F32)
        if P.ndim != 2 or P.shape[1] not in (2, 3):
            raise ValueError("positions must be (N,2) or (N,3)")
        self.P = P
        self.D = P.shape[1]
        self.cap = int(per_cell_cap)
        self.key_mode = key_mode
        self.inv = F32(F32(1.0) / F32(max(float(cell_size), 1e-12)))     # :67
        self.origin = P.min(axis=0)                                       # :68
        C = np.floor((P - self.origin) * self.inv).astype(np.int32)       # :70
        self.Cmin = C.min(axis=0)
        Csh = C - self.Cmin
        keys = self.pack(Csh)
        self.order = np.argsort(keys, kind="stable").astype(np.int64)    # :93
        self.sorted_keys = keys[self.order]
        uniq, first = np.unique(self.sorted_keys, return_index=True)      # :98
        self.unique_ids = uniq
        self.starts = first.astype(np.int64)
        self.counts = np.diff(np.concatenate([first, [self.sorted_keys.size]])).astype(np.int64)

    def pack(self, c):
        c = np.asarray(c)
        if self.key_mode == "x64":
            c = c.astype(np.int64)
            if self.D == 3:
                return (c[..., 0] << 42) | (c[..., 1] << 21) | c[..., 2]
            return (c[..., 0] << 32) | c[..., 1]
        c = c.astype(np.int32)
        # int32 keys; shifts >= 32 produce 0 (XLA ShiftLeft semantics)
        if self.D == 3:
            return ((c[..., 1] << np.int32(21)) | c[..., 2]).astype(np.int32)
        return c[..., 1].astype(np.int32)

    def offsets(self):
        r = (-1, 0, 1)
        if self.D == 3:
            return np.array([[a, b, c] for a in r for b in r for c in r], np.int32)
        return np.array([[a, b] for a in r for b in r], np.int32)

    def cell_of(self, Q):
        """neighbors.py:105-107."""
        return np.floor((np.asarray(Q, F32) - self.origin) * self.inv).astype(np.int32) - self.Cmin

    def candidates(self, q):
        """Candidate list of one query in the reference's order: 27 (or 9)
        probes in (di,dj,dk) lexicographic order, ``cap`` slots each.
        Returns (cand_pts, cand_valid) of length probes*cap."""
        cq = self.cell_of(q[None])[0]
        ngh = cq[None, :] + self.offsets()
        ids = self.pack(ngh)
        U = self.unique_ids.size
        pos = np.searchsorted(self.unique_ids, ids, side="left")
        present = (pos < U) & (self.unique_ids[np.minimum(pos, U - 1)] == ids)
        starts = np.where(present, self.starts[np.minimum(pos, U - 1)], -1)
        counts = np.where(present, self.counts[np.minimum(pos, U - 1)], 0)
        ar = np.arange(self.cap)
        base = starts[:, None] + ar[None, :]
        valid_slot = (starts[:, None] >= 0) & (ar[None, :] < counts[:, None])
        base_safe = np.clip(base, 0, max(self.order.size - 1, 0))
        cand_sorted = np.where(valid_slot, base_safe, 0)
        return self.order[cand_sorted].reshape(-1), valid_slot.reshape(-1)

    def query_many(self, Q, radius, max_neighbors):
        """neighbors.py:109-168 -> (idx int32 (M,K), dist f32 (M,K))."""
        Q = np.asarray(Q, dtype=F32)
        r = F32(radius)
        K = int(max_neighbors)
        M = Q.shape[0]
        out_i = np.full((M, K), -1, np.int32)
        out_d = np.full((M, K), np.inf, F32)
        for m in range(M):
            cp, cv = self.candidates(Q[m])
            diff = self.P[cp] - Q[m][None, :]
            sq = diff * diff
            s = sq[:, 0] + sq[:, 1]
            if self.D == 3:
                s = s + sq[:, 2]
            d = np.sqrt(s).astype(F32)
            valid = cv & (d <= r)
            dsel = np.where(valid, d, F32(np.inf)).astype(F32)
            order = np.argsort(dsel, kind="stable")[:K]
            k = order.size
            out_i[m, :k] = cp[order].astype(np.int32)
            out_d[m, :k] = dsel[order]
        return out_i, out_d


# ---------------------------------------------------------------------------
# SPH (sph.py)
# ---------------------------------------------------------------------------

def cubic_spline(r, h, D=3):
    """sph.py:227-236 in float32 (integer powers as lax.integer_pow: x*(x*x))."""
    r = np.asarray(r, F32)
    hh = np.maximum(np.asarray(h, F32), F32(1e-12))
    q = (r / hh).astype(F32)
    if D == 2:
        sigma = F32(10.0) / (F32(7.0 * np.pi) * (hh * hh))     # 7.0*jnp.pi is a Python product
    else:
        sigma = F32(1.0) / (F32(np.pi) * (hh * (hh * hh)))
    q2 = q * q
    q3 = q * q2
    with np.errstate(invalid="ignore", over="ignore"):
        w1 = sigma * ((F32(1.0) - F32(1.5) * q2) + F32(0.75) * q3)
        t = F32(2.0) - q
        w2 = (sigma * F32(0.25)) * (t * (t * t))
    return np.where(q <= 1.0, w1, np.where(q <= 2.0, w2, F32(0.0))).astype(F32)


PLANE_AXES = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}


def extract_2d(P3, plane="xy", position=0.0, slab_thickness=0.1):
    """sph.py:164-177 / kde.py:112-124: the points of the slab
    |p_k - position| <= thickness/2, projected on the plane's two axes (all
    points projected when the slab is empty)."""
    P3 = np.asarray(P3)
    i, j, k = PLANE_AXES[plane]
    mask = np.abs(P3[:, k] - float(position)) <= (float(slab_thickness) * 0.5)
    pts2 = P3[mask][:, [i, j]]
    if pts2.size == 0:
        pts2 = P3[:, [i, j]]
    return pts2.astype(F32, copy=False)


def sph_density(P, Q, h, per_cell_cap=64, max_neighbors=512, key_mode="x32", normalize=True,
                masses=None, kernel="cubic_spline", dimensions="3d", plane="xy", position=0.0,
                slab_thickness=0.1):
    """SPHDensityEstimator(dimensions, smoothing_length=h, kernel_type,
    neighbor_search='hashgrid').evaluate(Q) (sph.py:107-160, 269-279,
    333-380).  dimensions='2d' takes the slab of (N,3) positions
    (sph.py:112-117) and projects (M,3) queries on the plane (:354-361)."""
    P = np.asarray(P, F32)
    Q = np.asarray(Q, F32)
    Dm = 3
    if dimensions.lower() == "2d":
        P = extract_2d(P, plane, position, slab_thickness)
        Dm = 2
        if Q.shape[1] == 3:
            i, j, _ = PLANE_AXES[plane]
            Q = Q[:, [i, j]]
    N = P.shape[0]
    m = np.ones(N, F32) if masses is None else np.asarray(masses, F32)
    hs = np.full(N, F32(h), F32)
    mean_h = float(np.mean(hs))                                            # :142
    grid = HashGrid(P, max(mean_h, 1e-6), per_cell_cap, key_mode)          # :148-152
    idx, dist = grid.query_many(Q, 2.0 * mean_h, max_neighbors)            # :273
    valid = idx >= 0
    mass = np.where(valid, m[np.maximum(idx, 0)], F32(0))
    hi = np.where(valid, hs[np.maximum(idx, 0)], F32(mean_h))
    w = SPH_KERNELS[kernel](dist, hi, Dm)
    rho = (mass * w).sum(axis=1, dtype=F32)
    if normalize:                                                          # :365-378
        xmin, xmax = float(Q[:, 0].min()), float(Q[:, 0].max())
        ymin, ymax = float(Q[:, 1].min()), float(Q[:, 1].max())
        if Dm == 2:
            area = max((xmax - xmin) * (ymax - ymin), 1e-12)
            rho = (rho / F32(N / area)).astype(F32)
        else:
            zmin, zmax = float(Q[:, 2].min()), float(Q[:, 2].max())
            vol = max((xmax - xmin) * (ymax - ymin) * (zmax - zmin), 1e-12)
            rho = (rho / F32(N / vol)).astype(F32)
    return rho.astype(F32), idx, dist


# ---------------------------------------------------------------------------
# KDE (kde.py, kernels.py)
# ---------------------------------------------------------------------------

def scott_bandwidth(P):
    """kernels.py:30-47 (NumPy, float32 input)."""
    x = np.asarray(P)
    n, d = x.shape
    sigma = np.std(x, axis=0, ddof=1).mean()
    return sigma * n ** (-1.0 / (d + 4.0))


def silverman_bandwidth(P):
    """kernels.py:50-70."""
    x = np.asarray(P)
    n, d = x.shape
    sigma = np.std(x, axis=0, ddof=1).mean()
    return (4.0 / (d + 2.0)) ** (1.0 / (d + 4.0)) * sigma * n ** (-1.0 / (d + 4.0))


def make_grid_3d(P, resolution, bounds=None):
    """kde.py:188-201: f32 linspace axes, meshgrid(indexing='xy') -> (ny,nx,nz)."""
    P = np.asarray(P, F32)
    if bounds is None:
        lo, hi = P.min(axis=0), P.max(axis=0)
    else:
        lo = np.array([b[0] for b in bounds], dtype=float)
        hi = np.array([b[1] for b in bounds], dtype=float)
    res = resolution if not isinstance(resolution, int) else (resolution,) * 3
    nx, ny, nz = map(int, res)
    x = np.linspace(lo[0], hi[0], nx, dtype=np.float32)
    y = np.linspace(lo[1], hi[1], ny, dtype=np.float32)
    z = np.linspace(lo[2], hi[2], nz, dtype=np.float32)
    return np.meshgrid(x, y, z, indexing="xy")


def kde_eval(Q, P, h, dtype=F32):
    """kde.py:244-270 exhaustive: mean_n (2pi)^(-d/2) exp(-0.5 |q-p|^2 * f32(1/h^2))."""
    Q = np.asarray(Q, dtype)
    P = np.asarray(P, dtype)
    d = Q.shape[1]
    inv_h2 = dtype(1.0 / (h * h))
    norm = dtype((2.0 * math.pi) ** (-0.5 * d))
    out = np.empty(Q.shape[0], dtype)
    step = max(1, 2_000_000 // max(P.shape[0], 1))
    for s in range(0, Q.shape[0], step):
        diff = Q[s:s + step, None, :] - P[None, :, :]
        sq = diff * diff
        r2 = sq[..., 0] + sq[..., 1]
        if d == 3:
            r2 = r2 + sq[..., 2]
        k = norm * np.exp(dtype(-0.5) * (r2 * inv_h2))
        out[s:s + step] = k.mean(axis=1, dtype=dtype)
    return out


def kde_normalize_3d(X, Y, Z, D, N):
    """kde.py:229-239."""
    dx = (X.max() - X.min()) / max(X.shape[1] - 1, 1)
    dy = (Y.max() - Y.min()) / max(Y.shape[0] - 1, 1)
    dz = (Z.max() - Z.min()) / max(Z.shape[2] - 1, 1)
    cell_vol = float(dx * dy * dz)
    total = float((D * cell_vol).sum())
    if total > 0:
        D = D * (N / total)
    domain_vol = float((X.max() - X.min()) * (Y.max() - Y.min()) * (Z.max() - Z.min()))
    avg = N / max(domain_vol, 1e-12)
    return D / max(avg, 1e-12)


def make_grid_2d(pts, resolution, bounds=None):
    """kde.py:173-186 on the 2-D points: f32 linspace axes, meshgrid 'xy' -> (ny, nx)."""
    pts = np.asarray(pts, F32)
    if bounds is None:
        lo, hi = pts.min(axis=0), pts.max(axis=0)
    else:
        lo = np.array([bounds[0][0], bounds[1][0]], dtype=float)
        hi = np.array([bounds[0][1], bounds[1][1]], dtype=float)
    res = resolution if isinstance(resolution, int) else resolution[:2]
    nx = ny = int(res)
    x = np.linspace(lo[0], hi[0], nx, dtype=np.float32)
    y = np.linspace(lo[1], hi[1], ny, dtype=np.float32)
    return np.meshgrid(x, y, indexing="xy")


def kde_normalize_2d(X, Y, Z, N):
    """kde.py:210-220."""
    dx = (X.max() - X.min()) / max(X.shape[1] - 1, 1)
    dy = (Y.max() - Y.min()) / max(Y.shape[0] - 1, 1)
    cell_area = float(dx * dy)
    total = float((Z * cell_area).sum())
    if total > 0:
        Z = Z * (N / total)
    domain_area = float((X.max() - X.min()) * (Y.max() - Y.min()))
    avg = N / max(domain_area, 1e-12)
    return Z / max(avg, 1e-12)


def wendland(r, h, D=3):
    """sph.py:238-248 in float32 (t**4 as (t*t)*(t*t))."""
    r = np.asarray(r, F32)
    hh = np.maximum(np.asarray(h, F32), F32(1e-12))
    q = (r / hh).astype(F32)
    if D == 2:
        sigma = F32(7.0) / (F32(np.pi) * (hh * hh))
    else:
        sigma = F32(21.0) / (F32(2.0 * np.pi) * (hh * (hh * hh)))
    t = np.maximum(F32(1.0) - q, F32(0.0))
    t2 = t * t
    with np.errstate(invalid="ignore"):
        out = (sigma * (t2 * t2)) * (F32(1.0) + F32(4.0) * q)
    return np.where(np.isinf(q), F32(0), out).astype(F32)


def gaussian_sph(r, h, D=3):
    """sph.py:250-258 in float32."""
    r = np.asarray(r, F32)
    hh = np.maximum(np.asarray(h, F32), F32(1e-12))
    u = (r / hh).astype(F32)
    q2 = u * u
    if D == 2:
        sigma = F32(1.0) / (F32(np.pi) * (hh * hh))
    else:
        sigma = (F32(1.0) / np.power(F32(np.pi) * (hh * hh), F32(1.5))).astype(F32)
    return (sigma * np.exp(-q2)).astype(F32)


SPH_KERNELS = {"cubic_spline": cubic_spline, "wendland": wendland, "gaussian": gaussian_sph}
