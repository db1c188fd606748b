"""Oracle: temporal blend, trilinear sampling and nearest-node sampling.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows
  * jaxtrace/fields/time_series.py:155-245  (``_sample_all_nodes_at_time_jax``)
  * jaxtrace/fields/time_series.py:249-259  (``_nearest_neighbor_sample_jax``)
  * jaxtrace/fields/structured.py:123-161, 167-250  (metadata, ``_trilinear_jax``)

The configs' "structured trilinear in space, linear in time" field has no
single reference class: its reference semantics are the composition
``_trilinear_jax(_sample_all_nodes_at_time_jax(t))`` (blend every node, then
interpolate).  ``structured_series_sample`` restates exactly that; the GPU
kernel blends only the 8 gathered corners with the same per-node arithmetic.

Pinned in tests by: multilinear fields reproduced exactly by trilinear
interpolation, periodic wrap invariance, fields linear in t reproduced
exactly by the blend, snapshot identity at t = t_k, extrapolation special
cases, brute-force nearest node on tiny inputs.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def jmod(a, n):
    """``jnp.mod`` for floats: fmod with the sign fixed to the divisor's."""
    a = np.asarray(a)
    n = np.asarray(n, dtype=a.dtype)
    r = np.fmod(a, n)
    fix = (r != 0) & ((r < 0) != (n < 0))
    return np.where(fix, r + n, r).astype(a.dtype)


def clip(x, lo, hi):
    """``jnp.clip`` = minimum(maximum(x, lo), hi) with NaN propagation."""
    x = np.asarray(x)
    y = np.where(x < lo, lo, x)
    y = np.where(y > hi, hi, y)
    return np.where(np.isnan(x), x, y).astype(x.dtype)


# ---------------------------------------------------------------------------
# Temporal blend (A4)
# ---------------------------------------------------------------------------

def blend_params(times: np.ndarray, t, interpolation="linear", extrapolation="constant"):
    """Scalar part of time_series.py:155-245 for one query time ``t`` (f32).

    Returns (mode, i0, i1, alpha) where mode is one of
    'interp' -> (1-a)*V[i0] + a*V[i1]        (:179-188)
    'copy'   -> V[i0]                         (nearest, constant extrapolation)
    'left'   -> V0 + a*(V1-V0), V0=V[i0]      (:190-199)
    'right'  -> V1 + a*(V1-V0), V0=V[i0], V1=V[i1]   (:201-210)
    'zero' / 'nan'.
    """
    times = np.asarray(times, dtype=F32)
    t = F32(t)
    T = times.shape[0]
    interp = interpolation.lower()
    extra = extrapolation.lower()
    if interp == "nearest":                       # :171-174
        idx = int(np.argmin(np.abs(times - t)))
        return ("copy", idx, idx, F32(0))
    if interp != "linear":
        raise ValueError(interp)
    ir = int(np.searchsorted(times, t, side="right"))   # :162-164
    ir = min(max(ir, 0), T)
    il = min(max(ir - 1, 0), T - 1)
    t_min, t_max = times[0], times[-1]
    at_left = t < t_min
    at_right = t > t_max
    if at_left or at_right:
        if extra == "constant":                   # :217-219
            return ("copy", 0 if at_left else T - 1, 0 if at_left else T - 1, F32(0))
        if extra == "zero":
            return ("zero", 0, 0, F32(0))
        if extra == "nan":
            return ("nan", 0, 0, F32(0))
        if extra == "linear":
            if at_left:                           # :190-199
                i1 = min(1, T - 1)
                t0, t1 = times[0], times[i1]
                d = np.maximum(t1 - t0, F32(1e-12))
                a = F32((t - t0) / d)
                if T > 1:
                    return ("left", 0, i1, a)
                return ("copy", 0, 0, F32(0))
            i0 = max(T - 2, 0)                    # :201-210
            t0, t1 = times[i0], times[T - 1]
            d = np.maximum(t1 - t0, F32(1e-12))
            a = F32((t - t1) / d)
            if T > 1:
                return ("right", i0, T - 1, a)
            return ("copy", T - 1, T - 1, F32(0))
        raise ValueError(extra)
    irc = min(max(ir, 0), T - 1)                  # :178-188
    t0, t1 = times[il], times[irc]
    d = np.maximum(t1 - t0, F32(1e-12))
    a = clip(F32((t - t0) / d), F32(0), F32(1))
    return ("interp", il, irc, F32(a))


def apply_blend(params, V0, V1):
    """Apply ``blend_params`` output to two node-value arrays (f32)."""
    mode, _, _, a = params
    a = F32(a)
    if mode == "interp":
        return ((F32(1) - a) * V0 + a * V1).astype(F32)
    if mode == "copy":
        return np.array(V0, dtype=F32, copy=True)
    if mode == "left":
        return (V0 + a * (V1 - V0)).astype(F32)
    if mode == "right":
        return (V1 + a * (V1 - V0)).astype(F32)
    if mode == "zero":
        return np.zeros_like(V0, dtype=F32)
    if mode == "nan":
        return np.full_like(V0, np.nan, dtype=F32)
    raise ValueError(mode)


def sample_all_nodes_at_time(data, times, t, interpolation="linear", extrapolation="constant"):
    """time_series.py:155-245: blend every node, (T, ..., 3) -> (..., 3)."""
    p = blend_params(times, t, interpolation, extrapolation)
    return apply_blend(p, data[p[1]], data[p[2]])


# ---------------------------------------------------------------------------
# Trilinear (A6)
# ---------------------------------------------------------------------------

class GridMeta:
    """structured.py:101-161: f32 grid metadata exactly as the reference derives it."""

    def __init__(self, grid_x, grid_y, grid_z):
        gx = np.asarray(grid_x, dtype=F32)
        gy = np.asarray(grid_y, dtype=F32)
        gz = np.asarray(grid_z, dtype=F32)
        self.n = (gx.shape[0], gy.shape[0], gz.shape[0])
        # dx = float(grid_x[1] - grid_x[0]) (f32 subtraction), then f32 again
        self.d = tuple(F32(g[1] - g[0]) if g.shape[0] > 1 else F32(1.0) for g in (gx, gy, gz))
        self.lo = tuple(F32(g[0]) for g in (gx, gy, gz))


def trilinear(V, meta: GridMeta, pos, mode="clamp"):
    """structured.py:167-250 in float32.  V: (Nx,Ny,Nz,3) f32 (or any object
    supporting ``V[i, j, k]`` fancy indexing); pos (M,3) f32."""
    if isinstance(V, (np.ndarray, list)):
        V = np.asarray(V, dtype=F32)
    pos = np.asarray(pos, dtype=F32)
    mode = mode.lower()
    i0s, i1s, ws = [], [], []
    valid = np.ones(pos.shape[0], dtype=bool)
    for ax in range(3):
        x = pos[:, ax]
        n = meta.n[ax]
        f = ((x - meta.lo[ax]) / meta.d[ax]).astype(F32)          # :176-178
        if mode == "periodic":
            f = jmod(f, F32(n))                                    # :182-184
        else:
            hi = F32(meta.lo[ax] + F32(F32(n - 1) * meta.d[ax]))  # :188-190
            valid &= (x >= meta.lo[ax]) & (x <= hi)
            f = clip(f, F32(0), F32(max(n - 1, 0)))                # :193-201
        with np.errstate(invalid="ignore"):
            fl = np.floor(f)
            i0 = np.where(np.isnan(fl), 0, fl).astype(np.int64)   # NaN -> 0 (XLA convert)
        if mode == "periodic":
            i1 = np.mod(i0 + 1, n)                                 # :208-211
        else:
            i1 = np.clip(i0 + 1, 0, max(n - 1, 0))                 # :213-215
        w = (f - i0.astype(F32)).astype(F32)                       # :218-220
        # XLA gather clamps out-of-range indices (f may equal n after the mod)
        i0s.append(np.clip(i0, 0, n - 1))
        i1s.append(np.clip(i1, 0, n - 1))
        ws.append(w)
    (i0, j0, k0), (i1, j1, k1), (wx, wy, wz) = i0s, i1s, ws
    one = F32(1)
    v000 = V[i0, j0, k0]; v100 = V[i1, j0, k0]                     # :223-231
    v010 = V[i0, j1, k0]; v110 = V[i1, j1, k0]
    v001 = V[i0, j0, k1]; v101 = V[i1, j0, k1]
    v011 = V[i0, j1, k1]; v111 = V[i1, j1, k1]
    ax_, ay_, az_ = (one - wx)[:, None], (one - wy)[:, None], (one - wz)[:, None]
    bx, by, bz = wx[:, None], wy[:, None], wz[:, None]
    c00 = v000 * ax_ + v100 * bx                                   # :234-237
    c10 = v010 * ax_ + v110 * bx
    c01 = v001 * ax_ + v101 * bx
    c11 = v011 * ax_ + v111 * bx
    c0 = c00 * ay_ + c10 * by                                      # :239-240
    c1 = c01 * ay_ + c11 * by
    vals = (c0 * az_ + c1 * bz).astype(F32)                        # :242
    if mode == "zero":                                             # :244-247
        vals = np.where(valid[:, None], vals, F32(0))
    elif mode == "nan":
        vals = np.where(valid[:, None], vals, F32(np.nan))
    return vals.astype(F32)


def structured_series_sample(data, times, meta, pos, t, mode="clamp",
                             interpolation="linear", extrapolation="constant"):
    """Composition A4 then A6 over every node: data (T,Nx,Ny,Nz,3)."""
    V = sample_all_nodes_at_time(data, times, t, interpolation, extrapolation)
    return trilinear(V, meta, pos, mode)


class _CornerBlend:
    """Index proxy that blends only the nodes that are gathered."""

    def __init__(self, data, params):
        self.data, self.p = data, params

    def __getitem__(self, idx):
        p = self.p
        if p[0] in ("zero", "nan"):
            return apply_blend(p, self.data[0][idx], self.data[0][idx])
        return apply_blend(p, self.data[p[1]][idx], self.data[p[2]][idx])


def structured_series_sample_corners(data, times, meta, pos, t, mode="clamp",
                                     interpolation="linear", extrapolation="constant"):
    """Same values as ``structured_series_sample`` (the per-node blend
    arithmetic is identical, so results are bitwise equal) but blends only the
    gathered corners; used where blending every node would be too slow."""
    p = blend_params(times, t, interpolation, extrapolation)
    return trilinear(_CornerBlend(data, p), meta, pos, mode)


# ---------------------------------------------------------------------------
# Nearest-node sampling (A5)
# ---------------------------------------------------------------------------

def nearest_node_index(node_pos, query):
    """time_series.py:249-259: argmin_n sum((q-p_n)^2), first minimum wins.

    The squared distance is ((dx*dx + dy*dy) + dz*dz) in f32."""
    P = np.asarray(node_pos, dtype=F32)
    Q = np.asarray(query, dtype=F32)
    out = np.empty(Q.shape[0], dtype=np.int64)
    step = max(1, 4_000_000 // max(P.shape[0], 1))
    for s in range(0, Q.shape[0], step):
        d = Q[s:s + step, None, :] - P[None, :, :]
        d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        out[s:s + step] = np.argmin(d2, axis=1)
    return out


def nearest_node_sample(data, times, node_pos, query, t,
                        interpolation="linear", extrapolation="constant"):
    """TimeSeriesField.sample_at_positions (time_series.py:391-398)."""
    V = sample_all_nodes_at_time(data, times, t, interpolation, extrapolation)
    return V[nearest_node_index(node_pos, query)]
