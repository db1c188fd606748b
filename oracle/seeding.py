"""Oracle: uniform-grid seed positions in float32.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows jaxtrace/tracking/seeding.py:185-247 (``uniform_grid_seeds``): three
float32 coordinate arrays from np.linspace (the interior variant shifts both
ends by (max - min) / (n + 1); a z resolution of 1 gives [z_min]), then
np.meshgrid(..., indexing='ij') flattened into (N, 3) columns.  Bounds are the
(2, 3) [[min], [max]] form of seeding.py:88-137.

Pinned in tests by: a pure-Python triple loop over the coordinates (row
order i, j, k); exact integer lattices (bounds [0, n - 1] give coordinates
0..n-1, interior bounds [0, n + 1] give 1..n); corner rows equal to the
bounds; the 2-component resolution's z column equal to z_min.

Also ``random_seeds`` (seeding.py:140-182) and the inlet-face generator
``_generate_inlet_grid_particles`` (jaxtrace/tracking/boundary.py:1321-1441).
random_seeds draws jax.random.uniform(PRNGKey(seed), (n, 3)): threefry2x32
with 20 rounds (Salmon et al., "Parallel random numbers: as easy as 1, 2,
3", SC11; the Random123 reference), counters laid out as JAX's
non-partitionable random_bits (jax_threefry_partitionable=False, the default
before JAX 0.5): the words 0..m-1 are split into halves, block i hashes
(i, h + i) and yields word i and word h + i.  threefry2x32 is pinned by the
Random123 known-answer vectors; the JAX layout itself is parity unpinned
(no JAX here, and the reference's JAX version is unpinned: JAX >= 0.5
defaults to a different counter layout).
"""
import numpy as np

F32 = np.float32


def axes(resolution, bounds, include_boundaries=True):
    """seeding.py:209-236: the three coordinate arrays."""
    b = np.asarray(bounds, dtype=F32).reshape(2, 3)
    if isinstance(resolution, (int, np.integer)):
        n = (int(resolution),) * 3
    elif len(resolution) == 2:
        n = (int(resolution[0]), int(resolution[1]), 1)
    else:
        n = tuple(int(r) for r in resolution)
    out = []
    for a in range(3):
        lo, hi = b[0, a], b[1, a]
        if a == 2 and n[2] <= 1:
            out.append(np.array([lo], dtype=F32))
            continue
        if include_boundaries:
            out.append(np.linspace(lo, hi, n[a], dtype=F32))
        else:
            d = (hi - lo) / (n[a] + 1)
            out.append(np.linspace(lo + d, hi - d, n[a], dtype=F32))
    return out


def expand(xc, yc, zc):
    """seeding.py:239-245: meshgrid 'ij', flattened, stacked as columns."""
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    return np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]).astype(F32)


def uniform_grid_seeds(resolution, bounds, include_boundaries=True):
    return expand(*axes(resolution, bounds, include_boundaries))


# ---------------------------------------------------------------------------
# random_seeds (seeding.py:140-182)
# ---------------------------------------------------------------------------

U32 = np.uint32
_ROT = (13, 15, 26, 6, 17, 29, 16, 24)


def _rotl(x, r):
    return ((x << U32(r)) | (x >> U32(32 - r))).astype(U32)


def threefry2x32(k0, k1, x0, x1):
    """Threefry-2x32, 20 rounds, key injection every 4 rounds (Random123)."""
    k0, k1 = U32(k0), U32(k1)
    ks = (k0, k1, U32(0x1BD11BDA) ^ k0 ^ k1)
    x0 = np.asarray(x0, U32) + ks[0]
    x1 = np.asarray(x1, U32) + ks[1]
    for r in range(20):
        x0 = (x0 + x1).astype(U32)
        x1 = _rotl(x1, _ROT[r % 8]) ^ x0
        if r % 4 == 3:
            s = (r + 1) // 4
            x0 = (x0 + ks[s % 3]).astype(U32)
            x1 = (x1 + ks[(s + 1) % 3] + U32(s)).astype(U32)
    return x0, x1


def prng_key(seed):
    """PRNGKey(seed): the 32-bit halves (seed >> 32, seed & 0xffffffff)."""
    seed = int(seed)
    return (seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF


def random_bits32(key, m):
    """m 32-bit words: iota(m), padded with one 0 when m is odd, split into
    halves x0 | x1; block i hashes (x0[i], x1[i]); the words are y0 | y1."""
    h = (m + 1) // 2
    cnt = np.concatenate([np.arange(m, dtype=U32), np.zeros(m % 2, U32)])
    with np.errstate(over="ignore"):
        y0, y1 = threefry2x32(key[0], key[1], cnt[:h], cnt[h:])
    return np.concatenate([y0, y1])[:m]


def uniform01(key, shape):
    """jax.random.uniform(key, shape, float32) on [0, 1): the top 23 bits as
    the mantissa of a float in [1, 2), minus 1."""
    m = int(np.prod(shape))
    bits = random_bits32(key, m)
    f = ((bits >> U32(9)) | U32(0x3F800000)).view(np.float32) - F32(1.0)
    return np.maximum(F32(0.0), f * F32(1.0) + F32(0.0)).reshape(shape)


def random_seeds(n, bounds, rng_seed=0):
    """seeding.py:140-174 (JAX branch): lo + u * (hi - lo)."""
    if n <= 0:
        return np.zeros((0, 3), F32)
    b = np.asarray(bounds, F32).reshape(2, 3)
    u = uniform01(prng_key(rng_seed), (n, 3))
    lo, hi = b[0], b[1]
    return (lo + u * (hi - lo)).astype(F32)


# ---------------------------------------------------------------------------
# _generate_inlet_grid_particles (boundary.py:1321-1441)
# ---------------------------------------------------------------------------

def inlet_grid_indices(n_particles, total):
    """The inlet-face grid indices of boundary.py:1387-1424, as the reference
    computes them (Python loops included)."""
    if n_particles <= total:
        step = total / n_particles
        indices = np.array([int(i * step) for i in range(n_particles)])
        indices = np.clip(indices, 0, total - 1)
        indices = np.unique(indices)
        while len(indices) < n_particles and len(indices) < total:
            if len(indices) == 0:
                indices = np.array([0])
            else:
                gaps = np.diff(np.concatenate([[0], indices, [total]]))
                max_gap_idx = np.argmax(gaps[:-1])
                if max_gap_idx == 0:
                    new_idx = indices[0] // 2
                else:
                    new_idx = (indices[max_gap_idx - 1] + indices[max_gap_idx]) // 2
                indices = np.sort(np.append(indices, new_idx))
                indices = np.unique(indices)
        return indices[:n_particles]
    base = np.arange(total)
    extra = n_particles - total
    extra_step = total / extra
    return np.concatenate([base, np.array([int(i * extra_step) for i in range(extra)])])


def inlet_grid_particles(n_particles, inlet_position, axis_idx, bounds, concentrations):
    """boundary.py:1321-1441 for 3-D bounds (two inlet-face axes)."""
    bounds = np.asarray(bounds, F32).reshape(2, 3)
    names = ["x", "y", "z"]
    other = [i for i in range(3) if i != axis_idx]
    res = [max(1, int(concentrations[names[i]])) for i in other]
    c0 = np.linspace(bounds[0, other[0]], bounds[1, other[0]], res[0], dtype=F32)
    c1 = np.linspace(bounds[0, other[1]], bounds[1, other[1]], res[1], dtype=F32)
    G1, G2 = np.meshgrid(c0, c1, indexing="ij")
    g1, g2 = G1.ravel(), G2.ravel()
    idx = inlet_grid_indices(n_particles, len(g1))
    out = np.zeros((n_particles, 3), F32)
    out[:, axis_idx] = inlet_position
    for i in range(n_particles):
        out[i, other[0]] = g1[idx[i]]
        out[i, other[1]] = g2[idx[i]]
    return out
