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
