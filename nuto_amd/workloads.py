"""Synthetic mesh batches of the batched workloads (seeded, reproducible).

mixed_batch: meshes with element counts log-uniform in [lo, hi], element type drawn uniformly from
{tet4, tet10, hex8, hex20} (tet10 / hex20 are quadratic dof layers over tet4 / hex8 coordinate
elements), structured grids of 1 mm spacing with every node moved by up to jitter x spacing.
gdm_batch: hex8 meshes with log-uniform element counts (gradient damage).
"""
import numpy as np

from . import binding as B
from . import synthetic as S

SEED_COUNT = 0x4001
SEED_TYPE = 0x4002
SEED_JITTER = 0x4003

TYPES = ((B.TET4, B.TET4), (B.TET4, B.TET10), (B.HEX8, B.HEX8), (B.HEX8, B.HEX20))


def _grid(cells, per_brick):
    bricks = max(1.0, cells / per_brick)
    n = max(1, int(round(bricks ** (1.0 / 3.0))))
    nz = max(1, int(round(bricks / (n * n))))
    return n, n, nz


def mesh(coord_type, dof_type, cells, jitter, seed):
    """One structured mesh of about `cells` elements: (coords, conn, ranks, n_dof_nodes)."""
    nx, ny, nz = _grid(cells, 6.0 if coord_type == B.TET4 else 1.0)
    X, C = B.structured_mesh(coord_type, nx, ny, nz, (float(nx), float(ny), float(nz)), jitter, seed)
    ranks, nn = B.dof_layer(coord_type, C, dof_type)
    return X, C, ranks, nn


def mixed_batch(n_meshes, lo=1e3, hi=5e4, jitter=0.1):
    """[(coord_type, dof_type, coords, conn, ranks, n_dof_nodes)] of the mixed-type batch."""
    counts = np.exp(S.uniform(SEED_COUNT, n_meshes, np.log(lo), np.log(hi)))
    kinds = np.minimum((S.uniform(SEED_TYPE, n_meshes) * 4).astype(int), 3)
    out = []
    for m in range(n_meshes):
        ct, dt = TYPES[kinds[m]]
        out.append((ct, dt) + mesh(ct, dt, counts[m], jitter, SEED_JITTER + m))
    return out


def gdm_batch(n_meshes, lo=1e4, hi=2e5):
    """Hex8 meshes (no jitter) for the batched gradient-damage workload."""
    counts = np.exp(S.uniform(SEED_COUNT, n_meshes, np.log(lo), np.log(hi)))
    return [(B.HEX8, B.HEX8) + mesh(B.HEX8, B.HEX8, counts[m], 0.0, 0) for m in range(n_meshes)]
