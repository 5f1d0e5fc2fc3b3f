"""Isosurface pins for the marching-cubes kernel -- TEST INFRASTRUCTURE ONLY
(only tests/ use it).

``crossing_vertices`` restates the vertex set of the reference's meshing
(skimage marching cubes with method "lewiner",
gempy/modules/mesh_extranction/marching_cubes.py:86-93: one vertex per lattice
edge whose end values straddle the level, linearly interpolated, shared between
cubes) by brute force over all lattice edges with numpy, independently of any
case table; ``mesh_checks`` measures what the mathematics fixes for a
triangle mesh of a closed level set: every directed edge paired with its
reverse (watertight, consistently wound), the Euler characteristic, the area,
the enclosed volume (divergence theorem) and the orientation of the normals
against the field gradient."""
from __future__ import annotations

from collections import Counter

import numpy as np


def crossing_vertices(F, level, origin, spacing, mask=None):
    """Vertices (n, 3) of all lattice edges with F > level at exactly one end;
    with a mask (same shape as F), only the edges of cubes whose far corner
    (+1, +1, +1) is set."""
    F = np.asarray(F, dtype=np.float64)
    nx, ny, nz = F.shape
    above = F > level
    used = [np.zeros((nx, ny, nz), dtype=bool) for _ in range(3)]
    if mask is None:
        cube = np.ones((nx - 1, ny - 1, nz - 1), dtype=bool)
    else:
        cube = np.asarray(mask, dtype=bool)[1:, 1:, 1:]
    for axis in range(3):
        # the 4 edges of each cube along this axis
        for o1 in (0, 1):
            for o2 in (0, 1):
                off = [0, 0, 0]
                others = [a for a in range(3) if a != axis]
                off[others[0]], off[others[1]] = o1, o2
                sl = tuple(slice(off[d], off[d] + (F.shape[d] - 1)) for d in range(3))
                used[axis][sl] |= cube
    out = []
    for axis in range(3):
        a0 = [slice(None)] * 3
        a1 = [slice(None)] * 3
        a0[axis] = slice(0, -1)
        a1[axis] = slice(1, None)
        cross = (above[tuple(a0)] != above[tuple(a1)]) & used[axis][tuple(a0)]
        idx = np.argwhere(cross)
        v0 = F[tuple(a0)][cross]
        v1 = F[tuple(a1)][cross]
        t = (level - v0) / (v1 - v0)
        P = idx.astype(np.float64)
        P[:, axis] += t
        out.append(np.asarray(origin) + P * np.asarray(spacing))
    return np.concatenate(out) if out else np.zeros((0, 3))


def mesh_checks(V, Fc):
    """dict: unpaired (directed edges without their reverse, or repeated),
    euler (V - E + F), area, volume (enclosed, divergence theorem)."""
    V = np.asarray(V, dtype=np.float64)
    Fc = np.asarray(Fc)
    ed = Counter()
    for a, b, c in Fc:
        for u, w in ((a, b), (b, c), (c, a)):
            ed[(int(u), int(w))] += 1
    unpaired = sum(1 for (u, w), k in ed.items() if k != 1 or ed.get((w, u), 0) != 1)
    n = np.cross(V[Fc[:, 1]] - V[Fc[:, 0]], V[Fc[:, 2]] - V[Fc[:, 0]])
    area = 0.5 * np.linalg.norm(n, axis=1).sum()
    volume = np.einsum("ij,ij->", V[Fc[:, 0]], n) / 6.0
    return dict(unpaired=unpaired, euler=len(V) - len(ed) // 2 + len(Fc), area=area, volume=volume,
                normals=n, centroids=V[Fc].mean(1))
