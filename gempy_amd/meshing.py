"""Surface meshes from the evaluated fields: marching cubes on the GPU
(``gk_marching_cubes``) for every surface of every stack, at the surface's
threshold and restricted to the points the stack holds in the lith block --
what the reference does per structural element with skimage
(gempy/modules/mesh_extranction/marching_cubes.py:12-101)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .data import RegularGrid, Solutions
from .engine import _ptr, _stream, handle


@dataclass
class Mesh:
    """Triangle mesh of one surface (``vertices`` (n, 3), ``edges`` (m, 3)
    vertex indices, as the reference stores them on a structural element)."""
    vertices: np.ndarray
    edges: np.ndarray


def marching_cubes(field: torch.Tensor, grid: RegularGrid, level: float, mask: Optional[torch.Tensor] = None,
                   origin: Optional[Tuple[float, float, float]] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Isosurface ``field == level`` of a lattice field (device tensor of
    grid.n_points values, z fastest).  Vertices at origin + index * spacing
    (origin defaults to the first cell centre); mask (uint8/bool, same
    layout): cubes whose far corner (+1, +1, +1) is set."""
    lib = _lib.load()
    f = field.reshape(-1).to(dtype=torch.float64).contiguous()
    if f.numel() != grid.n_points:
        raise ValueError(f"field has {f.numel()} values, the grid {grid.n_points}")
    m = None
    if mask is not None:
        m = mask.reshape(-1).to(dtype=torch.uint8).contiguous()
        if m.numel() != grid.n_points:
            raise ValueError("mask and grid sizes differ")
    res = (C.c_int64 * 3)(*grid.resolution)
    org = (C.c_double * 3)(*(grid.origin if origin is None else origin))
    sp = (C.c_double * 3)(*grid.spacing)
    nv, nf = C.c_int64(0), C.c_int64(0)
    h = handle(f.device.index or 0)
    _lib.check("gk_marching_cubes", lib.gk_marching_cubes(h, _ptr(f), res, _ptr(m), float(level), org, sp, None, 0,
                                                          None, 0, C.byref(nv), C.byref(nf), _stream()))
    verts = torch.empty((max(nv.value, 1), 3), dtype=torch.float64, device=f.device)
    faces = torch.empty((max(nf.value, 1), 3), dtype=torch.int32, device=f.device)
    _lib.check("gk_marching_cubes", lib.gk_marching_cubes(h, _ptr(f), res, _ptr(m), float(level), org, sp,
                                                          _ptr(verts), verts.shape[0], _ptr(faces), faces.shape[0],
                                                          C.byref(nv), C.byref(nf), _stream()))
    return verts[:nv.value].cpu().numpy(), faces[:nf.value].cpu().numpy()


def compute_meshes(sol: Solutions, grid: RegularGrid, reference_origin: bool = True) -> List[List[Mesh]]:
    """One mesh per surface of every stack (stacks and surfaces in model
    order), each at its threshold and within the stack's mask; sets
    ``sol.dc_meshes`` (flat, element order).  reference_origin: vertices at
    extent_min + index * spacing as the reference places them
    (marching_cubes.py:94-98); otherwise at the cell centres' lattice."""
    origin = (grid.extent[0], grid.extent[2], grid.extent[4]) if reference_origin else None
    out: List[List[Mesh]] = []
    for s in sol.stacks:
        if s.scalar_field is None:
            raise ValueError("meshes need the dense scalar fields (compute_model(..., keep_scalar=True))")
        meshes = []
        for t in np.asarray(s.thresholds, dtype=np.float64):
            v, f = marching_cubes(s.scalar_field, grid, float(t), s.mask, origin)
            meshes.append(Mesh(v, f))
        out.append(meshes)
    sol.dc_meshes = [m for ms in out for m in ms]
    sol._raw = None
    return out
