"""Gravity forward of a model on centred grids around devices (the
reference's geophysics input: ``set_centered_grid`` + ``calculate_gravity_gradient``
+ densities -> ``Solutions.gravity``; gempy/API/grid_API.py:99-108,
test/test_modules/test_geophysics/test_gravity.py:67-89).

The kernel geometry of the engine's ``CenteredGrid`` is not in the reference
tree; :class:`CenteredGrid` restates GemPy v2's: cells at +-geomspace(0.01, 1,
res/2) * radius and 0 along x and y, at -(concat(0, geomspace(0.01, 1, res)) +
0.005) * radius along z (below the device), half-sizes from the spacings.  The
prism attraction of every cell (tz) and the contraction with the densities of
the model's units at the cells run on the GPU (gk_gravity_tz, gk_gravity)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .engine import _ptr, _stream, handle

G_UGAL = 6.674e-3


class CenteredGrid:
    """Centred grid(s): ``centers`` (n_dev, 3), ``resolution`` (3,), ``radius``
    (scalar or (3,)), in the input (physical) frame.  ``kernel_centers`` are
    relative to a device; ``values`` are all devices' cells (n_dev * n_kernel,
    3); ``transform`` (a Transform, optional) maps them into the model frame
    the stacks are given in."""

    def __init__(self, centers, resolution: Sequence[int], radius, transform=None):
        self.centers = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
        self.resolution = [int(r) for r in resolution]
        self.radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (3,)).copy()
        self.transform = transform
        g2, d = [], []
        for k in range(3):
            if k == 2:
                g = np.geomspace(0.01, 1, self.resolution[k])
                gk = (np.concatenate(([0.0], g)) + 0.005) * -self.radius[k]
            else:
                g = np.geomspace(0.01, 1, self.resolution[k] // 2)
                gk = np.concatenate((-g[::-1], [0.0], g)) * self.radius[k]
            g2.append(gk)
            d.append(np.diff(np.pad(gk, 1, "reflect", reflect_type="odd")))
        flat = lambda M: np.ascontiguousarray(np.stack([m.ravel() for m in M], 1))   # noqa: E731
        self.kernel_centers = flat(np.meshgrid(*g2))
        self.left_voxel_edges = flat(np.meshgrid(d[0][:-1] / 2, d[1][:-1] / 2, d[2][:-1] / 2))
        self.right_voxel_edges = flat(np.meshgrid(d[0][1:] / 2, d[1][1:] / 2, d[2][1:] / 2))

    @property
    def n_kernel(self) -> int:
        return self.kernel_centers.shape[0]

    @property
    def n_devices(self) -> int:
        return self.centers.shape[0]

    @property
    def values(self) -> np.ndarray:
        return (self.centers[:, None, :] + self.kernel_centers[None]).reshape(-1, 3)

    def model_values(self) -> np.ndarray:
        v = self.values
        return self.transform.apply(v) if self.transform is not None else v


@dataclass
class GeophysicsInput:
    """tz of the kernel (from calculate_gravity_gradient) and the density of
    every unit, by lith id (1-based: densities[0] is unit 1)
    (gempy/core/data/geophysics_input.py)."""
    grid: CenteredGrid
    densities: Sequence[float]
    tz: Optional[torch.Tensor] = None


def calculate_gravity_gradient(grid: CenteredGrid, ugal: bool = True) -> torch.Tensor:
    """tz of every kernel cell (device tensor), mGal per g/cm^3 with ugal."""
    lib = _lib.load()
    dev = torch.device("cuda", 0)
    c = torch.tensor(grid.kernel_centers, dtype=torch.float64, device=dev)
    lo = torch.tensor(grid.left_voxel_edges, dtype=torch.float64, device=dev)
    hi = torch.tensor(grid.right_voxel_edges, dtype=torch.float64, device=dev)
    tz = torch.empty(max(grid.n_kernel, 1), dtype=torch.float64, device=dev)
    G = G_UGAL if ugal else 6.674e-11
    _lib.check("gk_gravity_tz", lib.gk_gravity_tz(handle(0), grid.n_kernel, _ptr(c), _ptr(lo), _ptr(hi), G, _ptr(tz),
                                                  _stream()))
    return tz[:grid.n_kernel]


def gravity_from_ids(ids: torch.Tensor, tz: torch.Tensor, densities: Sequence[float], n_devices: int) -> np.ndarray:
    """sum_k tz[k] density(ids[d, k]) per device (ids: lith ids, 1-based)."""
    lib = _lib.load()
    if not ids.is_cuda or not tz.is_cuda:
        raise ValueError("gravity_from_ids: ids and tz must be device tensors")
    ids = ids.reshape(n_devices, -1).to(dtype=torch.int32).contiguous()
    tz = tz.to(torch.float64).contiguous()
    if ids.shape[1] != tz.numel():
        raise ValueError(f"{ids.shape[1]} ids per device, {tz.numel()} kernel cells")
    dens = (C.c_double * max(len(densities), 1))(*[float(v) for v in densities])
    out = torch.empty(max(n_devices, 1), dtype=torch.float64, device=tz.device)
    _lib.check("gk_gravity", lib.gk_gravity(handle(0), n_devices, tz.numel(), _ptr(tz), _ptr(ids), len(densities),
                                            dens, _ptr(out), _stream()))
    return out[:n_devices].cpu().numpy()


def gravity_of_model(m, geo: GeophysicsInput) -> np.ndarray:
    """Gravity of a solved model (compute._Model) at the devices of geo.grid:
    the lith block at the cells (model frame), contracted with tz."""
    if geo.tz is None:
        geo.tz = calculate_gravity_gradient(geo.grid)
    pts = torch.tensor(geo.grid.model_values(), dtype=torch.float64, device="cuda")
    lith, _, _ = m.at_points(pts)
    return gravity_from_ids(lith, geo.tz, geo.densities, geo.grid.n_devices)


def compute_gravity(model, grid: CenteredGrid, densities: Sequence[float], tz: Optional[torch.Tensor] = None,
                    engine_config=None) -> np.ndarray:
    """Gravity at every device of `grid` for a GeoModelInput."""
    from .compute import compute_model
    from .data import GeoModelInput
    m = GeoModelInput(stacks=model.stacks, options=model.options, grid=None, name=model.name,
                      geophysics=GeophysicsInput(grid, densities, tz))
    return compute_model(m, engine_config).gravity
