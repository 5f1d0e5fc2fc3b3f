"""Finite faults: the fault block's drift values tapered by an implicit
function (the reference's ``FiniteFaultData(implicit_function,
implicit_function_transform, pivot)``, gempy/core/data/__init__.py:20, and
``gp.implicit_functions.ellipsoid_3d_factory``,
gempy/modules/custom_implicit_functions/ellipsoid_implicit_function.py:5-35).

A fault stack carrying ``finite_fault`` enters the stacks it offsets with the
drift values (id - 1) f(x') instead of its ids, x' = R^T (x - pivot) + pivot
(the transform's inverse rotation about the pivot): the offset fades out with
f, and f == 1 reproduces the infinite fault (a constant shift of a drift column
changes no id).  The taper is applied on the GPU (gk_finite_fault_values); the
numpy ``__call__`` / ``apply`` below are host helpers for inspection."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib
from .adapters import Transform


@dataclass
class EllipsoidFunction:
    """1 - sigmoid(prod(k) (sum_d ((x_d - c_d) / (r_d k_d))^2 - 1))."""
    center: np.ndarray
    radius: np.ndarray
    max_slope: np.ndarray

    def __post_init__(self):
        self.center = np.broadcast_to(np.asarray(self.center, dtype=np.float64), (3,)).copy()
        self.radius = np.broadcast_to(np.asarray(self.radius, dtype=np.float64), (3,)).copy()
        self.max_slope = np.broadcast_to(np.asarray(self.max_slope, dtype=np.float64), (3,)).copy()

    def __call__(self, xyz: np.ndarray) -> np.ndarray:
        d = np.asarray(xyz, dtype=np.float64).reshape(-1, 3) - self.center
        v = ((d / (self.radius * self.max_slope)) ** 2).sum(1) - 1.0
        with np.errstate(over="ignore"):        # exp -> inf gives f = 1
            return 1.0 - 1.0 / (1.0 + np.exp(-v * np.prod(self.max_slope)))


def ellipsoid_3d_factory(center, radius, max_slope) -> EllipsoidFunction:
    return EllipsoidFunction(center, radius, max_slope)


@dataclass
class FiniteFaultData:
    implicit_function: EllipsoidFunction
    implicit_function_transform: Transform = field(default_factory=Transform)
    pivot: np.ndarray = field(default_factory=lambda: np.zeros(3))

    def __post_init__(self):
        self.pivot = np.asarray(self.pivot, dtype=np.float64).reshape(3)
        if not isinstance(self.implicit_function, EllipsoidFunction):
            raise NotImplementedError("finite faults support the ellipsoid implicit function")
        self.implicit_function_transform._pivot_only()

    def apply(self, points: np.ndarray) -> np.ndarray:
        """f at the points (host helper)."""
        return self.implicit_function(self.implicit_function_transform.apply_inverse_with_pivot(points, self.pivot))

    def as_oracle_dict(self) -> dict:
        f = self.implicit_function
        return {"rotation": self.implicit_function_transform.rotation_matrix(), "pivot": self.pivot,
                "center": f.center, "radii": f.radius, "k": f.max_slope}

    def taper(self, h, values: torch.Tensor, pts: Optional[torch.Tensor] = None, lattice=None) -> torch.Tensor:
        """values (device, FP64, one per point) -> (values - 1) f(x'), in place;
        points: pts (n, 3) on the device, or lattice = (origin, spacing, res)."""
        from .engine import _ptr, _stream
        lib = _lib.load()
        dbl = lambda a, n: (C.c_double * n)(*[float(v) for v in np.asarray(a, dtype=np.float64).ravel()])  # noqa: E731
        f = self.implicit_function
        lat = None
        if pts is None:
            o, sp, res = lattice
            lat = dbl(list(o) + list(sp) + [float(r) for r in res], 9)
        _lib.check("gk_finite_fault_values", lib.gk_finite_fault_values(
            h, int(values.numel()), _ptr(pts) if pts is not None else None, lat,
            dbl(self.implicit_function_transform.rotation_matrix(), 9), dbl(self.pivot, 3), dbl(f.center, 3),
            dbl(f.radius, 3), dbl(f.max_slope, 3), _ptr(values), _stream()))
        return values
