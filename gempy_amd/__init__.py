"""gempy_amd -- MI355X-native implicit co-kriging engine (the ``hip`` backend).

Public surface (mirrors the reference's compute seam, gempy/API/compute_API.py):
``compute_model``, ``compute_model_at``, ``GemPyEngineConfig``,
``AvailableBackends``, ``BackendTensor`` and the input/output records in
:mod:`gempy_amd.data`.  Numerics run in libgkhip.so (HIP, gfx950).
"""
from .adapters import Transform, model_input_from_geomodel
from .config import AvailableBackends, BackendTensor, GemPyEngineConfig
from .data import (EvaluationOptions, GeoModelInput, InterpolationOptions, KernelOptions,
                   RegularGrid, Solutions, SolverForm, StackInput, StackRelationType)
from .faults import EllipsoidFunction, FiniteFaultData, ellipsoid_3d_factory


def compute_model(*args, **kwargs):
    from .compute import compute_model as _cm
    return _cm(*args, **kwargs)


def compute_model_at(*args, **kwargs):
    from .compute import compute_model_at as _cma
    return _cma(*args, **kwargs)


__all__ = ["AvailableBackends", "BackendTensor", "GemPyEngineConfig", "EvaluationOptions",
           "GeoModelInput", "InterpolationOptions", "KernelOptions", "RegularGrid", "Solutions",
           "SolverForm", "StackInput", "StackRelationType", "Transform", "compute_model", "compute_model_at",
           "model_input_from_geomodel", "EllipsoidFunction", "FiniteFaultData", "ellipsoid_3d_factory"]
