"""Engine configuration and the tensor-backend registry.

Mirrors the reference's seam: ``GemPyEngineConfig(backend, use_gpu, dtype,
compute_grads)`` (gempy/core/data/gempy_engine_config.py:9-14) and
``BackendTensor.change_backend_gempy(engine_backend, use_gpu, dtype, grads)``
(called at gempy/API/compute_API.py:45-50), with one more backend, ``hip``.
"""
from __future__ import annotations

import enum
import os
from dataclasses import dataclass
from typing import Optional


class AvailableBackends(enum.Enum):
    numpy = "numpy"
    PYTORCH = "PYTORCH"
    hip = "hip"


@dataclass
class GemPyEngineConfig:
    """``backend``: which tensor backend computes the model; ``use_gpu``;
    ``dtype``: None (backend default: 'float64' for the direct solve),
    'float32' selects the FP32-factor + FP64-refinement solve form;
    ``compute_grads``: tape the inputs for differentiation."""
    backend: AvailableBackends = AvailableBackends.hip
    use_gpu: bool = os.environ.get("GEMPY_USE_GPU", "True").lower() not in ("0", "false", "no")
    dtype: Optional[str] = None
    compute_grads: bool = False


class BackendTensor:
    """Global tensor-backend state (engine_backend, dtype, device, grads).

    Only the ``hip`` backend is implemented by this library; numpy and PyTorch
    belong to the reference's engine and are not provided here.
    """
    engine_backend: Optional[AvailableBackends] = None
    dtype: str = "float64"
    device: str = "cpu"
    COMPUTE_GRADS: bool = False
    PYKEOPS: bool = False

    @classmethod
    def change_backend_gempy(cls, engine_backend: AvailableBackends, use_gpu: bool = True,
                             dtype: Optional[str] = None, grads: bool = False) -> None:
        if engine_backend != AvailableBackends.hip:
            raise ValueError(f"Backend {engine_backend} is not provided by gempy_amd "
                             "(it implements the 'hip' backend only)")
        if not use_gpu:
            raise RuntimeError("The hip backend runs on the GPU only (use_gpu=False requested)")
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("The hip backend needs a ROCm GPU; none is visible")
        from . import _lib
        _lib.load()
        cls.engine_backend = engine_backend
        cls.dtype = dtype or "float64"
        if cls.dtype not in ("float32", "float64"):
            raise ValueError(f"dtype must be 'float32' or 'float64', got {dtype!r}")
        cls.device = "cuda"
        cls.COMPUTE_GRADS = bool(grads)
        cls.PYKEOPS = False
