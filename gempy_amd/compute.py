"""``compute_model`` / ``compute_model_at`` of the hip backend.

Dispatch follows gempy/API/compute_API.py:20-114: select the backend through
``BackendTensor.change_backend_gempy`` (RuntimeError when the GPU is missing,
ValueError for a backend this library does not provide), package the input,
then run, per stack (youngest first; fault stacks before the stacks they
offset): assemble -> solve -> evaluate at the surface points (thresholds) ->
evaluate on the grid (scalar field, optional gradient, ids), and finally the
erosion combination of the non-fault stacks into the lith block.
"""
from __future__ import annotations

import ctypes as C
import time
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .config import AvailableBackends, BackendTensor, GemPyEngineConfig
from .data import (GeoModelInput, RegularGrid, Solutions, SolverForm, StackRelationType,
                   StackSolution)
from .engine import DeviceStack, _ptr, _stream, handle


def _form(model: GeoModelInput, cfg: GemPyEngineConfig) -> int:
    if cfg.dtype == "float32":
        return _lib.GK_SOLVE_FP32_REFINE
    if cfg.dtype == "float64":
        return _lib.GK_SOLVE_FP64_LU
    return (_lib.GK_SOLVE_FP64_LU if model.options.solver == SolverForm.FP64_DIRECT
            else _lib.GK_SOLVE_FP32_REFINE)


def _validate(model: GeoModelInput) -> None:
    """Semantic checks done before compute (the reference validates at
    compute_API.py:38-39)."""
    if not model.stacks:
        raise ValueError("model has no structural group")
    for k, st in enumerate(model.stacks):
        if st.n_surfaces == 0:
            raise ValueError(f"stack {st.name!r} has no surface")
        if st.n_orientations == 0 and st.relation != StackRelationType.FAULT and st.n_rest == 0:
            raise ValueError(f"stack {st.name!r} has neither orientations nor rest points")
        if st.n_surfaces > _lib.GK_MAX_THRESHOLDS:
            raise ValueError(f"stack {st.name!r}: at most {_lib.GK_MAX_THRESHOLDS} surfaces")
        for f in st.faults:
            if not (0 <= f < k) or model.stacks[f].relation != StackRelationType.FAULT:
                raise ValueError(f"stack {st.name!r}: fault index {f} must name an earlier fault stack")


def compute_model(model: GeoModelInput, engine_config: Optional[GemPyEngineConfig] = None,
                  eval_precision: int = _lib.GK_F64, keep_scalar: bool = True) -> Solutions:
    """Compute every stack of ``model`` on the GPU and combine the blocks."""
    cfg = engine_config or GemPyEngineConfig()
    _validate(model)
    if cfg.backend in (AvailableBackends.numpy, AvailableBackends.PYTORCH, AvailableBackends.hip):
        BackendTensor.change_backend_gempy(engine_backend=cfg.backend, use_gpu=cfg.use_gpu,
                                           dtype=cfg.dtype, grads=cfg.compute_grads)
    else:
        raise ValueError(f"Backend {cfg.backend} not supported")
    form = _form(model, cfg)
    grid = model.grid
    opts = model.options
    gradient = opts.evaluation_options.compute_scalar_gradient
    center = _center(model)
    t0 = time.perf_counter()
    stack_sols: List[StackSolution] = []
    dev_stacks: List[DeviceStack] = []
    fault_grid = {}         # fault stack index -> fault block on the grid (float64, device)
    fault_custom = {}
    custom = None if model.custom_grid is None else torch.tensor(
        np.asarray(model.custom_grid, dtype=np.float64).reshape(-1, 3), dtype=torch.float64, device="cuda")
    ids_rows, nsurf, custom_rows = [], [], []
    for k, st in enumerate(model.stacks):
        fault_sp = None
        fv_grid = fv_custom = None
        if st.faults:
            cols = []
            for j in st.faults:
                ds = dev_stacks[j]
                pts = torch.tensor(st.sp_xyz, dtype=torch.float64, device="cuda")
                _, _, fid = ds.evaluate_points(pts, precision=_lib.GK_F64, thresholds=ds.thresholds)
                cols.append(fid.double().cpu().numpy())
            fault_sp = np.stack(cols, 1)
            if grid is not None:
                fv_grid = torch.stack([fault_grid[j] for j in st.faults], 1).contiguous()
            if custom is not None:
                fv_custom = torch.stack([fault_custom[j] for j in st.faults], 1).contiguous()
        ds = DeviceStack(st, opts.kernel_options, center, fault_sp)
        w, info = ds.solve(form)
        thr, s_sp = ds.surface_thresholds(w)
        sol = StackSolution(weights=w, scalar_field_at_surface_points=s_sp.cpu().numpy(),
                            thresholds=thr, solver_info=info)
        if grid is not None:
            S, G, ids = ds.evaluate_lattice(grid, w, eval_precision, gradient, thr, fv_grid)
            sol.scalar_field = S if keep_scalar else None
            sol.gradient = G
            sol.ids_block = ids
            if st.relation == StackRelationType.FAULT:
                fault_grid[k] = ids.double()
            else:
                ids_rows.append(ids)
        if custom is not None:
            _, _, cids = ds.evaluate_points(custom, w, _lib.GK_F64, thresholds=thr, fault_values=fv_custom)
            if st.relation == StackRelationType.FAULT:
                fault_custom[k] = cids.double()
            else:
                custom_rows.append(cids)
        if st.relation != StackRelationType.FAULT:
            nsurf.append(st.n_surfaces)
        dev_stacks.append(ds)
        stack_sols.append(sol)
    out = Solutions(stacks=stack_sols)
    if grid is not None and ids_rows:
        out.lith_block = _combine(ids_rows, nsurf, grid.n_points)
        out.fault_blocks = [stack_sols[k].ids_block for k in sorted(fault_grid)]
        if keep_scalar:
            out.scalar_field_matrix = torch.stack([s.scalar_field for s in stack_sols])
    if custom is not None and custom_rows:
        out.custom = _combine(custom_rows, nsurf, custom.shape[0])
    torch.cuda.synchronize()
    out.timings["model_s"] = time.perf_counter() - t0
    return out


def _combine(rows, nsurf, n) -> torch.Tensor:
    lib = _lib.load()
    ids = torch.stack(rows).contiguous()
    lith = torch.empty(n, dtype=torch.int32, device="cuda")
    ns = (C.c_int32 * len(nsurf))(*nsurf)
    _lib.check("gk_combine_lith", lib.gk_combine_lith(handle(0), len(rows), _ptr(ids), n, ns, _ptr(lith),
                                                      _stream()))
    return lith


def _center(model: GeoModelInput):
    if model.grid is not None:
        e = model.grid.extent
        return [(e[2 * d] + e[2 * d + 1]) / 2 for d in range(3)]
    pts = np.concatenate([s.sp_xyz for s in model.stacks] + [s.ori_xyz for s in model.stacks])
    return list((pts.min(0) + pts.max(0)) / 2)


def compute_model_at(model: GeoModelInput, at: np.ndarray,
                     engine_config: Optional[GemPyEngineConfig] = None) -> np.ndarray:
    """Lith ids at arbitrary points (compute_API.py:89-114: sets a custom grid
    and returns ``raw_arrays.custom``)."""
    m = GeoModelInput(stacks=model.stacks, options=model.options, grid=None, name=model.name,
                      custom_grid=np.asarray(at, dtype=np.float64).reshape(-1, 3))
    sol = compute_model(m, engine_config)
    return sol.custom.cpu().numpy().astype(np.float64)
