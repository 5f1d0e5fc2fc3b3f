"""``compute_model`` / ``compute_model_at`` of the hip backend.

Dispatch follows gempy/API/compute_API.py:20-114: select the backend through
``BackendTensor.change_backend_gempy`` (RuntimeError when the GPU is missing,
ValueError for a backend this library does not provide), package the input,
then run, per stack (youngest first; fault stacks before the stacks they
offset): assemble -> solve -> evaluate at the surface points (thresholds) ->
evaluate on the grid (scalar field, optional gradient, ids), and finally the
erosion combination of the non-fault stacks into the lith block.  With more
than one octree level the grid is also evaluated level by level on the
voxels whose probe ids differ (``Solutions.octrees_output``).
"""
from __future__ import annotations

import ctypes as C
import time
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .config import AvailableBackends, BackendTensor, GemPyEngineConfig
from .data import (GeoModelInput, OctreeLevel, Solutions, SolverForm, StackRelationType,
                   StackSolution)
from .engine import DeviceStack, _ptr, _stream, handle

# GK_SOLVE_* of each solver form; the learned form is driven from the host
# (network-preconditioned matrix-free GMRES, then the FP32 form on stagnation)
GMRES_LEARNED = -1
_FORMS = {SolverForm.FP64_DIRECT: _lib.GK_SOLVE_FP64_LU, SolverForm.FP32_REFINE: _lib.GK_SOLVE_FP32_REFINE,
          SolverForm.FP64_CHOL: _lib.GK_SOLVE_FP64_CHOL, SolverForm.GMRES_LEARNED: GMRES_LEARNED}


def _form(model: GeoModelInput, cfg: GemPyEngineConfig) -> int:
    """engine dtype 'float32' -> FP32 factor + FP64 refinement; 'float64' -> the
    options' FP64 form (Cholesky + Schur unless the LU is asked for)."""
    if cfg.dtype == "float32":
        return GMRES_LEARNED if model.options.solver == SolverForm.GMRES_LEARNED else _lib.GK_SOLVE_FP32_REFINE
    if cfg.dtype == "float64":
        return (_lib.GK_SOLVE_FP64_LU if model.options.solver == SolverForm.FP64_DIRECT
                else _lib.GK_SOLVE_FP64_CHOL)
    return _FORMS[model.options.solver]


def _validate(model: GeoModelInput) -> None:
    """Semantic checks done before compute (the reference validates at
    compute_API.py:38-39)."""
    if not model.stacks:
        raise ValueError("model has no structural group")
    for k, st in enumerate(model.stacks):
        if st.n_surfaces == 0:
            raise ValueError(f"stack {st.name!r} has no surface")
        if st.n_orientations == 0 and st.n_rest == 0:
            raise ValueError(f"stack {st.name!r} has neither orientations nor rest points")
        if st.n_surfaces > _lib.GK_MAX_THRESHOLDS:
            raise ValueError(f"stack {st.name!r}: at most {_lib.GK_MAX_THRESHOLDS} surfaces")
        for f in st.faults:
            if not (0 <= f < k) or model.stacks[f].relation != StackRelationType.FAULT:
                raise ValueError(f"stack {st.name!r}: fault index {f} must name an earlier fault stack")
        if st.finite_fault is not None and st.relation != StackRelationType.FAULT:
            raise ValueError(f"stack {st.name!r}: finite fault data on a stack that is not a fault")
    if model.grid is not None:
        L = model.options.evaluation_options.number_octree_levels
        if L > 1 and any(r % (1 << (L - 1)) for r in model.grid.resolution):
            raise ValueError("with octree levels the grid resolution must be base * 2^(levels-1)")


class _Model:
    """The solved stacks of one model on the device."""

    def __init__(self, model: GeoModelInput, form: int):
        self.model = model
        self.stacks: List[DeviceStack] = []
        self.sols: List[StackSolution] = []
        center = _center(model)
        slope = float(model.options.sigmoid_slope)
        for k, st in enumerate(model.stacks):
            fault_sp = None
            if st.faults:
                pts = torch.tensor(st.sp_xyz, dtype=torch.float64, device="cuda")
                fault_sp = np.stack([self._fault_ids(j, pts).cpu().numpy() for j in st.faults], 1)
            ds = DeviceStack(st, model.options.kernel_options, center, fault_sp)
            x0 = None
            if model.weights is not None and k < len(model.weights) and model.weights[k] is not None \
                    and np.asarray(model.weights[k]).size == ds.N:
                x0 = model.weights[k]                 # warm start from the previous solution
            if form == GMRES_LEARNED:
                from .learned import solve_learned
                w, info = solve_learned(ds)
            else:
                w, info = ds.solve(form, x0=x0)
            info["warm_start"] = x0 is not None
            ko = model.options.kernel_options
            if ko.compute_condition_number:
                if info["form"] not in (_lib.GK_SOLVE_FP64_CHOL, _lib.GK_SOLVE_FP32_REFINE):
                    ds.solve(_lib.GK_SOLVE_FP64_CHOL)         # a kept factor for the inverse iteration
                info["condition_number"], info["condition"] = ds.condition_number()
                # one value on the options, as the reference keeps it: the last stack's
                ko.condition_number = info["condition_number"]
            thr, s_sp = ds.surface_thresholds(w)
            self.stacks.append(ds)
            self.sols.append(StackSolution(weights=w, scalar_field_at_surface_points=s_sp.cpu().numpy(),
                                           thresholds=thr, solver_info=info, sigmoid_slope=slope))

    def _fault_ids(self, j, pts):
        """Block values of fault stack j at pts; a fault offset by older faults
        is evaluated with their block values at the same points."""
        ds, st = self.stacks[j], self.model.stacks[j]
        fv = (torch.stack([self._fault_ids(i, pts) for i in st.faults], 1).contiguous() if st.faults else None)
        _, _, fid = ds.evaluate_points(pts, precision=_lib.GK_F64, thresholds=ds.thresholds, fault_values=fv)
        return self._fault_values(j, fid, pts=pts)

    def _fault_values(self, j, ids, pts=None, lattice=None):
        """Drift values of fault stack j from its block at the points: the ids,
        or, for a finite fault, (id - 1) tapered by its implicit function."""
        ff = self.model.stacks[j].finite_fault
        if ff is None:
            return ids.double()
        # a copy: the block itself keeps its ids
        v = ids.to(torch.float64, copy=True).contiguous()
        return ff.taper(self.stacks[j].h, v, pts=pts.contiguous() if pts is not None else None, lattice=lattice)

    def at_points(self, pts: torch.Tensor, precision=_lib.GK_F64, keep_scalar=False, per_stack=None):
        """Ids of every stack, the combined lith block and the fault blocks at
        pts.  per_stack (a list): receives (S, ids, mask) of every stack."""
        ids_rows, nsurf, rels, faults, fault_of, scalars, rows_of = [], [], [], [], {}, [], []
        per = []
        for k, (st, ds) in enumerate(zip(self.model.stacks, self.stacks)):
            fv = (torch.stack([fault_of[j] for j in st.faults], 1).contiguous() if st.faults else None)
            S, _, ids = ds.evaluate_points(pts, precision=precision, thresholds=ds.thresholds, fault_values=fv)
            if keep_scalar:
                scalars.append(S)
            per.append([S, ids, None])
            if st.relation == StackRelationType.FAULT:
                fault_of[k] = self._fault_values(k, ids, pts=pts)
                faults.append(ids)
            else:
                ids_rows.append(ids)
                nsurf.append(st.n_surfaces)
                rels.append(st.relation)
                rows_of.append(k)
        masks = None
        lith = None
        if ids_rows:
            lith, masks = _combine(ids_rows, nsurf, pts.shape[0], rels, want_masks=per_stack is not None)
        if per_stack is not None:
            for i, k in enumerate(rows_of):
                per[k][2] = masks[i]
            per_stack.extend(per)
        return lith, faults, scalars


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
    t0 = time.perf_counter()
    grid, opts = model.grid, model.options
    gradient = opts.evaluation_options.compute_scalar_gradient
    m = _Model(model, _form(model, cfg))
    out = Solutions(stacks=m.sols)
    if grid is not None:
        ids_rows, nsurf, rels, rows_of, fault_grid = [], [], [], [], {}
        for k, (st, ds, sol) in enumerate(zip(model.stacks, m.stacks, m.sols)):
            fv = torch.stack([fault_grid[j] for j in st.faults], 1).contiguous() if st.faults else None
            S, G, ids = ds.evaluate_lattice(grid, sol.weights, eval_precision, gradient, sol.thresholds, fv)
            sol.scalar_field = S if keep_scalar else None
            sol.gradient = G
            sol.ids_block = ids
            if st.relation == StackRelationType.FAULT:
                fault_grid[k] = m._fault_values(k, ids, lattice=(grid.origin, grid.spacing, grid.resolution))
                sol.mask = torch.ones(grid.n_points, dtype=torch.uint8, device="cuda")
            else:
                ids_rows.append(ids)
                nsurf.append(st.n_surfaces)
                rels.append(st.relation)
                rows_of.append(k)
        if ids_rows:
            out.lith_block, masks = _combine(ids_rows, nsurf, grid.n_points, rels, want_masks=True)
            for i, k in enumerate(rows_of):
                m.sols[k].mask = masks[i]
        out.fault_blocks = [m.sols[k].ids_block for k in sorted(fault_grid)]
        if keep_scalar:
            out.scalar_field_matrix = torch.stack([s.scalar_field for s in m.sols])
        levels = opts.evaluation_options.number_octree_levels
        if levels > 1:
            out.octrees_output = compute_octree(m, levels)
    if model.custom_grid is not None:
        pts = torch.tensor(np.asarray(model.custom_grid, dtype=np.float64).reshape(-1, 3), dtype=torch.float64,
                           device="cuda")
        out.custom, _, _ = m.at_points(pts)
    if model.geophysics is not None:
        from .gravity import gravity_of_model
        out.gravity = gravity_of_model(m, model.geophysics)
    torch.cuda.synchronize()
    out.timings["model_s"] = time.perf_counter() - t0
    return out


def compute_octree(m: _Model, levels: int) -> List[OctreeLevel]:
    """Level-by-level evaluation from the base grid (resolution / 2^(levels-1))
    down to the finest level, probing centre + corners of each active voxel."""
    lib = _lib.load()
    grid = m.model.grid
    ext = (C.c_double * 6)(*grid.extent)
    res = [r >> (levels - 1) for r in grid.resolution]
    vox = torch.arange(res[0] * res[1] * res[2], dtype=torch.int64, device="cuda")
    out: List[OctreeLevel] = []
    for lvl in range(levels):
        n = int(vox.numel())
        cres = (C.c_int64 * 3)(*res)
        last = lvl == levels - 1
        npts = n if last else 9 * n
        pts = torch.empty((max(npts, 1), 3), dtype=torch.float64, device="cuda")
        _lib.check("gk_octree_points", lib.gk_octree_points(handle(0), ext, cres, _ptr(vox), n, int(last),
                                                            _ptr(pts), _stream()))
        pts = pts[:npts]
        per: list = []
        lith, faults, scal = m.at_points(pts, keep_scalar=True, per_stack=per)
        step = 1 if last else 9                 # voxel centres: every 9th probe point
        outs = [StackSolution(weights=sol.weights, scalar_field_at_surface_points=sol.scalar_field_at_surface_points,
                              thresholds=sol.thresholds, scalar_field=S[0::step], ids_block=ids[0::step],
                              solver_info=sol.solver_info, sigmoid_slope=sol.sigmoid_slope,
                              mask=None if msk is None else msk[0::step])
                for sol, (S, ids, msk) in zip(m.sols, per)]
        if last:
            out.append(OctreeLevel(tuple(res), vox, lith, scal[-1] if scal else None, None, outs))
            break
        fields = torch.stack([lith] + faults).contiguous() if lith is not None else torch.stack(faults)
        refine = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
        children = torch.empty(max(8 * n, 1), dtype=torch.int64, device="cuda")
        nch = C.c_int64(0)
        _lib.check("gk_octree_refine", lib.gk_octree_refine(handle(0), fields.shape[0], _ptr(fields), n, _ptr(vox),
                                                            cres, _ptr(refine), _ptr(children), C.byref(nch),
                                                            _stream()))
        centre_scalar = scal[-1][0::9] if scal else None
        out.append(OctreeLevel(tuple(res), vox, lith[0::9] if lith is not None else None, centre_scalar,
                               refine[:n].bool(), outs))
        vox = children[:nch.value].clone()
        res = [2 * r for r in res]
    return out


def _combine(rows, nsurf, n, relations=None, want_masks=False):
    """Lith block of the non-fault stacks (youngest first) and, with
    want_masks, each stack's mask of the points it holds."""
    lib = _lib.load()
    ids = torch.stack(rows).contiguous()
    lith = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    masks = torch.empty((len(rows), max(n, 1)), dtype=torch.uint8, device="cuda") if want_masks else None
    ns = (C.c_int32 * len(nsurf))(*nsurf)
    rel = [_lib.GK_REL_ONLAP if r == StackRelationType.ONLAP else _lib.GK_REL_ERODE
           for r in (relations or [StackRelationType.ERODE] * len(rows))]
    rl = (C.c_int32 * len(rel))(*rel)
    _lib.check("gk_combine_stacks", lib.gk_combine_stacks(handle(0), len(rows), _ptr(ids), n, ns, rl, _ptr(lith),
                                                          _ptr(masks), _stream()))
    return lith[:n], (None if masks is None else masks[:, :n])


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
