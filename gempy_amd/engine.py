"""Device engine of the ``hip`` backend: one stack's data resident in HBM and
the assemble / solve / evaluate calls into libgkhip.so.

Layout in HBM per stack (FP64, row-major): orientation positions and
gradients (n_o x 3), rest points (R x 3) with their surface index (int32) and
increment nugget, reference points (S x 3), fault-drift increments (R x F).
The solution vector is FP64 of length N = 3 n_o + R + L + F.  Regular grids are
never materialised: the evaluation kernels regenerate lattice coordinates.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .data import KernelOptions, RegularGrid, StackInput

_HANDLES = {}


def handle(device: int = 0):
    """Per-device library handle (created once per process)."""
    lib = _lib.load()
    if device not in _HANDLES:
        h = C.c_void_p()
        _lib.check("gk_create", lib.gk_create(device, C.byref(h)))
        _HANDLES[device] = h
    return _HANDLES[device]


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else C.c_void_p(0)


def split_ref_rest(counts: Sequence[int]):
    counts = np.asarray(counts, dtype=np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
    n = int(counts.sum())
    mask = np.ones(n, dtype=bool)
    mask[starts] = False
    rest_idx = np.nonzero(mask)[0]
    surf = np.repeat(np.arange(len(counts)), counts)[rest_idx].astype(np.int32)
    return starts, rest_idx, surf


class DeviceStack:
    """One stack's co-kriging data in HBM and its gk_system descriptor."""

    # incremented by every solve: the library keeps the factorization of the
    # last solve of the handle, which the adjoint reuses
    solve_epoch = 0

    def __init__(self, stack: StackInput, kopts: KernelOptions, center: Sequence[float],
                 fault_sp: Optional[np.ndarray] = None, device: int = 0):
        if any(c < 1 for c in stack.counts):
            raise ValueError("every surface needs at least one point (its reference point)")
        self.stack = stack
        self.kopts = kopts
        self.dev = torch.device("cuda", device)
        self.h = handle(device)
        ref_idx, rest_idx, surf = split_ref_rest(stack.counts)
        f64 = dict(dtype=torch.float64, device=self.dev)
        sp = stack.sp_xyz
        self.ref_idx, self.rest_idx = ref_idx, rest_idx
        self.ori_xyz = torch.tensor(stack.ori_xyz, **f64).contiguous()
        self.ori_grad = torch.tensor(stack.ori_grad, **f64).contiguous()
        self.ori_nugget = torch.tensor(stack.ori_nugget, **f64).contiguous()
        self.rest_xyz = torch.tensor(sp[rest_idx], **f64).contiguous()
        self.rest_surf = torch.tensor(surf, dtype=torch.int32, device=self.dev)
        nug = stack.sp_nugget[rest_idx] + stack.sp_nugget[ref_idx][surf]
        self.rest_nugget = torch.tensor(nug, **f64)
        self.ref_xyz = torch.tensor(sp[ref_idx], **f64).contiguous()
        self.sp_xyz = torch.tensor(sp, **f64).contiguous()
        # fault columns: a fault whose block is the same at every rest point and
        # its reference point leaves an all-zero drift column (a singular
        # system) and says nothing about this stack; it is dropped (weight 0)
        self.n_fault_in = 0 if fault_sp is None else int(np.asarray(fault_sp).reshape(len(sp), -1).shape[1])
        self.fault_cols = np.zeros(0, dtype=np.int64)
        self.fault_rest = None
        if self.n_fault_in:
            fs = np.asarray(fault_sp, dtype=np.float64).reshape(len(sp), -1)
            inc = fs[rest_idx] - fs[ref_idx][surf]
            self.fault_cols = np.nonzero(np.any(inc != 0.0, axis=0))[0]
            fs, inc = fs[:, self.fault_cols], inc[:, self.fault_cols]
            self.fault_rest = torch.tensor(inc, **f64).contiguous()
            self.fault_sp = torch.tensor(fs, **f64).contiguous()
        self.n_fault = int(len(self.fault_cols))
        s = _lib.GkSystem()
        s.n_ori = stack.n_orientations
        s.n_rest = len(rest_idx)
        s.n_surf = len(ref_idx)
        s.uni_degree = int(kopts.uni_degree)
        s.n_fault = self.n_fault
        s.range = float(kopts.range)
        s.c_o = float(kopts.c_o)
        s.i_res = float(kopts.i_res)
        s.gi_res = float(kopts.gi_res)
        s.ori_xyz, s.ori_grad, s.ori_nugget = _ptr(self.ori_xyz), _ptr(self.ori_grad), _ptr(self.ori_nugget)
        s.rest_xyz, s.rest_surf, s.rest_nugget = _ptr(self.rest_xyz), _ptr(self.rest_surf), _ptr(self.rest_nugget)
        s.ref_xyz = _ptr(self.ref_xyz)
        s.fault_rest = _ptr(self.fault_rest)
        for k in range(3):
            s.center[k] = float(center[k])
        self.sys = s
        lib = _lib.load()
        self.N = int(lib.gk_system_n(C.byref(s)))
        self.weights: Optional[torch.Tensor] = None
        self.thresholds: Optional[np.ndarray] = None

    # ------------------------------------------------------------------ solve
    def solve(self, form: int = _lib.GK_SOLVE_FP32_REFINE, tol: float = 1e-6, max_iter: int = 300,
              x0: Optional[torch.Tensor] = None):
        """Solve the stack's system; x0 (length N): initial guess of the
        FP32 form's refinement (e.g. the weights of a previous solve).  tol:
        relative FP64 residual at which the FP32 form's refinement stops (at
        1e-6 the N = 35k precision workload's field is within 2e-8 of the FP64
        direct solve's, relative to its range); it also stops earlier at the
        residual floor of the FP64 operator."""
        lib = _lib.load()
        x = torch.empty(self.N, dtype=torch.float64, device=self.dev)
        info = _lib.GkSolveInfo()
        if x0 is not None:
            x0 = torch.as_tensor(x0, dtype=torch.float64, device=self.dev).reshape(-1).contiguous()
            if x0.numel() != self.N:
                raise ValueError(f"initial guess of length {x0.numel()}, the system has {self.N} unknowns")
            _lib.check("gk_solve_x0", lib.gk_solve_x0(self.h, C.byref(self.sys), int(form), float(tol),
                                                      int(max_iter), _ptr(x0), _ptr(x), C.byref(info), _stream()))
        else:
            _lib.check("gk_solve", lib.gk_solve(self.h, C.byref(self.sys), int(form), float(tol),
                                                 int(max_iter), _ptr(x), C.byref(info), _stream()))
        DeviceStack.solve_epoch += 1
        self.weights = x
        return x, {"form": int(info.form), "iterations": int(info.iterations),
                   "fallback": int(info.fallback), "info": int(info.info),
                   "relres": float(info.relres), "t_assemble_ms": float(info.t_assemble_ms),
                   "t_factor_ms": float(info.t_factor_ms), "t_refine_ms": float(info.t_refine_ms),
                   "shift": float(info.shift)}

    def condition_number(self, max_steps: int = 120, tol: float = 1e-10):
        """2-norm condition number of the system matrix (needs the factor
        kept by the last solve of this stack: FP64_CHOL or FP32_REFINE)."""
        lib = _lib.load()
        cond = C.c_double()
        sig = (C.c_double * 2)()
        steps = (C.c_int32 * 2)()
        _lib.check("gk_condition_number", lib.gk_condition_number(self.h, C.byref(self.sys), int(max_steps),
                                                                  float(tol), C.byref(cond), sig, steps, _stream()))
        return float(cond.value), {"sigma_max": float(sig[0]), "sigma_min": float(sig[1]),
                                   "lanczos_steps": [int(steps[0]), int(steps[1])]}

    # ------------------------------------------------------------ assembly
    def assemble(self, dtype: int = _lib.GK_F64):
        """Dense matrix (FP64 full system or FP32 equilibrated covariance block)."""
        lib = _lib.load()
        if dtype == _lib.GK_F64:
            lda = int(lib.gk_system_padded(C.byref(self.sys)))
            A = torch.empty((lda, lda), dtype=torch.float64, device=self.dev)
            b = torch.empty(self.N, dtype=torch.float64, device=self.dev)
            _lib.check("gk_assemble", lib.gk_assemble(self.h, C.byref(self.sys), _lib.GK_F64, _ptr(A),
                                                      lda, _ptr(b), None, _stream()))
            return A, b, None
        nc = 3 * self.sys.n_ori + self.sys.n_rest
        lda = max(_lib.GK_BLOCK, -(-nc // _lib.GK_BLOCK) * _lib.GK_BLOCK)
        A = torch.empty((lda, lda), dtype=torch.float32, device=self.dev)
        d = torch.empty(max(nc, 1), dtype=torch.float64, device=self.dev)
        _lib.check("gk_assemble", lib.gk_assemble(self.h, C.byref(self.sys), _lib.GK_F32, _ptr(A), lda,
                                                  None, _ptr(d), _stream()))
        return A, None, d

    def matvec(self, x: torch.Tensor, precision: int = _lib.GK_F64) -> torch.Tensor:
        lib = _lib.load()
        y = torch.empty(self.N, dtype=torch.float64, device=self.dev)
        _lib.check("gk_matvec", lib.gk_matvec(self.h, C.byref(self.sys), int(precision), _ptr(x), _ptr(y),
                                              _stream()))
        return y

    # ------------------------------------------------------------- evaluate
    def used_fault_values(self, fault_values):
        """The columns of the fault-block values (one per fault of the stack)
        that enter this stack's system."""
        if fault_values is None or self.n_fault == 0:
            return None
        fv = fault_values.reshape(fault_values.shape[0], -1)
        if fv.shape[1] == self.n_fault:            # already the used columns
            return fault_values
        if fv.shape[1] != self.n_fault_in:
            raise ValueError(f"fault values with {fv.shape[1]} columns; the stack has {self.n_fault_in} faults")
        cols = torch.as_tensor(self.fault_cols, device=fault_values.device)
        return fv[:, cols].contiguous()

    def _eval(self, desc, weights):
        lib = _lib.load()
        _lib.check("gk_eval", lib.gk_eval(self.h, C.byref(self.sys), _ptr(weights), C.byref(desc), _stream()))

    def evaluate_points(self, xyz: torch.Tensor, weights: Optional[torch.Tensor] = None,
                        precision: int = _lib.GK_F64, gradient: bool = False,
                        thresholds: Optional[Sequence[float]] = None,
                        fault_values: Optional[torch.Tensor] = None):
        w = self.weights if weights is None else weights
        n = xyz.shape[0]
        S = torch.empty(n, dtype=torch.float64, device=self.dev)
        G = torch.empty((n, 3), dtype=torch.float64, device=self.dev) if gradient else None
        ids = torch.empty(n, dtype=torch.int32, device=self.dev) if thresholds is not None else None
        d = _lib.GkEvalDesc()
        d.kind = _lib.GK_POINTS_EXPLICIT
        d.precision = int(precision)
        d.n_points = n
        d.xyz = _ptr(xyz)
        d.scalar, d.gradient, d.ids = _ptr(S), _ptr(G), _ptr(ids)
        fault_values = self.used_fault_values(fault_values)
        d.fault_values = _ptr(fault_values)
        self._set_thresholds(d, thresholds)
        if n:
            self._eval(d, w)
        return S, G, ids

    def evaluate_lattice(self, grid: RegularGrid, weights: Optional[torch.Tensor] = None,
                         precision: int = _lib.GK_F64, gradient: bool = False,
                         thresholds: Optional[Sequence[float]] = None,
                         fault_values: Optional[torch.Tensor] = None, want_scalar: bool = True):
        w = self.weights if weights is None else weights
        n = grid.n_points
        S = torch.empty(n, dtype=torch.float64, device=self.dev)
        G = torch.empty((n, 3), dtype=torch.float64, device=self.dev) if gradient else None
        ids = torch.empty(n, dtype=torch.int32, device=self.dev) if thresholds is not None else None
        d = _lib.GkEvalDesc()
        d.kind = _lib.GK_POINTS_LATTICE
        d.precision = int(precision)
        for k in range(3):
            d.origin[k] = grid.origin[k]
            d.spacing[k] = grid.spacing[k]
            d.res[k] = grid.resolution[k]
        d.scalar, d.gradient, d.ids = _ptr(S), _ptr(G), _ptr(ids)
        fault_values = self.used_fault_values(fault_values)
        d.fault_values = _ptr(fault_values)
        self._set_thresholds(d, thresholds)
        self._eval(d, w)
        return S, G, ids

    @staticmethod
    def _set_thresholds(d, thresholds):
        if thresholds is None:
            d.n_thresholds = 0
            return
        t = list(thresholds)
        if len(t) > _lib.GK_MAX_THRESHOLDS:
            raise ValueError(f"at most {_lib.GK_MAX_THRESHOLDS} surfaces per stack")
        d.n_thresholds = len(t)
        for k, v in enumerate(t):
            d.thresholds[k] = float(v)

    def surface_thresholds(self, weights: Optional[torch.Tensor] = None) -> Tuple[np.ndarray, torch.Tensor]:
        """S at every surface point (FP64) and the thresholds (S at reference points)."""
        fv = self.fault_sp if self.n_fault else None
        S, _, _ = self.evaluate_points(self.sp_xyz, weights, _lib.GK_F64, fault_values=fv)
        thr = S[torch.as_tensor(self.ref_idx, device=self.dev)].cpu().numpy()
        self.thresholds = thr
        return thr, S


def block_values(scalar: torch.Tensor, thresholds, sigmoid_slope: float) -> torch.Tensor:
    """Smooth ids of one stack (gk_block_values) at the points of `scalar`."""
    lib = _lib.load()
    t = [float(v) for v in thresholds]
    if len(t) > _lib.GK_MAX_THRESHOLDS:
        raise ValueError(f"at most {_lib.GK_MAX_THRESHOLDS} surfaces per stack")
    S = scalar.reshape(-1).contiguous()
    out = torch.empty_like(S)
    arr = (C.c_double * max(len(t), 1))(*t)
    _lib.check("gk_block_values", lib.gk_block_values(handle(S.device.index or 0), _ptr(S), S.numel(), len(t), arr,
                                                      float(sigmoid_slope), _ptr(out), _stream()))
    return out
