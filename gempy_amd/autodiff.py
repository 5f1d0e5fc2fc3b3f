"""Differentiation of the co-kriging field with a hand-written adjoint.

The reference differentiates by taping torch autograd through its engine
(gempy/API/compute_API.py:49,66; gradients w.r.t. the surface-point
coordinates and orientation data, test_ad_I.py:41-51,118-123).  Here the
forward is assemble -> FP64 Cholesky + Schur solve -> FP64 evaluation on the
GPU, and the backward is ``gk_adjoint``: one more solve with the kept factor
(lam = A^-1 Phi^T e, Phi^T e being the observation field evaluated at the data
points) and three tiled field-gradient evaluations; no kernel tile is stored.

    dL/dG_g     = lam_G
    dL/dp_a     = i_res Omega_a grad E(p_a) - Lambda_a grad S_w(p_a) - Omega_a grad S_lam(p_a)
    dL/dq_g     = gi_res Hess E(q_g) w_g - (Hess S_w(q_g) lam_g + Hess S_lam(q_g) w_g) / gi_res
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .data import KernelOptions, StackInput
from .engine import DeviceStack, _ptr, _stream


class ScalarFieldAt(torch.autograd.Function):
    """S(obs) of one stack as a differentiable function of its surface-point
    coordinates (n_sp, 3), orientation gradients (n_o, 3) and orientation
    positions (n_o, 3; the reference's dip_positions, test_ad_I.py:48-50)."""

    @staticmethod
    def forward(ctx, sp_xyz: torch.Tensor, ori_grad: torch.Tensor, ori_xyz: torch.Tensor, template: StackInput,
                kopts: KernelOptions, obs: torch.Tensor, center: Sequence[float], fault_sp=None, fault_obs=None):
        """fault_sp (n_sp, F) / fault_obs (n_obs, F): frozen fault-block values
        of the fault stacks offsetting this stack (piecewise constant)."""
        st = template.copy()
        st.sp_xyz = sp_xyz.detach().double().cpu().numpy().reshape(-1, 3)
        st.ori_grad = ori_grad.detach().double().cpu().numpy().reshape(-1, 3)
        st.ori_xyz = ori_xyz.detach().double().cpu().numpy().reshape(-1, 3)
        ds = DeviceStack(st, kopts, center, fault_sp)
        w, _ = ds.solve(_lib.GK_SOLVE_FP64_CHOL)
        obs_d = obs.detach().to(device="cuda", dtype=torch.float64).contiguous()
        fo = None
        if ds.n_fault:
            if fault_obs is None:
                raise ValueError("fault_obs is required when the stack has fault-drift columns")
            fo = torch.as_tensor(np.asarray(fault_obs, dtype=np.float64).reshape(obs_d.shape[0], -1),
                                 device="cuda").contiguous()
            fo = ds.used_fault_values(fo)
        S, _, _ = ds.evaluate_points(obs_d, w, _lib.GK_F64, fault_values=fo)
        ctx.ds, ctx.w, ctx.obs, ctx.fo = ds, w, obs_d, fo
        ctx.epoch = DeviceStack.solve_epoch
        ctx.meta = [(t.device, t.dtype) for t in (sp_xyz, ori_grad, ori_xyz)]
        return S.to(obs.device) if obs.device.type != "cuda" else S

    @staticmethod
    def backward(ctx, gS: torch.Tensor):
        ds: DeviceStack = ctx.ds
        if DeviceStack.solve_epoch != ctx.epoch:        # another solve replaced the kept factor
            ds.solve(_lib.GK_SOLVE_FP64_CHOL)
        lib = _lib.load()
        e = gS.detach().to(device="cuda", dtype=torch.float64).contiguous()
        R, S, no = ds.sys.n_rest, ds.sys.n_surf, ds.sys.n_ori
        d_pts = torch.empty((R + S, 3), dtype=torch.float64, device="cuda")
        d_g = torch.empty((max(no, 1), 3), dtype=torch.float64, device="cuda")
        want_q = ctx.needs_input_grad[2]
        d_q = torch.empty((max(no, 1), 3), dtype=torch.float64, device="cuda") if want_q else None
        _lib.check("gk_adjoint_full", lib.gk_adjoint_full(ds.h, C.byref(ds.sys), _ptr(ctx.w), e.numel(),
                                                          _ptr(ctx.obs), _ptr(ctx.fo), _ptr(e), _ptr(d_pts),
                                                          _ptr(d_g), _ptr(d_q), None, _stream()))
        n_sp = len(ds.ref_idx) + len(ds.rest_idx)
        d_sp = torch.zeros((n_sp, 3), dtype=torch.float64, device="cuda")
        d_sp[torch.as_tensor(ds.rest_idx, device="cuda")] = d_pts[:R]
        d_sp[torch.as_tensor(ds.ref_idx, device="cuda")] = d_pts[R:]
        (dv, dt), (gdv, gdt), (qdv, qdt) = ctx.meta
        return (d_sp.to(device=dv, dtype=dt), d_g[:no].to(device=gdv, dtype=gdt),
                None if d_q is None else d_q[:no].to(device=qdv, dtype=qdt), None, None, None, None, None, None)


def scalar_field_at(sp_xyz, ori_grad, template: StackInput, kopts: KernelOptions, obs, center,
                    fault_sp=None, fault_obs=None, ori_xyz=None):
    """S(obs); ori_xyz (n_o, 3), optional: differentiable orientation
    positions (the template's are used, without a gradient, when omitted)."""
    if ori_xyz is None:
        ori_xyz = torch.tensor(template.ori_xyz, dtype=torch.float64)
    return ScalarFieldAt.apply(sp_xyz, ori_grad, ori_xyz, template, kopts, obs, center, fault_sp, fault_obs)


def misfit_gradient(template: StackInput, kopts: KernelOptions, obs: np.ndarray, target: np.ndarray, center,
                    fault_sp=None, fault_obs=None, positions: bool = False):
    """Loss sum (S(obs) - target)^2 and its gradients w.r.t. the surface-point
    coordinates and the orientation gradients (numpy arrays); with positions,
    also w.r.t. the orientation positions (a fourth array)."""
    sp = torch.tensor(template.sp_xyz, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(template.ori_grad, dtype=torch.float64, requires_grad=True)
    q = torch.tensor(template.ori_xyz, dtype=torch.float64, requires_grad=positions)
    o = torch.tensor(obs, dtype=torch.float64, device="cuda")
    S = scalar_field_at(sp, g, template, kopts, o, center, fault_sp, fault_obs, ori_xyz=q)
    loss = ((S - torch.tensor(target, dtype=torch.float64, device="cuda")) ** 2).sum()
    loss.backward()
    out = (float(loss.detach()), sp.grad.numpy(), g.grad.numpy())
    return out + (q.grad.numpy(),) if positions else out
