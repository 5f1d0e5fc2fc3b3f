"""Timed workloads beside the headline benchmark: the multi-stack model with
a fault and octree levels, and the adjoint of a misfit."""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib, synthetic
from .compute import _Model, compute_model
from .config import GemPyEngineConfig
from .engine import DeviceStack


def _sync_time(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def multistack_octree(steps=1, warmup=1, resolution=200, levels=4):
    """Fault + 3 stacks (5,000 points / 1,000 orientations), dense grid and
    `levels` octree levels (base resolution / 2^(levels-1))."""
    m = synthetic.config2(resolution=(resolution,) * 3, octree_levels=levels)
    t, sol = _sync_time(lambda: compute_model(m, GemPyEngineConfig(dtype="float32"), keep_scalar=False),
                        steps, warmup)
    return {"model_s": t, "grid_points": m.grid.n_points,
            "octree_level_voxels": [int(o.voxel_index.numel()) for o in sol.octrees_output],
            "stacks": [{"N": int(3 * s.n_orientations + s.n_rest + 3 + len(s.faults)),
                        "form": sol.stacks[i].solver_info["form"],
                        "iterations": sol.stacks[i].solver_info["iterations"],
                        "fallback": sol.stacks[i].solver_info["fallback"]} for i, s in enumerate(m.stacks)]}


def adjoint_misfit(steps=1, warmup=1, resolution=128, n_obs=10_000):
    """Gradient of sum (S(obs) - S_target(obs))^2 w.r.t. the interface
    coordinates and orientation gradients of the faulted 'lower' stack."""
    from .autodiff import scalar_field_at
    model, k, pert, obs = synthetic.config5(resolution=(resolution,) * 3, n_obs=n_obs)
    base = _Model(synthetic.GeoModelInput(stacks=model.stacks[:1], options=model.options, grid=model.grid),
                  _lib.GK_SOLVE_FP64_CHOL)           # the fault stack only
    fault = base.stacks[0]
    ko = model.options.kernel_options
    center = [500.0, 500.0, 500.0]
    o = torch.tensor(obs, dtype=torch.float64, device="cuda")

    def fault_at(xyz):
        _, _, ids = fault.evaluate_points(torch.as_tensor(xyz, dtype=torch.float64, device="cuda"),
                                          thresholds=fault.thresholds)
        return ids.double().cpu().numpy()[:, None]
    f_obs = fault_at(obs)
    tgt = scalar_field_at(torch.tensor(pert.sp_xyz), torch.tensor(pert.ori_grad), pert, ko, o, center,
                          fault_at(pert.sp_xyz), f_obs).detach()
    st = model.stacks[k]
    f_sp = fault_at(st.sp_xyz)
    grid = model.grid

    def fwd_bwd():
        sp = torch.tensor(st.sp_xyz, dtype=torch.float64, requires_grad=True)
        g = torch.tensor(st.ori_grad, dtype=torch.float64, requires_grad=True)
        S = scalar_field_at(sp, g, st, ko, o, center, f_sp, f_obs)
        loss = ((S - tgt) ** 2).sum()
        loss.backward()
        return float(loss), sp.grad, g.grad
    t, (loss, gsp, gg) = _sync_time(fwd_bwd, steps, warmup)
    # forward grid evaluation of the same stack (128^3) for the record
    ds = DeviceStack(st, ko, center, f_sp)
    w, _ = ds.solve(_lib.GK_SOLVE_FP64_CHOL)
    fv = torch.tensor(fault_at(grid.values()), device="cuda")
    tg, _ = _sync_time(lambda: ds.evaluate_lattice(grid, w, _lib.GK_F64, False, None, fv), 1, 1)
    return {"fwd_bwd_s": t, "loss": loss, "n_params": int(gsp.numel() + gg.numel()),
            "grad_norm_sp": float(gsp.norm()), "grad_norm_G": float(gg.norm()), "grid_eval_s": tg,
            "grid_points": grid.n_points, "n_obs": n_obs}


def learned_gmres(n_points=100_000, n_ori=10_000, judge_points=None, tol=1e-9, budget=200,
                  weights=None, grid_res=256, eval_grid=True):
    """Config 4.  Full system (N ~ 130k): matrix-free FP32-Arnoldi GMRES
    preconditioned by the network's sparse approximate inverse alone, against
    the same GMRES without it (identity blocks), for a fixed iteration budget
    (a purely local inverse cannot remove the smooth error of this dense
    kernel matrix, so neither reaches 1e-6 at this size); then the learned
    form end to end (that attempt, stagnation, the FP64 direct solve: the
    system is solved) and the 256^3 grid evaluated with its weights.  Judge: a
    ~35k sub-sample of the same model, solved by the FP64 direct (Cholesky)
    form, the shifted FP32 factor alone, the factor with the network's
    correction and with the exact local inverse; field error on 200k points
    against the direct solve."""
    from . import learned
    net = learned.SpaiNetwork.load(weights or learned.WEIGHTS)
    m = synthetic.config4(n_points=n_points, n_ori=n_ori, resolution=(grid_res,) * 3)
    st, ko = m.stacks[0], m.options.kernel_options
    center = [500.0, 500.0, 500.0]
    out = {}
    ds = DeviceStack(st, ko, center)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    idx, y = learned.spai_blocks(net, ds)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    full = {"N": ds.N, "inference_s": t_inf}
    for name, blocks in (("network", (idx, y)), ("identity", learned.identity_blocks(ds))):
        t0 = time.perf_counter()
        w, info = learned.solve_spai(ds, *blocks, tol=tol, max_iter=budget, krylov_precision=_lib.GK_F32,
                                     no_fallback=True)
        torch.cuda.synchronize()
        full[name] = {"solve_s": time.perf_counter() - t0, **info}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w, info = learned.solve_learned(ds)
    torch.cuda.synchronize()
    full["learned_form"] = {"solve_s": time.perf_counter() - t0, **{k: v for k, v in info.items() if k != "learned"},
                            "network_attempt": info.get("learned")}
    out["full"] = full
    if eval_grid:
        ds.evaluate_lattice(m.grid, w, _lib.GK_F64, want_scalar=True)      # warm-up of the evaluation path
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        S, _, _ = ds.evaluate_lattice(m.grid, w, _lib.GK_F64)
        torch.cuda.synchronize()
        out["grid_eval_s"] = time.perf_counter() - t0
        out["grid_points"] = m.grid.n_points
        out["grid_gpts_s"] = m.grid.n_points / out["grid_eval_s"] / 1e9
        del S
    nj = judge_points or (20_000, 5_000)
    sub = synthetic.config4(n_points=nj[0], n_ori=nj[1], resolution=(8, 8, 8)).stacks[0]
    dj = DeviceStack(sub, ko, center)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wd, infod = dj.solve(_lib.GK_SOLVE_FP64_CHOL)
    torch.cuda.synchronize()
    judge = {"N": dj.N, "direct": {"solve_s": time.perf_counter() - t0, **infod}}
    X = torch.tensor(np.random.default_rng(0).uniform(0, 1000, (200_000, 3)), device="cuda")
    Sd, _, _ = dj.evaluate_points(X, wd, _lib.GK_F64)
    rng = float(Sd.max() - Sd.min())
    for name, fn in (("factor alone", lambda: (None, None)), ("factor + network", lambda: learned.spai_blocks(net, dj)),
                     ("factor + exact local inverse", lambda: learned.local_inverse_blocks(dj))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bi, by = fn()
        wl, infol = learned.solve_spai(dj, bi, by, tol=tol, max_iter=400, krylov_precision=_lib.GK_F64,
                                       with_factor=True)
        torch.cuda.synchronize()
        Sl, _, _ = dj.evaluate_points(X, wl, _lib.GK_F64)
        judge[name] = {"solve_s": time.perf_counter() - t0, **infol,
                       "max_rel_err_vs_direct": float((Sl - Sd).abs().max()) / rng}
    out["judge"] = judge
    return out
