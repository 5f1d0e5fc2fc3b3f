"""Learned sparse approximate inverse for the matrix-free GMRES of large
systems (the reference's iterative solver is plain CG / KeOps,
docs/developers_notes/dev_log/log_2024-05.md:10-50; it has no learned part).

A small network (8 pair features -> 32 -> 32 -> 3x3 block, tanh) predicts,
for every data point P and each of its 16 nearest data points Q, the block
M_PQ of an approximate inverse of the equilibrated covariance block H.  It is
trained on systems generated from seeded models (``synthetic.training_system``,
seeds 1000-2999) to minimise the local SPAI residual ||M_P H_loc - E_P||^2.
Training and inference run in libgkhip.so (knn, features, local blocks, MLP
forward/backward, loss, Adam); the network is used by ``gk_solve_spai``, alone
(matrix-free GMRES) or as the local correction of the shifted FP32 Cholesky
factor (y = P u; y += M (u - H y); y += P (u - H y)), where it removes the local error modes that the
shift leaves.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib, synthetic
from .data import KernelOptions
from .engine import DeviceStack, _ptr, _stream

K = _lib.GK_SPAI_K
F = _lib.GK_SPAI_FEATURES
HIDDEN = _lib.GK_SPAI_HIDDEN
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spai_weights.npz")


def n_params() -> int:
    return int(_lib.load().gk_spai_nparams())


class SpaiNetwork:
    """Parameters (FP32, on the GPU) and Adam state of the preconditioner network."""

    def __init__(self, seed: int = 0):
        rng = np.random.default_rng(seed)
        shapes = [(HIDDEN, F), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (9, HIDDEN), (9,)]
        parts = []
        for shp in shapes:
            if len(shp) == 2:
                parts.append(rng.normal(0, 1.0 / np.sqrt(shp[1]), shp).ravel())
            else:
                parts.append(np.zeros(shp))
        parts[4] *= 0.01                       # start close to the Jacobi preconditioner
        p = np.concatenate(parts).astype(np.float32)
        assert p.size == n_params()
        self.params = torch.tensor(p, device="cuda")
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0

    def save(self, path: str = WEIGHTS) -> None:
        np.savez(path, params=self.params.cpu().numpy(), step=self.step)

    @classmethod
    def load(cls, path: str = WEIGHTS) -> "SpaiNetwork":
        d = np.load(path)
        net = cls()
        net.params = torch.tensor(d["params"], device="cuda")
        net.step = int(d["step"])
        return net


def neighbourhoods(ds: DeviceStack):
    lib = _lib.load()
    n = ds.sys.n_ori + ds.sys.n_rest
    idx = torch.empty((n, K), dtype=torch.int32, device="cuda")
    hk = torch.empty(n, dtype=torch.float32, device="cuda")
    _lib.check("gk_knn", lib.gk_knn(ds.h, C.byref(ds.sys), _ptr(idx), _ptr(hk), _stream()))
    return idx, hk


def features(ds: DeviceStack, idx, hk, pts: Optional[torch.Tensor] = None):
    lib = _lib.load()
    nq = int(pts.numel()) if pts is not None else idx.shape[0]
    f = torch.empty((nq * K, F), dtype=torch.float32, device="cuda")
    _lib.check("gk_spai_features", lib.gk_spai_features(ds.h, C.byref(ds.sys), nq, _ptr(pts), _ptr(idx), _ptr(hk),
                                                        _ptr(f), _stream()))
    return f


def forward(net: SpaiNetwork, feat, keep_acts: bool = False):
    lib = _lib.load()
    n = feat.shape[0]
    y = torch.empty((n, 9), dtype=torch.float32, device="cuda")
    acts = torch.empty((n, 2 * HIDDEN), dtype=torch.float32, device="cuda") if keep_acts else None
    _lib.check("gk_mlp_forward", lib.gk_mlp_forward(_handle(), n, _ptr(net.params), _ptr(feat), _ptr(y),
                                                    _ptr(acts), _stream()))
    return y, acts


def _handle():
    from .engine import handle
    return handle(0)


def spai_blocks(net: SpaiNetwork, ds: DeviceStack):
    """Inference: the 3x3 blocks (n*K, 9) and the neighbour lists of a system."""
    idx, hk = neighbourhoods(ds)
    y, _ = forward(net, features(ds, idx, hk))
    return idx, y


def identity_blocks(ds: DeviceStack):
    """The preconditioner without the network: identity self blocks (the
    diagonal scaling of the equilibrated system)."""
    idx, _ = neighbourhoods(ds)
    n = idx.shape[0]
    y = torch.zeros((n, K, 9), dtype=torch.float32, device="cuda")
    self_mask = idx == torch.arange(n, device="cuda", dtype=torch.int32)[:, None]
    y[..., 0][self_mask] = 1.0
    y[..., 4][self_mask] = 1.0
    y[..., 8][self_mask] = 1.0
    return idx, y.reshape(n * K, 9).contiguous()


def loss_and_grad(net: SpaiNetwork, ds: DeviceStack, idx, hk, dscale, pts):
    """Mean local SPAI residual over points `pts` and its parameter gradient."""
    lib = _lib.load()
    h = _handle()
    nq = int(pts.numel())
    Hloc = torch.empty((nq, 3 * K, 3 * K), dtype=torch.float32, device="cuda")
    _lib.check("gk_spai_local_blocks", lib.gk_spai_local_blocks(h, C.byref(ds.sys), nq, _ptr(pts), _ptr(idx),
                                                                _ptr(dscale), _ptr(Hloc), _stream()))
    feat = features(ds, idx, hk, pts)
    y, acts = forward(net, feat, keep_acts=True)
    dy = torch.empty_like(y)
    loss = torch.empty(nq, dtype=torch.float32, device="cuda")
    _lib.check("gk_spai_loss", lib.gk_spai_loss(h, C.byref(ds.sys), nq, _ptr(pts), _ptr(idx), _ptr(Hloc), _ptr(y),
                                                _ptr(dy), _ptr(loss), _stream()))
    grad = torch.empty_like(net.params)
    _lib.check("gk_mlp_backward", lib.gk_mlp_backward(h, y.shape[0], _ptr(net.params), _ptr(feat), _ptr(acts),
                                                      _ptr(dy), _ptr(grad), _stream()))
    return loss, grad, (Hloc, feat, y, dy)


def diag_scale(ds: DeviceStack):
    lib = _lib.load()
    nc = 3 * ds.sys.n_ori + ds.sys.n_rest
    d = torch.empty(nc, dtype=torch.float64, device="cuda")
    _lib.check("gk_diag_scale", lib.gk_diag_scale(_handle(), C.byref(ds.sys), _ptr(d), _stream()))
    return d


def training_stack(seed: int, n_target: int):
    m = synthetic.training_system(seed, n_target)
    return DeviceStack(m.stacks[0], m.options.kernel_options, [500.0, 500.0, 500.0]), m


def train(net: SpaiNetwork, steps: int, seeds=range(1000, 3000), size: Callable[[int], int] = None,
          batch: int = 1024, lr: float = 2e-3, log: Optional[Callable[[str], None]] = None):
    """Adam on the local SPAI residual; one generated system per step."""
    lib = _lib.load()
    h = _handle()
    seeds = list(seeds)
    size = size or (lambda s: int(np.random.default_rng(s).integers(2000, 20001)))
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        seed = seeds[it % len(seeds)]
        ds, _ = training_stack(seed, size(seed))
        idx, hk = neighbourhoods(ds)
        dscale = diag_scale(ds)
        n = idx.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        pts = torch.randperm(n, generator=g)[:min(batch, n)].to(device="cuda", dtype=torch.int32)
        loss, grad, _ = loss_and_grad(net, ds, idx, hk, dscale, pts)
        net.step += 1
        _lib.check("gk_adam", lib.gk_adam(h, _ptr(net.params), _ptr(grad), _ptr(net.m), _ptr(net.v),
                                          net.params.numel(), float(lr), net.step, 1.0 / pts.numel(), _stream()))
        mean = float(loss.mean())
        hist.append(mean)
        if log and (it % 100 == 0 or it == steps - 1):
            log(f"spai train step {it} seed {seed} N~{n} loss {mean:.4e} ({time.perf_counter() - t0:.1f} s)")
    return hist


def local_inverse_blocks(ds: DeviceStack):
    """The exact local sparse approximate inverse (what the network is trained
    towards): rows of each neighbourhood's inverted equilibrated block."""
    lib = _lib.load()
    idx, _ = neighbourhoods(ds)
    y = torch.empty((idx.shape[0] * K, 9), dtype=torch.float32, device="cuda")
    d = diag_scale(ds)
    _lib.check("gk_spai_local_inverse", lib.gk_spai_local_inverse(_handle(), C.byref(ds.sys), _ptr(idx), _ptr(d),
                                                                  _ptr(y), _stream()))
    return idx, y


def solve_spai(ds: DeviceStack, idx=None, y=None, tol: float = 1e-8, max_iter: int = 500,
               krylov_precision: int = _lib.GK_F32, with_factor: bool = False, no_fallback: bool = False,
               sandwich: bool = True):
    """GMRES preconditioned by the blocks (idx, y) alone (with_factor False),
    or by the shifted FP32 factor combined with them (with_factor True; the
    factor alone when idx/y are None).  sandwich: factor, blocks, factor (the
    form for the network's inexact blocks) instead of blocks then factor.
    no_fallback: return the last iterate instead of falling back to the FP64
    direct solve on stagnation."""
    lib = _lib.load()
    x = torch.empty(ds.N, dtype=torch.float64, device="cuda")
    info = _lib.GkSolveInfo()
    _lib.check("gk_solve_spai", lib.gk_solve_spai(ds.h, C.byref(ds.sys), _ptr(idx), _ptr(y), float(tol),
                                                  int(max_iter), int(krylov_precision),
                                                  (_lib.GK_SPAI_WITH_FACTOR if with_factor else 0)
                                                  | (_lib.GK_SPAI_NO_FALLBACK if no_fallback else 0)
                                                  | (_lib.GK_SPAI_SANDWICH if sandwich else 0),
                                                  _ptr(x), C.byref(info), _stream()))
    DeviceStack.solve_epoch += 1
    return x, {"form": int(info.form), "iterations": int(info.iterations), "fallback": int(info.fallback),
               "relres": float(info.relres), "shift": float(info.shift), "t_setup_ms": float(info.t_factor_ms),
               "t_krylov_ms": float(info.t_refine_ms)}


_NET = {}


def default_network() -> SpaiNetwork:
    """The trained network shipped with the package (spai_weights.npz), loaded once."""
    if "net" not in _NET:
        _NET["net"] = SpaiNetwork.load(WEIGHTS)
    return _NET["net"]


def solve_learned(ds: DeviceStack, tol: float = 1e-10, max_iter: int = 48,
                  fallback: int = _lib.GK_SOLVE_FP64_CHOL):
    """The learned solve form: matrix-free GMRES (FP32 Krylov matvecs on the
    recomputed kernel tiles) preconditioned by the network's blocks; when it
    stagnates above 1e-6, the plain direct solve takes over (``fallback``:
    the FP64 Cholesky + Schur form by default; GK_SOLVE_FP32_REFINE, the
    shifted FP32 factor with FP64 refinement, is the faster choice up to a
    few 10^4 unknowns, but at config-4 size its factor needs a larger shift and
    the refinement stagnates).  ``info['learned']`` records the
    network-preconditioned attempt."""
    if ds.sys.n_ori + ds.sys.n_rest < K:             # too few points for a neighbourhood
        return ds.solve(fallback)
    idx, y = spai_blocks(default_network(), ds)
    x, first = solve_spai(ds, idx, y, tol=tol, max_iter=max_iter, krylov_precision=_lib.GK_F32,
                          with_factor=False, no_fallback=True, sandwich=False)
    if first["relres"] <= 1e-6:
        first["learned"] = dict(first)
        first["fallback"] = 0
        return x, first
    x, info = ds.solve(fallback)
    info["learned"] = first
    info["fallback"] = 1
    return x, info
