"""Differentiable FP64 restatement of the oracle in torch -- TEST INFRASTRUCTURE ONLY.

It states the same system and evaluation as oracle/cokriging.py (whose
conventions are pinned there) with torch operations, so that torch.autograd
gives the gradient of a misfit w.r.t. the surface-point coordinates, the
orientation positions and the orientation gradients (fault-drift columns
included, their block values held fixed): the reference obtains such gradients by taping its
PyTorch backend (gempy/API/compute_API.py:49,66; test_ad_I.py:41-51), which is
not available here.  Used only to check the hand-written adjoint.
"""
from __future__ import annotations

import torch


def _rho(h, a):
    r2 = (h * h).sum(-1)
    r = torch.sqrt(torch.clamp(r2, min=1e-300))
    return r / a


def _cov(rho, c_o):
    v = c_o * (1 - 7 * rho**2 + 35 / 4 * rho**3 - 3.5 * rho**5 + 0.75 * rho**7)
    return torch.where(rho < 1, v, torch.zeros_like(v))


def _dcov_r(rho, a, c_o):
    v = c_o / a**2 * (-14 + 105 / 4 * rho - 17.5 * rho**3 + 5.25 * rho**5)
    return torch.where(rho < 1, v, torch.zeros_like(v))


def _dd(rho, a, c_o):   # (C'' - C'/r) / r^2 * r^2 handled by the caller through unit vectors
    v = 105 / 4 * c_o / a**2 * rho * (1 - rho**2) ** 2
    return torch.where(rho < 1, v, torch.zeros_like(v))


def build_and_solve(sp, counts, sp_nug, q, G, q_nug, a, c_o, i_res=4.0, gi_res=2.0, fault_sp=None):
    """fault_sp (n_sp, F), optional: fault-block values at the surface points
    (piecewise constant: constants of the differentiation)."""
    counts = [int(c) for c in counts]
    starts = [0]
    for c in counts[:-1]:
        starts.append(starts[-1] + c)
    ref_idx = torch.tensor(starts)
    rest_idx = torch.tensor([i for i in range(sp.shape[0]) if i not in set(starts)])
    surf = torch.repeat_interleave(torch.arange(len(counts)), torch.tensor(counts))
    rs = surf[rest_idx]
    rest, ref = sp[rest_idx], sp[ref_idx]
    refp = ref[rs]
    n_o, R = q.shape[0], rest.shape[0]
    NG = 3 * n_o
    F = 0 if fault_sp is None else fault_sp.shape[1]
    # GG
    h = q[:, None, :] - q[None, :, :]
    rho = _rho(h, a)
    d1 = _dcov_r(rho, a, c_o)
    dd = _dd(rho, a, c_o)
    r = rho * a
    hh = torch.where(r[..., None] > 0, h / torch.clamp(r, min=1e-300)[..., None], torch.zeros_like(h))
    rows = []
    for u in range(3):
        cols = []
        for v in range(3):
            blk = -(dd * hh[..., u] * hh[..., v] + (d1 if u == v else 0.0))
            cols.append(blk)
        rows.append(torch.cat(cols, 1))
    GG = torch.cat(rows, 0) + torch.diag(q_nug.repeat(3))
    # GI
    h1 = q[:, None, :] - rest[None, :, :]
    h2 = q[:, None, :] - refp[None, :, :]
    k1 = _dcov_r(_rho(h1, a), a, c_o)
    k2 = _dcov_r(_rho(h2, a), a, c_o)
    GI = torch.cat([gi_res * (h1[..., u] * k1 - h2[..., u] * k2) for u in range(3)], 0)
    # II
    def C(x, y):
        return _cov(_rho(x[:, None, :] - y[None, :, :], a), c_o)
    II = i_res * (C(rest, rest) - C(rest, refp) - C(refp, rest) + C(refp, refp))
    II = II + torch.diag(sp_nug[rest_idx] + sp_nug[ref_idx][rs])
    UG = torch.cat([torch.eye(3).repeat(n_o, 1).reshape(n_o, 3, 3)[:, :, u] for u in range(3)], 0)
    UI = gi_res * (rest - refp)
    if F:
        fs = torch.as_tensor(fault_sp, dtype=GG.dtype)
        UG = torch.cat([UG.to(GG), torch.zeros(NG, F, dtype=GG.dtype)], 1)
        UI = torch.cat([UI, gi_res * (fs[rest_idx] - fs[ref_idx][rs])], 1)
    Lt = 3 + F
    top = torch.cat([GG, GI, UG.to(GG)], 1)
    mid = torch.cat([GI.T, II, UI], 1)
    bot = torch.cat([UG.T.to(GG), UI.T, torch.zeros(Lt, Lt, dtype=GG.dtype)], 1)
    A = torch.cat([top, mid, bot], 0)
    b = torch.cat([G.T.reshape(-1), torch.zeros(R + Lt, dtype=G.dtype)])
    w = torch.linalg.solve(A, b)
    return w, rest, ref, rs


def evaluate(x, w, rest, ref, rs, q, a, c_o, i_res=4.0, gi_res=2.0, fault_x=None):
    """fault_x (P, F), optional: fault-block values at x."""
    n_o, R = q.shape[0], rest.shape[0]
    NG = 3 * n_o
    wG = w[:NG].reshape(3, n_o).T
    wI = w[NG:NG + R]
    lam = w[NG + R:NG + R + 3]
    lamF = w[NG + R + 3:]
    W = torch.zeros(ref.shape[0], dtype=w.dtype).index_add(0, rs, wI)
    h = x[:, None, :] - q[None, :, :]
    d1 = _dcov_r(_rho(h, a), a, c_o)
    s = gi_res * ((-h * d1[..., None]) * wG[None]).sum((1, 2))
    s = s + i_res * (_cov(_rho(x[:, None, :] - rest[None], a), c_o) @ wI)
    s = s - i_res * (_cov(_rho(x[:, None, :] - ref[None], a), c_o) @ W)
    s = s + gi_res * (x @ lam)
    if fault_x is not None and lamF.numel():
        s = s + gi_res * (torch.as_tensor(fault_x, dtype=w.dtype) @ lamF)
    return s
