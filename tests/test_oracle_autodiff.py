"""The torch restatement of the oracle (used to check the hand-written
adjoint) agrees with the NumPy oracle and its autograd gradient with central
finite differences of the NumPy oracle."""
import numpy as np
import torch

from oracle import cokriging as ok
from oracle import torch_cokriging as tk
from test_oracle import A_, C_, _rand_stack


def _t(x):
    return torch.tensor(x, dtype=torch.float64)


def _loss_np(st, sp, G, x, target):
    A, b = ok.build_system(sp, st["counts"], st["sp_nugget"], st["ori_xyz"], G, st["ori_nugget"], A_, C_)
    w = np.linalg.solve(A, b)
    S = ok.evaluate(x, sp, st["counts"], st["ori_xyz"], w, A_, C_)
    return float(((S - target) ** 2).sum())


def test_torch_oracle_gradient_matches_fd():
    st = _rand_stack(11)
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1000, (60, 3))
    target = rng.normal(size=60) * 50 + 800
    sp = _t(st["sp_xyz"]).requires_grad_(True)
    G = _t(st["ori_grad"]).requires_grad_(True)
    w, rest, ref, rs = tk.build_and_solve(sp, st["counts"], _t(st["sp_nugget"]), _t(st["ori_xyz"]), G,
                                          _t(st["ori_nugget"]), A_, C_)
    S = tk.evaluate(_t(x), w, rest, ref, rs, _t(st["ori_xyz"]), A_, C_)
    loss = ((S - _t(target)) ** 2).sum()
    loss.backward()
    assert abs(float(loss) - _loss_np(st, st["sp_xyz"], st["ori_grad"], x, target)) <= 1e-9 * float(loss)
    for (i, u) in [(0, 2), (1, 0), (5, 1), (len(st["sp_xyz"]) - 1, 2)]:
        h = 1e-3
        sp_p, sp_m = st["sp_xyz"].copy(), st["sp_xyz"].copy()
        sp_p[i, u] += h; sp_m[i, u] -= h
        fd = (_loss_np(st, sp_p, st["ori_grad"], x, target) - _loss_np(st, sp_m, st["ori_grad"], x, target)) / (2 * h)
        assert abs(sp.grad[i, u].item() - fd) <= 1e-5 * max(1.0, abs(fd))
    for (g, u) in [(0, 0), (2, 2), (4, 1)]:
        h = 1e-5
        Gp, Gm = st["ori_grad"].copy(), st["ori_grad"].copy()
        Gp[g, u] += h; Gm[g, u] -= h
        fd = (_loss_np(st, st["sp_xyz"], Gp, x, target) - _loss_np(st, st["sp_xyz"], Gm, x, target)) / (2 * h)
        assert abs(G.grad[g, u].item() - fd) <= 1e-5 * max(1.0, abs(fd))


def _loss_np_f(st, sp, q, G, x, target, Fsp, Fx):
    A, b = ok.build_system(sp, st["counts"], st["sp_nugget"], q, G, st["ori_nugget"], A_, C_, fault_sp=Fsp)
    w = np.linalg.solve(A, b)
    S = ok.evaluate(x, sp, st["counts"], q, w, A_, C_, fault_x=Fx)
    return float(((S - target) ** 2).sum())


def test_torch_oracle_fault_columns_and_orientation_positions_match_fd():
    """With a fault-drift column (block values fixed): the torch restatement's
    loss equals the NumPy oracle's, and its autograd gradients w.r.t. the
    surface points, the orientation positions and the orientation gradients
    equal central differences of the NumPy oracle."""
    st = _rand_stack(12)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1000, (60, 3))
    target = rng.normal(size=60) * 50 + 800
    hang = lambda p: ((np.asarray(p) - 500.0) @ np.array([0.8, 0.0, 0.6])) > 0   # noqa: E731
    Fsp = np.where(hang(st["sp_xyz"]), 1.0, 2.0)[:, None]
    Fx = np.where(hang(x), 1.0, 2.0)[:, None]
    sp = _t(st["sp_xyz"]).requires_grad_(True)
    q = _t(st["ori_xyz"]).requires_grad_(True)
    G = _t(st["ori_grad"]).requires_grad_(True)
    w, rest, ref, rs = tk.build_and_solve(sp, st["counts"], _t(st["sp_nugget"]), q, G, _t(st["ori_nugget"]),
                                          A_, C_, fault_sp=_t(Fsp))
    S = tk.evaluate(_t(x), w, rest, ref, rs, q, A_, C_, fault_x=_t(Fx))
    loss = ((S - _t(target)) ** 2).sum()
    loss.backward()
    L0 = _loss_np_f(st, st["sp_xyz"], st["ori_xyz"], st["ori_grad"], x, target, Fsp, Fx)
    assert abs(float(loss) - L0) <= 1e-9 * L0
    assert abs(float(w[-1])) > 1e-6                     # the fault column is used
    def fd(which, i, u, h):
        arrs = {"sp": st["sp_xyz"], "q": st["ori_xyz"], "G": st["ori_grad"]}
        p, m = {k: v.copy() for k, v in arrs.items()}, {k: v.copy() for k, v in arrs.items()}
        p[which][i, u] += h
        m[which][i, u] -= h
        return (_loss_np_f(st, p["sp"], p["q"], p["G"], x, target, Fsp, Fx)
                - _loss_np_f(st, m["sp"], m["q"], m["G"], x, target, Fsp, Fx)) / (2 * h)
    for (i, u) in [(0, 2), (3, 0), (len(st["sp_xyz"]) - 1, 1)]:
        f = fd("sp", i, u, 1e-3)
        assert abs(sp.grad[i, u].item() - f) <= 1e-5 * max(1.0, abs(f))
    for (g, u) in [(0, 0), (2, 2), (4, 1)]:
        f = fd("q", g, u, 1e-3)
        assert abs(q.grad[g, u].item() - f) <= 1e-5 * max(1.0, abs(f)), (q.grad[g, u].item(), f)
        f = fd("G", g, u, 1e-5)
        assert abs(G.grad[g, u].item() - f) <= 1e-5 * max(1.0, abs(f))
