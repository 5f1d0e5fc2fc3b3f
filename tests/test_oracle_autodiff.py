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
