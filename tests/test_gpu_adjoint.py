"""The hand-written adjoint on the differentiation workload (config 5: the
config-2 geometry, misfit on the faulted 'lower' stack at observation points
against a perturbed model), fault-drift column included: gradients w.r.t. the
interface coordinates, the orientation gradients and the orientation positions
(the reference's dip_positions, test_ad_I.py:48-50) against torch autograd of
the oracle restatement (<= 1e-6) and Richardson-extrapolated central
differences of the GPU forward (h = 1e-3 extent)."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_adjoint_config5_faulted_stack(gpu):
    say("adjoint: config5 scale 0.1, faulted 'lower' stack, 2000 observations, with orientation positions")
    from gempy_amd import _lib, synthetic
    from gempy_amd.autodiff import misfit_gradient, scalar_field_at
    from gempy_amd.compute import _Model
    from gempy_amd.data import GeoModelInput
    from oracle import torch_cokriging as tk
    model, k, pert, obs = synthetic.config5(scale=0.1, resolution=(16, 16, 16), n_obs=2000)
    ko = model.options.kernel_options
    center = [500.0, 500.0, 500.0]
    base = _Model(GeoModelInput(stacks=model.stacks[:1], options=model.options, grid=model.grid),
                  _lib.GK_SOLVE_FP64_CHOL)
    fault = base.stacks[0]

    def fault_at(xyz):
        _, _, ids = fault.evaluate_points(torch.as_tensor(np.asarray(xyz), dtype=torch.float64, device="cuda"),
                                          thresholds=fault.thresholds)
        return ids.double().cpu().numpy()[:, None]
    st = model.stacks[k]
    f_sp, f_obs = fault_at(st.sp_xyz), fault_at(obs)
    assert len(np.unique(f_sp)) == 2 and len(np.unique(f_obs)) == 2       # the fault cuts the stack
    o = torch.tensor(obs, dtype=torch.float64, device="cuda")
    tgt = scalar_field_at(torch.tensor(pert.sp_xyz), torch.tensor(pert.ori_grad), pert, ko, o, center,
                          fault_at(pert.sp_xyz), f_obs).detach().cpu().numpy()
    loss, d_sp, d_g, d_q = misfit_gradient(st, ko, obs, tgt, center, f_sp, f_obs, positions=True)
    # torch autograd through the oracle restatement (float64, CPU), fault column included
    t = lambda x: torch.tensor(np.asarray(x), dtype=torch.float64)  # noqa: E731
    sp = t(st.sp_xyz).requires_grad_(True)
    G = t(st.ori_grad).requires_grad_(True)
    q = t(st.ori_xyz).requires_grad_(True)
    w, rest, ref, rs = tk.build_and_solve(sp, st.counts, t(st.sp_nugget), q, G, t(st.ori_nugget), ko.range, ko.c_o,
                                          fault_sp=t(f_sp))
    S = tk.evaluate(t(obs), w, rest, ref, rs, q, ko.range, ko.c_o, fault_x=t(f_obs))
    lt = ((S - t(tgt)) ** 2).sum()
    lt.backward()
    say(f"config5 adjoint: loss gpu {loss:.10e} torch {float(lt):.10e}; fault weight {float(w[-1]):.4e}")
    assert abs(loss - float(lt)) <= 1e-8 * abs(float(lt))
    errs = {}
    for name, mine, ref_g in (("sp", d_sp, sp.grad), ("G", d_g, G.grad), ("q", d_q, q.grad)):
        r = ref_g.numpy()
        errs[name] = np.abs(mine - r).max() / np.abs(r).max()
    say(f"config5 adjoint vs autograd: max rel err d/dp {errs['sp']:.3e}, d/dG {errs['G']:.3e}, "
        f"d/dq {errs['q']:.3e}")
    assert max(errs.values()) <= 1e-6

    # Richardson-extrapolated central differences of the GPU forward (fault
    # block values held fixed, as in the gradient)
    def L(spv, gv, qv):
        tmpl = st.copy()
        tmpl.ori_xyz = qv
        with torch.no_grad():
            Sv = scalar_field_at(torch.tensor(spv), torch.tensor(gv), tmpl, ko, o, center, f_sp, f_obs,
                                 ori_xyz=torch.tensor(qv))
        return float(((Sv.cpu().numpy() - tgt) ** 2).sum())

    def fd(which, i, u, h):
        base_v = {"sp": st.sp_xyz, "q": st.ori_xyz}
        p, m = {a: b.copy() for a, b in base_v.items()}, {a: b.copy() for a, b in base_v.items()}
        p[which][i, u] += h
        m[which][i, u] -= h
        return (L(p["sp"], st.ori_grad, p["q"]) - L(m["sp"], st.ori_grad, m["q"])) / (2 * h)
    h = 1.0                                   # 1e-3 of the 1000 m extent
    for which, grad, cases in (("q", d_q, [(0, 0), (5, 2), (17, 1)]), ("sp", d_sp, [(3, 2), (40, 0)])):
        scale = np.abs(grad).max()
        for (i, u) in cases:
            f1, f2 = fd(which, i, u, h), fd(which, i, u, h / 2)
            rich = (4 * f2 - f1) / 3
            say(f"fd d/d{which}[{i},{u}]: h {f1:.6e}, extrapolated {rich:.6e}, adjoint {grad[i, u]:.6e}")
            assert abs(rich - grad[i, u]) <= 1e-4 * max(abs(rich), 1e-3 * scale)
