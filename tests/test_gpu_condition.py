"""Condition number of the co-kriging matrix (gk_condition_number: Lanczos on
A and on A^-1 through the kept factor) against numpy's SVD-based
np.linalg.cond of the oracle's matrix, for both kept-factor forms, with a
fault column, at drift degrees 0-2, and through compute_model's
kernel_options.compute_condition_number (the reference's option, read at
gempy/modules/optimize_nuggets/_ops.py:40).

Tolerance: sigma_min of a backward-stable computation (the SVD, or a solve
with A) moves by up to eps * sigma_max, i.e. eps * cond relative; both sides
carry that, so the bound is max(1e-6, 100 eps cond)."""
import sys

import numpy as np
import pytest

from oracle import cokriging as ok

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _tol(cond):
    return max(1e-6, 100 * np.finfo(float).eps * cond)


@pytest.mark.parametrize("form,deg,fault", [(2, 1, False), (1, 1, False), (2, 2, False), (2, 0, False),
                                            (2, 1, True), (1, 1, True)])
def test_condition_number_matches_svd(gpu, form, deg, fault):
    from gempy_amd.engine import DeviceStack
    from test_gpu_parity import _kopts, _oracle_A, _small_stack
    say(f"condition number: form {form} degree {deg} fault {fault}")
    st, fsp = _small_stack(3, fault=fault)
    k = _kopts(deg)
    ds = DeviceStack(st, k, [500, 500, 500], fsp)
    ds.solve(form)
    cond, info = ds.condition_number()
    Ao, _ = _oracle_A(st, k, fsp)
    ref = np.linalg.cond(Ao)
    s = np.linalg.svd(Ao, compute_uv=False)
    say(f"cond: gpu {cond:.10e} svd {ref:.10e} (N {len(Ao)}), sigma max {info['sigma_max']:.6e} ({s[0]:.6e}), "
        f"min {info['sigma_min']:.6e} ({s[-1]:.6e}), steps {info['lanczos_steps']}")
    assert abs(cond / ref - 1) <= _tol(ref)


def test_compute_model_condition_number_option(gpu):
    from gempy_amd import GemPyEngineConfig, compute_model, synthetic
    say("condition number through compute_model: config2 scale 0.1, every stack")
    m = synthetic.config2(scale=0.1, resolution=(8, 8, 8), octree_levels=1)
    m.options.kernel_options.compute_condition_number = True
    sol = compute_model(m, GemPyEngineConfig())
    ko = m.options.kernel_options
    res = ok.compute_model([x.as_oracle_dict() for x in m.stacks], m.options.as_dict(), m.stacks[0].sp_xyz[:1])
    for st, s, r in zip(m.stacks, sol.stacks, res["stacks"]):
        ref = np.linalg.cond(r["A"])
        say(f"cond {st.name}: gpu {s.solver_info['condition_number']:.10e} svd {ref:.10e} (N {len(r['A'])}, "
            f"form {s.solver_info['form']}, steps {s.solver_info['condition']['lanczos_steps']})")
        assert abs(s.solver_info["condition_number"] / ref - 1) <= _tol(ref)
    assert ko.condition_number == sol.stacks[-1].solver_info["condition_number"]
