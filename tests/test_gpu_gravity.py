"""Gravity on centred grids (gk_gravity_tz, gk_gravity) against the oracle's
prism restatement and its lith block: the kernel's tz cell by cell, the
reference's 2-layers gravity case (test_gravity.py:67-89; its recorded
-1624.1714 comes from an engine kernel whose source is not in the reference
tree, so the value is reported, not asserted), a layered model whose cells
cross every unit, and the uniform-density identity g = rho * sum(tz)."""
import sys

import numpy as np
import pytest

from oracle import cokriging as ok
from oracle import gravity as og

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _two_layers():
    from gempy_amd import Transform
    from gempy_amd.data import GeoModelInput, InterpolationOptions, SolverForm, StackInput
    sp = np.array([[3, 0, 3.05], [9, 0, 3.05], [3, 0, 1.02], [9, 0, 1.02]], dtype=np.float64)
    q = np.array([[6, 0, 4.0]])
    tr = Transform.from_input_points(sp, q)
    st = StackInput(name="default", sp_xyz=tr.apply(sp), counts=[2, 2], sp_nugget=np.full(4, 2e-5),
                    ori_xyz=tr.apply(q), ori_grad=tr.transform_gradient(np.array([[0, 0, 1.0]])),
                    ori_nugget=np.array([0.01]), surface_names=["surface1", "surface2"])
    return GeoModelInput(stacks=[st], options=InterpolationOptions(solver=SolverForm.FP64_CHOL), name="2-layers"), tr


def _oracle_gravity(model, grid, densities):
    X = grid.model_values()
    res = ok.compute_model([s.as_oracle_dict() for s in model.stacks], model.options.as_dict(), X)
    ids = np.asarray(res["lith_block"]).astype(np.int64).reshape(grid.n_devices, -1)
    tz = og.kernel_tz(*og.centered_kernel(grid.resolution, grid.radius))
    dens = np.asarray(densities, dtype=np.float64)
    return (dens[ids - 1] * tz[None]).sum(1), ids, tz


@pytest.mark.parametrize("res,radius", [((10, 10, 100), 16000.0), ((10, 10, 20), 100.0), ((6, 4, 9), (30.0, 20.0, 50.0))])
def test_kernel_tz_matches_oracle(gpu, res, radius):
    from gempy_amd.gravity import CenteredGrid, calculate_gravity_gradient
    say(f"gravity tz: resolution {res}, radius {radius}")
    g = CenteredGrid([[0, 0, 0]], res, radius)
    tz = calculate_gravity_gradient(g).cpu().numpy()
    ref = og.kernel_tz(*og.centered_kernel(res, radius))
    err = np.abs(tz - ref).max() / np.abs(ref).max()
    say(f"gravity tz: {len(tz)} cells, sum {tz.sum():.10f}, max rel err {err:.2e}")
    # tz is a difference of 8 corner terms of size G |x| ln(2 r) ~ 1e3 times
    # larger than tz itself for the outer cells: libm-level differences of the
    # device's log / atan leave ~1e-12 of max |tz|
    assert len(tz) == g.n_kernel and err <= 1e-10


def test_two_layers_reference_case(gpu):
    from gempy_amd import compute_model
    from gempy_amd.gravity import CenteredGrid, GeophysicsInput
    say("gravity: 2-layers model, device (6, 0, 4), kernel 10x10x100, radius 16000")
    model, tr = _two_layers()
    grid = CenteredGrid([[6, 0, 4.0]], (10, 10, 100), 16000.0, transform=tr)
    dens = [2.6, 2.4, 3.2]
    model.geophysics = GeophysicsInput(grid, dens)
    sol = compute_model(model)
    ref, ids, tz = _oracle_gravity(model, grid, dens)
    say(f"2-layers gravity: gpu {sol.gravity[0]:.6f}, oracle {ref[0]:.6f}, units at the cells "
        f"{np.unique(ids).tolist()}, record of the reference's engine -1624.1714")
    assert abs(sol.gravity[0] - ref[0]) <= 1e-10 * abs(ref[0])


def test_layered_model_cells_cross_units(gpu):
    """Three devices above/inside the layers, a kernel small enough that its
    cells fall in every unit; then the uniform-density identity."""
    from gempy_amd import compute_model
    from gempy_amd.gravity import CenteredGrid, GeophysicsInput, calculate_gravity_gradient
    model, tr = _two_layers()
    grid = CenteredGrid([[6, 0, 4.0], [4, 0.5, 3.5], [8, -0.5, 2.0]], (8, 8, 24), (3.0, 3.0, 4.0), transform=tr)
    dens = [1.9, 2.45, 3.1]
    say(f"gravity: layered model, {grid.n_devices} devices x {grid.n_kernel} cells")
    model.geophysics = GeophysicsInput(grid, dens)
    sol = compute_model(model)
    ref, ids, tz = _oracle_gravity(model, grid, dens)
    say(f"layered gravity: gpu {np.array2string(sol.gravity, precision=6)}, oracle {np.array2string(ref, precision=6)}, "
        f"cells per unit {[int((ids == u).sum()) for u in (1, 2, 3)]}")
    assert all(int((ids == u).sum()) > 0 for u in (1, 2, 3))
    np.testing.assert_allclose(sol.gravity, ref, rtol=1e-10, atol=0)
    rho = 2.67
    tzg = calculate_gravity_gradient(grid)
    model.geophysics = GeophysicsInput(grid, [rho] * 3, tzg)
    sol = compute_model(model)
    np.testing.assert_allclose(sol.gravity, rho * float(tzg.sum()), rtol=1e-12, atol=0)
