"""Input adapter and Transform (CPU).

* Transform.from_input_points reproduces the recorded input_transform of the
  2-layers, anticline and 2.5-D models (tests/golden/transforms.npz).
* model_input_from_geomodel reads a GeoModel through the attribute names the
  reference uses (structural_frame.structural_groups[g].elements[e].
  surface_points / orientations, fault_relations, input_transform, grid,
  interpolation_options, solutions.root_output) -- here a stand-in object --
  and yields the transformed stacks the 2-layers fixture holds.
"""
import os
from types import SimpleNamespace as NS

import numpy as np
import pytest

from gempy_amd import Transform, model_input_from_geomodel
from gempy_amd.data import Solutions, StackRelationType, StackSolution

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["two_layers", "fold", "model1"])
def test_transform_from_input_points_matches_records(name):
    d = np.load(os.path.join(GOLD, "transforms.npz"))
    t = Transform.from_input_points(d[f"{name}_sp"], d[f"{name}_ori"])
    np.testing.assert_allclose(t.position, d[f"{name}_position"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(t.scale, d[f"{name}_scale"], rtol=1e-12)


def test_transform_apply_inverse_and_gradient():
    t = Transform(position=[-6.0, 0.0, -2.51], scale=[1 / 12, 1 / 5.96, 1 / 12])
    x = np.random.default_rng(0).uniform(-5, 5, (10, 3))
    np.testing.assert_allclose(t.apply_inverse(t.apply(x)), x, rtol=1e-14, atol=1e-12)
    # a gradient of f(x) = a.x becomes that of f(x') = (a / s).x', at unchanged length
    a = np.array([[0.3, -0.4, 0.866]])
    g = t.transform_gradient(a)
    np.testing.assert_allclose(np.linalg.norm(g), np.linalg.norm(a), rtol=1e-14)
    u = (a / t.scale)[0]
    np.testing.assert_allclose(g[0], u / np.linalg.norm(u) * np.linalg.norm(a), rtol=1e-14)
    # isotropic: unchanged
    ti = Transform(position=[1, 2, 3], scale=[0.5, 0.5, 0.5])
    np.testing.assert_allclose(ti.transform_gradient(a), a, rtol=1e-15)


def _table(xyz, nugget, grads=None):
    t = NS(xyz=np.asarray(xyz, dtype=float), nugget=np.full(len(xyz), nugget))
    if grads is not None:
        t.grads = np.asarray(grads, dtype=float)
    return t


def _two_layers_geomodel():
    e1 = NS(name="surface1", surface_points=_table([[3, 0, 3.05], [9, 0, 3.05]], 2e-5),
            orientations=_table([[6, 0, 4]], 0.01, [[0, 0, 1]]))
    e2 = NS(name="surface2", surface_points=_table([[3, 0, 1.02], [9, 0, 1.02]], 2e-5),
            orientations=_table(np.zeros((0, 3)), 0.01, np.zeros((0, 3))))
    group = NS(name="default", elements=[e1, e2], structural_relation=NS(name="ERODE"))
    frame = NS(structural_groups=[group], fault_relations=np.zeros((1, 1), dtype=bool))
    sp = np.concatenate([e1.surface_points.xyz, e2.surface_points.xyz])
    tr = Transform.from_input_points(sp, e1.orientations.xyz)
    opts = NS(kernel_options=NS(range=1.7, c_o=10.0, uni_degree=1, i_res=4.0, gi_res=2.0),
              number_octree_levels=1, evaluation_options=NS(compute_scalar_gradient=False), sigmoid_slope=5e6)
    grid = NS(regular_grid=NS(extent=[0, 12, -2, 2, 0, 4], resolution=[50, 5, 50]))
    return NS(structural_frame=frame, input_transform=tr, interpolation_options=opts, grid=grid,
              solutions=None, meta=NS(name="2-layers"))


def test_model_input_from_geomodel_two_layers():
    gm = _two_layers_geomodel()
    m = model_input_from_geomodel(gm)
    d = np.load(os.path.join(GOLD, "two_layers.npz"))
    assert len(m.stacks) == 1
    st = m.stacks[0]
    np.testing.assert_allclose(st.sp_xyz, d["sp_xyz"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(st.ori_xyz, d["ori_xyz"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(st.ori_grad, d["ori_grad"], rtol=0, atol=1e-15)
    assert st.counts == [2, 2] and st.relation == StackRelationType.ERODE and st.faults == []
    np.testing.assert_allclose(st.sp_nugget, d["sp_nugget"])
    assert m.options.kernel_options.range == 1.7 and m.options.kernel_options.c_o == 10.0
    # the dense grid's extent is transformed like the points
    np.testing.assert_allclose(m.grid.extent, [-0.5, 0.5, -2 / 5.96, 2 / 5.96, -2.51 / 12, 1.49 / 12], atol=1e-15)
    assert m.grid.resolution == (50, 5, 50)
    assert m.weights is None


def test_model_input_from_geomodel_faults_and_warm_start():
    """Fault relation matrix -> per-stack fault indices; previous weights
    (solutions.root_output.outputs[s].weights) -> the initial guesses."""
    gm = _two_layers_geomodel()
    fault_el = NS(name="fault", surface_points=_table([[6, 0, 0.5], [6, 0, 3.5]], 2e-5),
                  orientations=_table([[6, 0, 2]], 0.01, [[1, 0, 0]]))
    fault = NS(name="fault", elements=[fault_el], structural_relation=NS(name="FAULT"))
    layers = gm.structural_frame.structural_groups[0]
    gm.structural_frame = NS(structural_groups=[fault, layers],
                             fault_relations=np.array([[False, True], [False, False]]))
    w_prev = [np.arange(4.0), np.arange(9.0)]
    gm.solutions = Solutions(stacks=[StackSolution(weights=w, scalar_field_at_surface_points=np.zeros(1),
                                                   thresholds=np.zeros(1)) for w in w_prev])
    m = model_input_from_geomodel(gm)
    assert [s.relation for s in m.stacks] == [StackRelationType.FAULT, StackRelationType.ERODE]
    assert m.stacks[0].faults == [] and m.stacks[1].faults == [0]
    assert [len(w) for w in m.weights] == [4, 9]


def test_solutions_view_fields():
    """The Solutions record exposes the reference's field names."""
    s1 = StackSolution(weights=np.ones(3), scalar_field_at_surface_points=np.zeros(4),
                       thresholds=np.array([0.5, 0.2]))
    s2 = StackSolution(weights=np.ones(5), scalar_field_at_surface_points=np.zeros(4),
                       thresholds=np.array([0.1, 0.3, 0.9]))
    sol = Solutions(stacks=[s1, s2])
    np.testing.assert_array_equal(sol.scalar_field_at_surface_points, [0.5, 0.2, 0.1, 0.3, 0.9])
    assert [list(o) for o in sol._ordered_elements] == [[0, 1], [0, 1, 2]]
    assert [len(o.weights) for o in sol.root_output.outputs] == [3, 5]
    assert sol.block_solution_type == "dense"
    assert s1.scalar_fields.exported_fields.scalar_field_at_surface_points.tolist() == [0.5, 0.2]
