"""Finite faults on the CPU: the ellipsoid implicit function, the transform's
rotation about a pivot, the adapter, and the oracle's tapered fault drift.

* Ellipsoid (ellipsoid_implicit_function.py:21-35): f = 1/2 on the ellipsoid
  scaled by k, 1 - sigmoid(-prod k) at the centre, monotone in the distance.
* Rotation (R = Rz Ry Rx, degrees): orthogonal, det 1, a quarter turn about z
  takes x to y, the pivot is fixed, apply / apply_inverse round-trip.
* The ONE_FAULT example (examples_generator.py:198-240, its model5 tables):
  the input transform equals the reference's serialized one (position -500,
  scale 0.0005, test_finite_faults verify file), the fault is tapered by the
  reference test's ellipsoid (test_finite_faults.py:11-52).
* Oracle invariants: f == 1 everywhere (an ellipsoid much larger than the
  model with steep k) gives the infinite fault's ids exactly (a constant shift
  of a drift column changes no id); a finite fault changes the ids of the
  offset stack near the fault's tip and leaves the fault stack's own ids.
"""
import functools
import os

import numpy as np
import pytest

from gempy_amd.adapters import Transform
from gempy_amd.data import StackRelationType
from gempy_amd.faults import EllipsoidFunction, FiniteFaultData, ellipsoid_3d_factory
from oracle import cokriging as ok

ONE = os.path.join(os.path.dirname(__file__), "golden", "one_fault")

# the reference test's ellipsoid, in input coordinates (test_finite_faults.py:11-13)
CENTER = np.array([500.0, 500.0, 500.0])
RADIUS = np.array([500.0, 100.0, 200.0]) / 2
K = np.ones(3)


def one_fault_model(resolution=(20, 20, 20), finite=None):
    """ONE_FAULT: a fault offsetting one two-surface series, on a dense grid of
    the 1000^3 box; finite = (rotation degrees, k) adds the reference test's
    finite fault (centre and radius scaled into the model frame, pivot at the
    centre)."""
    from gempy_amd.io import read_csv_model
    lm = read_csv_model(os.path.join(ONE, "model5_surface_points.csv"), os.path.join(ONE, "model5_orientations.csv"),
                        {"Fault_Series": ("fault",), "Strat_Series": ("rock2", "rock1")},
                        [0, 1000, 0, 1000, 0, 1000], resolution,
                        relations={"Fault_Series": StackRelationType.FAULT},
                        fault_relations=[[0, 1], [0, 0]], name="fault")
    if finite is not None:
        rot, k = finite
        c = lm.transform.apply(CENTER[None])[0]
        fn = ellipsoid_3d_factory(c, RADIUS * lm.transform.scale, k)
        lm.model.stacks[0].finite_fault = FiniteFaultData(fn, Transform(rotation=np.asarray(rot, dtype=float)), c)
    return lm


def test_ellipsoid_closed_forms():
    f = ellipsoid_3d_factory(CENTER, RADIUS, K)
    d = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=float)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    on = CENTER + d * RADIUS / np.sqrt(((d * RADIUS / RADIUS) ** 2).sum(1, keepdims=True))
    np.testing.assert_allclose(f(on), 0.5, rtol=0, atol=1e-15)
    np.testing.assert_allclose(f(CENTER[None]), 1 / (1 + np.exp(-1.0)), rtol=1e-15)
    k = np.array([2.0, 0.5, 3.0])
    g = ellipsoid_3d_factory(CENTER, RADIUS, k)
    # on the ellipsoid with semi-axes r k: 1/2; at the centre 1 - sigmoid(-prod k)
    np.testing.assert_allclose(g(CENTER + np.diag(RADIUS * k)), 0.5, atol=1e-15)
    np.testing.assert_allclose(g(CENTER[None]), 1 / (1 + np.exp(-3.0)), rtol=1e-15)
    t = np.linspace(0, 3, 50)[:, None] * np.array([[1.0, 0.3, -0.2]]) * RADIUS + CENTER
    assert np.all(np.diff(f(t)) < 0)


def test_rotation_about_pivot():
    R = Transform(rotation=np.array([10.0, -35.0, 70.0])).rotation_matrix()
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-15)
    assert abs(np.linalg.det(R) - 1) < 1e-14
    np.testing.assert_allclose(Transform(rotation=[0, 0, 90]).rotation_matrix() @ [1, 0, 0], [0, 1, 0], atol=1e-15)
    np.testing.assert_allclose(Transform(rotation=[90, 0, 0]).rotation_matrix() @ [0, 1, 0], [0, 0, 1], atol=1e-15)
    np.testing.assert_allclose(Transform(rotation=[0, 90, 0]).rotation_matrix() @ [0, 0, 1], [1, 0, 0], atol=1e-15)
    # the composition order: z after y after x
    a = Transform(rotation=[90, 0, 90]).rotation_matrix()
    np.testing.assert_allclose(a @ [0, 1, 0], Transform(rotation=[0, 0, 90]).rotation_matrix() @ [0, 0, 1],
                               atol=1e-15)
    tr = Transform(rotation=np.array([0.0, 60.0, 0.0]))
    p = np.array([0.3, -0.2, 0.1])
    x = np.random.default_rng(0).uniform(-1, 1, (40, 3))
    np.testing.assert_allclose(tr.apply_with_pivot(tr.apply_inverse_with_pivot(x, p), p), x, atol=1e-15)
    np.testing.assert_allclose(tr.apply_inverse_with_pivot(p[None], p), p[None], atol=1e-16)
    # distances to the pivot are kept
    np.testing.assert_allclose(np.linalg.norm(tr.apply_with_pivot(x, p) - p, axis=1), np.linalg.norm(x - p, axis=1))
    with pytest.raises(NotImplementedError):
        Transform(position=[1, 0, 0], rotation=[0, 0, 10]).apply_with_pivot(x, p)
    with pytest.raises(NotImplementedError):
        Transform(rotation=[0, 0, 10]).apply(x)


def test_finite_fault_data_and_adapter():
    lm = one_fault_model()
    # Transform.from_input_points of the model5 tables = the reference's serialized transform
    np.testing.assert_allclose(lm.transform.position, [-500, -500, -500])
    np.testing.assert_allclose(lm.transform.scale, 0.0005)
    c = lm.transform.apply(CENTER[None])[0]
    np.testing.assert_allclose(c, 0, atol=1e-15)
    ff = FiniteFaultData(ellipsoid_3d_factory(c, RADIUS * 5e-4, K), Transform(rotation=[0, 60, 0]), c)
    d = ff.as_oracle_dict()
    np.testing.assert_allclose(d["radii"], [0.125, 0.025, 0.05])
    # f of the rotated points = the ellipsoid at the inverse-rotated points
    x = np.random.default_rng(1).uniform(-0.5, 0.5, (30, 3))
    R = Transform(rotation=[0, 60, 0]).rotation_matrix()
    np.testing.assert_allclose(ff.apply(x @ R.T), ff.implicit_function(x), rtol=1e-13)

    # the adapter takes the reference's faults_input_data.finite_fault_data
    class Obj:
        def __init__(self, **kw):
            self.__dict__.update(kw)
    part = functools.partial(lambda xyz, center, radii, k_factors: None, center=c, radii=RADIUS * 5e-4, k_factors=K)
    from gempy_amd.adapters import _finite_fault
    g = Obj(faults_input_data=Obj(finite_fault_data=Obj(
        implicit_function=part, implicit_function_transform=Obj(position=np.zeros(3), rotation=np.array([0, 60, 0]),
                                                                scale=np.ones(3)), pivot=c)))
    got = _finite_fault(g)
    assert isinstance(got.implicit_function, EllipsoidFunction)
    np.testing.assert_allclose(got.apply(x), ff.apply(x), rtol=0, atol=0)
    assert _finite_fault(Obj(faults_input_data=None)) is None
    with pytest.raises(NotImplementedError):
        _finite_fault(Obj(faults_input_data=Obj(finite_fault_data=Obj(
            implicit_function=np.sin, implicit_function_transform=None, pivot=c))))


def _oracle(lm, X):
    m = lm.model
    return ok.compute_model([s.as_oracle_dict() for s in m.stacks], m.options.as_dict(), X)


def test_oracle_unit_taper_is_the_infinite_fault():
    """An ellipsoid 100x the model with k = 10: f = 1 at every point, the drift
    column is ids - 1, and every id equals the infinite fault's."""
    inf = one_fault_model()
    fin = one_fault_model()
    fin.model.stacks[0].finite_fault = FiniteFaultData(ellipsoid_3d_factory(np.zeros(3), np.full(3, 100.0),
                                                                            np.full(3, 10.0)), Transform(), np.zeros(3))
    X = inf.model.grid.values()
    a, b = _oracle(inf, X), _oracle(fin, X)
    np.testing.assert_array_equal(a["lith_block"], b["lith_block"])
    for ra, rb in zip(a["stacks"], b["stacks"]):
        np.testing.assert_array_equal(ra["ids"], rb["ids"])


def test_oracle_finite_fault_changes_the_offset_stack_only():
    inf = one_fault_model()
    fin = one_fault_model(finite=([0, 60, 0], K))
    X = inf.model.grid.values()
    a, b = _oracle(inf, X), _oracle(fin, X)
    np.testing.assert_array_equal(a["stacks"][0]["ids"], b["stacks"][0]["ids"])
    changed = a["stacks"][1]["ids"] != b["stacks"][1]["ids"]
    assert 0 < changed.sum() < 0.5 * len(X)
    # the values a finite fault feeds the offset stack: (id - 1) f at the points
    ids = a["stacks"][0]["ids"].astype(np.float64)
    ff = fin.model.stacks[0].finite_fault
    np.testing.assert_allclose(ok.finite_fault_values(ids, X, ff.as_oracle_dict()), (ids - 1) * ff.apply(X),
                               rtol=1e-14, atol=1e-15)
    for r in b["stacks"]:
        assert np.linalg.norm(r["A"] @ r["weights"] - r["b"]) <= 1e-9 * np.linalg.norm(r["b"])


def test_finite_fault_only_on_fault_stacks():
    from gempy_amd.compute import _validate
    lm = one_fault_model()
    lm.model.stacks[1].finite_fault = FiniteFaultData(ellipsoid_3d_factory(np.zeros(3), np.ones(3), K))
    with pytest.raises(ValueError):
        _validate(lm.model)
