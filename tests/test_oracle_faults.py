"""Pins of the oracle's fault-drift path (CPU).

* The fault columns of A are the increments of the fault block between each
  rest point and its surface's reference point (the drift-column convention).
* The co-kriging interpolation identities hold with fault columns present:
  they tie the F columns of build_system to the fault term of evaluate (a
  transposed block, a missing gi_res or a sign error on either side breaks them).
* Closed form: planar layers offset by a throw t along their normal across a
  planar fault, every gradient equal to the normal n.  The data then lie in the
  span of the drift {x, y, z, F}: the covariance weights vanish, the drift
  weights are n and the fault weight is t / (F_hanging - F_foot), so that
  S(x) = gi_res (n.x + t F(x) / (F_hanging - F_foot)) is constant on every
  offset surface.
* The same closed form through the multi-stack driver: a planar fault stack
  (its block is then the exact half-space indicator) offsetting a planar
  layer stack; ids match the analytic geometry.
* Two faults, the second offset by the first (fault on fault).
"""
import numpy as np
import pytest

from oracle import cokriging as ok

A_, C_ = 1732.0508075688772, 71428.57142857142
N_LAYER = np.array([0.25, -0.15, 1.0]) / np.linalg.norm([0.25, -0.15, 1.0])
N_FAULT = np.array([np.sin(np.deg2rad(60.0)), 0.0, np.cos(np.deg2rad(60.0))])
P_FAULT = np.array([500.0, 500.0, 500.0])
THROW = 70.0


# dip-slip direction: in the fault plane, so a displaced point stays on its wall
SLIP = np.array([-np.cos(np.deg2rad(60.0)), 0.0, np.sin(np.deg2rad(60.0))])


def _hanging(x):
    return (np.asarray(x) - P_FAULT) @ N_FAULT > 0


def _fault_block(x):
    """Block of the planar fault as the oracle's segmentation gives it: 1 on
    the side the fault's gradient points to (larger S), 2 on the other."""
    return np.where(_hanging(x), 1.0, 2.0)


def _planar_offset_stack(rng, n_surf=3, per=12, n_ori=6):
    """Layers n.x = c_k on the foot wall, n.x = c_k - THROW on the hanging wall."""
    pts, counts = [], []
    for k, c0 in enumerate((700.0, 520.0, 330.0)[:n_surf]):
        xy = rng.uniform(0, 1000, (per, 2))
        z = (c0 - N_LAYER[0] * xy[:, 0] - N_LAYER[1] * xy[:, 1]) / N_LAYER[2]
        p = np.column_stack([xy, z])
        hang = _hanging(p)
        p[hang] -= (THROW / (SLIP @ N_LAYER)) * SLIP   # down-dip slip: n.p drops by THROW
        # keep the reference (first) point on the foot wall
        order = np.argsort(hang, kind="stable")
        pts.append(p[order])
        counts.append(per)
    sp = np.concatenate(pts)
    q = rng.uniform(100, 900, (n_ori, 3))
    return dict(sp_xyz=sp, counts=counts, sp_nugget=np.full(len(sp), 1e-6), ori_xyz=q,
                ori_grad=np.tile(N_LAYER, (n_ori, 1)), ori_nugget=np.full(n_ori, 0.01))


def _rand_faulted_stack(seed):
    rng = np.random.default_rng(seed)
    st = _planar_offset_stack(rng)
    st["sp_xyz"] = st["sp_xyz"] + rng.normal(0, 8.0, st["sp_xyz"].shape)     # not planar any more
    g = st["ori_grad"] + rng.normal(0, 0.15, st["ori_grad"].shape)
    st["ori_grad"] = g / np.linalg.norm(g, axis=1, keepdims=True)
    return st


def test_fault_columns_are_block_increments():
    st = _rand_faulted_stack(1)
    F = _fault_block(st["sp_xyz"])[:, None]
    A, _ = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_, fault_sp=F)
    _, rest, rs, ref_idx, rest_idx = ok.split_ref_rest(st["sp_xyz"], st["counts"])
    n_o, R = len(st["ori_xyz"]), len(rest)
    NG, col = 3 * n_o, 3 * n_o + R + 3
    assert A.shape == (col + 1, col + 1)
    inc = F[rest_idx, 0] - F[ref_idx][rs, 0]
    assert np.count_nonzero(inc) > 0
    np.testing.assert_array_equal(A[NG:NG + R, col], 2.0 * inc)
    np.testing.assert_array_equal(A[col, NG:NG + R], 2.0 * inc)
    np.testing.assert_array_equal(A[:NG, col], 0.0)          # the block has no gradient
    np.testing.assert_array_equal(A[col, NG + R:], 0.0)


def test_interpolation_identities_with_fault_column():
    st = _rand_faulted_stack(2)
    F = _fault_block(st["sp_xyz"])[:, None]
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_, fault_sp=F)
    w = ok.solve_system(A, b)
    _, rest, rs, ref_idx, rest_idx = ok.split_ref_rest(st["sp_xyz"], st["counts"])
    n_o, R = len(st["ori_xyz"]), len(rest)
    NG = 3 * n_o
    assert abs(w[-1]) > 1e-3                                   # the fault column is used
    S = ok.evaluate(st["sp_xyz"], st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_, fault_x=F)
    nug = st["sp_nugget"][rest_idx] + st["sp_nugget"][ref_idx][rs]
    np.testing.assert_allclose(S[rest_idx] - S[ref_idx][rs], -nug * w[NG:NG + R], atol=1e-9)
    Fq = _fault_block(st["ori_xyz"])[:, None]
    _, dS = ok.evaluate(st["ori_xyz"], st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_,
                        fault_x=Fq, gradient=True)
    wG = w[:NG].reshape(3, n_o).T
    np.testing.assert_allclose(dS, 2.0 * (st["ori_grad"] - 0.01 * wG), rtol=1e-8, atol=1e-9)
    # the same evaluation without the fault term misses exactly gi_res lam_F F(x)
    S0 = ok.evaluate(st["sp_xyz"], st["sp_xyz"], st["counts"], st["ori_xyz"], w[:-1], A_, C_)
    np.testing.assert_allclose(S - S0, 2.0 * w[-1] * F[:, 0], rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("fault_values", [(1.0, 2.0), (0.0, 1.0), (2.0, 1.0)])
def test_planar_offset_closed_form(fault_values):
    """Covariance weights vanish, drift weights = n, fault weight = t/dF."""
    rng = np.random.default_rng(3)
    st = _planar_offset_stack(rng)
    f_hang, f_foot = fault_values
    F = np.where(_hanging(st["sp_xyz"]), f_hang, f_foot)[:, None]
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_, fault_sp=F)
    w = ok.solve_system(A, b)
    n_o, R = len(st["ori_xyz"]), len(st["sp_xyz"]) - len(st["counts"])
    Nc = 3 * n_o + R
    assert np.abs(w[:Nc]).max() < 1e-9
    np.testing.assert_allclose(w[Nc:Nc + 3], N_LAYER, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(w[-1], THROW / (f_hang - f_foot), rtol=1e-9)
    x = rng.uniform(0, 1000, (200, 3))
    Fx = np.where(_hanging(x), f_hang, f_foot)
    S = ok.evaluate(x, st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_, fault_x=Fx[:, None])
    expect = 2.0 * (x @ N_LAYER + THROW * Fx / (f_hang - f_foot))
    np.testing.assert_allclose(S, expect, rtol=1e-9, atol=1e-7)
    # every point of a surface, on either wall, has the surface's threshold
    Ssp = ok.evaluate(st["sp_xyz"], st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_, fault_x=F)
    for k, seg in enumerate(np.split(Ssp, np.cumsum(st["counts"])[:-1])):
        np.testing.assert_allclose(seg, seg[0], rtol=0, atol=1e-7)


def _planar_fault_stack(rng, n=10, n_ori=4, normal=N_FAULT, point=P_FAULT):
    yz = rng.uniform(50, 950, (n, 2))
    x = point[0] - (normal[1] * (yz[:, 0] - point[1]) + normal[2] * (yz[:, 1] - point[2])) / normal[0]
    yzo = rng.uniform(100, 900, (n_ori, 2))
    xo = point[0] - (normal[1] * (yzo[:, 0] - point[1]) + normal[2] * (yzo[:, 1] - point[2])) / normal[0]
    return dict(sp_xyz=np.column_stack([x, yz]), counts=[n], sp_nugget=np.full(n, 1e-6),
                ori_xyz=np.column_stack([xo, yzo]), ori_grad=np.tile(normal, (n_ori, 1)),
                ori_nugget=np.full(n_ori, 0.01), relation="fault", faults=[])


def test_planar_fault_model_through_driver():
    """compute_model with a planar fault stack offsetting a planar layer stack:
    the fault block is the exact half-space indicator (1 on the side the
    fault's gradient points to), the layer stack's covariance weights vanish
    and its ids are those of the analytic offset geometry."""
    rng = np.random.default_rng(4)
    fault = _planar_fault_stack(rng)
    layers = _planar_offset_stack(rng)
    layers.update(relation="erode", faults=[0])
    x = ok.grid_points([0, 1000, 0, 1000, 0, 1000], [12, 12, 12])
    res = ok.compute_model([fault, layers], dict(range=A_, c_o=C_), x)
    np.testing.assert_array_equal(res["fault_blocks"][0], _fault_block(x).astype(np.int32))
    r = res["stacks"][1]
    n_o, R = len(layers["ori_xyz"]), len(layers["sp_xyz"]) - 3
    assert np.abs(r["weights"][:3 * n_o + R]).max() < 1e-9
    np.testing.assert_allclose(r["weights"][-1], THROW / (1.0 - 2.0), rtol=1e-9)
    # analytic ids: unfold the hanging wall, then count the planes below
    z_eff = x @ N_LAYER + THROW * _hanging(x)
    planes = np.array([700.0, 520.0, 330.0])
    expect = 1 + (z_eff[:, None] < planes[None, :]).sum(1)
    margin = np.min(np.abs(z_eff[:, None] - planes[None, :]), 1) > 1e-6
    np.testing.assert_array_equal(res["lith_block"][margin], expect[margin])


def test_fault_on_fault_block_uses_older_fault():
    """A second planar fault whose points are offset by the first fault: its
    block at a point depends on the first fault's block there (the driver
    evaluates it with the first fault's values)."""
    rng = np.random.default_rng(5)
    f1 = _planar_fault_stack(rng)
    n2 = np.array([0.0, np.sin(np.deg2rad(70.0)), np.cos(np.deg2rad(70.0))])
    f2 = _planar_fault_stack(rng, n=14, normal=np.array([0.3, n2[1], n2[2]]) / np.linalg.norm([0.3, n2[1], n2[2]]),
                             point=np.array([500.0, 420.0, 500.0]))
    hang = _hanging(f2["sp_xyz"])
    f2["sp_xyz"][hang] -= 80.0 * SLIP              # the first fault drops its hanging wall
    f2["faults"] = [0]
    x = ok.grid_points([0, 1000, 0, 1000, 0, 1000], [10, 10, 10])
    res = ok.compute_model([f1, f2], dict(range=A_, c_o=C_), x)
    r2 = res["stacks"][1]
    assert abs(r2["weights"][-1]) > 1.0             # the second fault's own fault column is used
    # the block returned for the grid equals the one the driver's block function
    # gives (it evaluates the second fault with the first fault's block)
    S = ok.evaluate(x, f2["sp_xyz"], f2["counts"], f2["ori_xyz"], r2["weights"], A_, C_,
                    fault_x=res["fault_blocks"][0][:, None].astype(np.float64))
    np.testing.assert_array_equal(ok.segment(S, r2["thresholds"]), res["fault_blocks"][1])
