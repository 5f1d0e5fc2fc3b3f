"""Pins of the FP64 oracle (CPU).

Each check is independent of the formula it checks: closed forms of the
cubic covariance, finite differences of C(|x - y|) for the cross-covariance
blocks, the co-kriging interpolation identities, the planar closed form, scale
covariance, and the reference's own recorded outputs (2-layers thresholds,
anticline corner ids).
"""
import os

import numpy as np
import pytest

from oracle import cokriging as ok

GOLD = os.path.join(os.path.dirname(__file__), "golden")
A_, C_ = 1732.0508075688772, 71428.57142857142


def _rand_stack(seed=0, n_surf=3, per=6, n_ori=5, box=1000.0):
    rng = np.random.default_rng(seed)
    pts, counts = [], []
    for k in range(n_surf):
        x = rng.uniform(0, box, (per, 2))
        z = 700 - 200 * k + 30 * np.sin(x[:, 0] / 200.0) + rng.normal(0, 3, per)   # top surface first
        pts.append(np.column_stack([x, z]))
        counts.append(per)
    sp = np.concatenate(pts)
    q = np.column_stack([rng.uniform(0, box, (n_ori, 2)), rng.uniform(300, 700, n_ori)])
    g = rng.normal(size=(n_ori, 3)) * 0.2 + [0, 0, 1]
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return dict(sp_xyz=sp, counts=counts, sp_nugget=np.full(len(sp), 1e-6), ori_xyz=q, ori_grad=g,
                ori_nugget=np.full(n_ori, 0.01))


# ------------------------------------------------------------- A.1 kernels
def test_cov_closed_forms():
    a, c = A_, C_
    assert ok.cov(0.0, a, c) == pytest.approx(c)
    assert ok.dcov_r(0.0, a, c) == pytest.approx(-14 * c / a**2)
    assert ok.d2cov(0.0, a, c) == pytest.approx(-14 * c / a**2)
    for f in (ok.cov, ok.d2cov):
        assert abs(f(a * (1 - 1e-9), a, c)) < 1e-6 * c
    assert abs(ok.dcov_r(a * (1 - 1e-9), a, c) * a) < 1e-6 * c
    assert ok.cov(1.5 * a, a, c) == 0.0


def test_cov_derivatives_fd():
    a, c = A_, C_
    r = np.linspace(0.05, 0.95, 19) * a
    h = 1e-5 * a      # truncation h^2/6 C''' ~ 1e-9 relative
    d1 = (ok.cov(r + h, a, c) - ok.cov(r - h, a, c)) / (2 * h)
    np.testing.assert_allclose(d1, ok.dcov_r(r, a, c) * r, rtol=1e-7, atol=1e-9 * c / a)
    h = 1e-4 * a
    d2 = (ok.cov(r + h, a, c) - 2 * ok.cov(r, a, c) + ok.cov(r - h, a, c)) / h**2
    np.testing.assert_allclose(d2, ok.d2cov(r, a, c), rtol=1e-5, atol=1e-6 * c / a**2)


# -------------------------------------------------------- A.2 blocks by FD
def _C(x, y, a=A_, c=C_):
    return float(ok.cov(np.linalg.norm(np.asarray(x) - np.asarray(y)), a, c))


def test_gradient_blocks_match_finite_differences():
    st = _rand_stack(1)
    A, _ = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_)
    ref, rest, rs, _, _ = ok.split_ref_rest(st["sp_xyz"], st["counts"])
    q = st["ori_xyz"]
    n_o, NG, h = len(q), 3 * len(q), 1e-2
    E = np.eye(3) * h
    # GG: Cov(d_u Z(q), d_v Z(q')) = d^2/dq_u dq'_v C(|q - q'|)
    for g1, g2 in [(0, 1), (2, 4), (3, 0)]:
        for u in range(3):
            for v in range(3):
                fd = (_C(q[g1] + E[u], q[g2] + E[v]) - _C(q[g1] + E[u], q[g2] - E[v])
                      - _C(q[g1] - E[u], q[g2] + E[v]) + _C(q[g1] - E[u], q[g2] - E[v])) / (4 * h * h)
                assert A[u * n_o + g1, v * n_o + g2] == pytest.approx(fd, rel=1e-5, abs=1e-8)
    # GI: gi_res * d/dq_u [C(q, p_i) - C(q, p_ref)]
    for g in range(n_o):
        for i in (0, 3, len(rest) - 1):
            for u in range(3):
                fd = ((_C(q[g] + E[u], rest[i]) - _C(q[g] + E[u], ref[rs[i]]))
                      - (_C(q[g] - E[u], rest[i]) - _C(q[g] - E[u], ref[rs[i]]))) / (2 * h)
                assert A[u * n_o + g, NG + i] == pytest.approx(2.0 * fd, rel=1e-6, abs=1e-9)
                assert A[NG + i, u * n_o + g] == A[u * n_o + g, NG + i]


def test_gg_diagonal_and_symmetry_and_pd():
    st = _rand_stack(2)
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_)
    n_o = len(st["ori_xyz"])
    assert np.allclose(A, A.T, rtol=0, atol=1e-9 * np.abs(A).max())
    np.testing.assert_allclose(np.diag(A)[:3 * n_o], 14 * C_ / A_**2 + 0.01)
    Nc = len(b) - 3
    assert np.linalg.eigvalsh(A[:Nc, :Nc]).min() > 0
    np.testing.assert_array_equal(A[Nc:, Nc:], 0.0)


def test_drift_rows_are_derivatives_and_increments():
    st = _rand_stack(3)
    A, _ = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_, uni_degree=2)
    ref, rest, rs, _, _ = ok.split_ref_rest(st["sp_xyz"], st["counts"])
    q = st["ori_xyz"]
    n_o, NG, R = len(q), 3 * len(q), len(rest)
    h = 1e-3
    for l in range(9):
        for u in range(3):
            e = np.zeros(3); e[u] = h
            fd = (ok.drift_functions((q + e), 2)[0][:, l] - ok.drift_functions((q - e), 2)[0][:, l]) / (2 * h)
            np.testing.assert_allclose(A[u * n_o:(u + 1) * n_o, NG + R + l], fd, rtol=1e-6, atol=1e-6)
        fr = ok.drift_functions(rest, 2)[0][:, l] - ok.drift_functions(ref[rs], 2)[0][:, l]
        np.testing.assert_allclose(A[NG:NG + R, NG + R + l], 2.0 * fr, rtol=1e-12)


# ------------------------------------------------- interpolation identities
@pytest.mark.parametrize("degree", [1, 2])
def test_interpolation_identities(degree):
    st = _rand_stack(4)
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_, uni_degree=degree)
    w = ok.solve_system(A, b)
    ref, rest, rs, ref_idx, rest_idx = ok.split_ref_rest(st["sp_xyz"], st["counts"])
    n_o = len(st["ori_xyz"])
    NG, R = 3 * n_o, len(rest)
    S = ok.evaluate(st["sp_xyz"], st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_, degree)
    nug = st["sp_nugget"][rest_idx] + st["sp_nugget"][ref_idx][rs]
    # S(p_i) - S(p_ref) = -(nu_i + nu_ref) w_i   (the interface rows of A w = b)
    np.testing.assert_allclose(S[rest_idx] - S[ref_idx][rs], -nug * w[NG:NG + R], atol=1e-9)
    # grad S(q_g) = gi_res (G - nu_g w_G)
    _, dS = ok.evaluate(st["ori_xyz"], st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_, degree,
                        gradient=True)
    wG = w[:NG].reshape(3, n_o).T
    np.testing.assert_allclose(dS, 2.0 * (st["ori_grad"] - 0.01 * wG), rtol=1e-8, atol=1e-9)


def test_gradient_matches_fd_of_scalar():
    st = _rand_stack(5)
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_)
    w = ok.solve_system(A, b)
    x = np.random.default_rng(0).uniform(50, 950, (20, 3))
    S, dS = ok.evaluate(x, st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_, gradient=True)
    h = 1e-2
    for u in range(3):
        e = np.zeros(3); e[u] = h
        fd = (ok.evaluate(x + e, st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_)
              - ok.evaluate(x - e, st["sp_xyz"], st["counts"], st["ori_xyz"], w, A_, C_)) / (2 * h)
        np.testing.assert_allclose(dS[:, u], fd, rtol=1e-6, atol=1e-6 * np.abs(dS).max())


def test_solve_direct_matches_cholesky_schur():
    import scipy.linalg as sl
    st = _rand_stack(6)
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], A_, C_)
    x1 = ok.solve_system(A, b)
    Nc = len(b) - 3
    K, U = A[:Nc, :Nc], A[:Nc, Nc:]
    c = sl.cho_factor(K, lower=True)
    Y, y = sl.cho_solve(c, U), sl.cho_solve(c, b[:Nc])
    lam = np.linalg.solve(U.T @ Y, U.T @ y)
    x2 = np.concatenate([y - Y @ lam, lam])
    np.testing.assert_allclose(x1, x2, rtol=1e-7, atol=1e-9 * np.abs(x1).max())
    assert np.linalg.norm(A @ x1 - b) <= 1e-10 * np.linalg.norm(A) * np.linalg.norm(x1)


# ------------------------------------------------------ closed forms / pins
@pytest.mark.parametrize("degree", [1, 2])
def test_planar_closed_form(degree):
    """Planes orthogonal to n with all gradients n: w = 0 and S = 2 n.x (A.7)."""
    rng = np.random.default_rng(7)
    n = np.array([0.3, -0.2, 1.0]); n /= np.linalg.norm(n)
    pts, counts = [], []
    for c0 in (200.0, 400.0, 650.0):
        xy = rng.uniform(0, 1000, (7, 2))
        z = (c0 - n[0] * xy[:, 0] - n[1] * xy[:, 1]) / n[2]
        pts.append(np.column_stack([xy, z])); counts.append(7)
    sp = np.concatenate(pts)
    q = np.column_stack([rng.uniform(0, 1000, (4, 2)), rng.uniform(0, 1000, 4)])
    G = np.tile(n, (4, 1))
    A, b = ok.build_system(sp, counts, np.full(len(sp), 1e-6), q, G, np.full(4, 0.01), A_, C_, degree)
    w = ok.solve_system(A, b)
    assert np.abs(w[:3 * 4 + len(sp) - 3]).max() < 1e-9
    x = rng.uniform(0, 1000, (50, 3))
    S = ok.evaluate(x, sp, counts, q, w, A_, C_, degree)
    np.testing.assert_allclose(S, 2.0 * x @ n, rtol=1e-9, atol=1e-7)


def test_two_layers_recorded_thresholds():
    """2-layers.approved.txt:29,46 -- scalar_field_at_interface 0.09 / -0.248333."""
    d = np.load(os.path.join(GOLD, "two_layers.npz"))
    st = dict(sp_xyz=d["sp_xyz"], counts=list(d["counts"]), sp_nugget=d["sp_nugget"],
              ori_xyz=d["ori_xyz"], ori_grad=d["ori_grad"], ori_nugget=d["ori_nugget"])
    r = ok.compute_stack(st, dict(range=float(d["range"]), c_o=float(d["c_o"])), d["sp_xyz"])
    np.testing.assert_allclose(r["thresholds"], d["expected_thresholds"], rtol=0, atol=1e-12)


def test_anticline_recorded_corner_ids():
    """test_custom_grid.py:50-65 -- ids [3,3,3,3,1,1,1,1] at the box corners."""
    d = np.load(os.path.join(GOLD, "anticline.npz"))
    st = dict(sp_xyz=d["sp_xyz"], counts=list(d["counts"]), sp_nugget=d["sp_nugget"],
              ori_xyz=d["ori_xyz"], ori_grad=d["ori_grad"], ori_nugget=d["ori_nugget"])
    r = ok.compute_stack(st, dict(range=float(d["range"]), c_o=float(d["c_o"])), d["corners"])
    np.testing.assert_array_equal(r["ids"], d["expected_ids"])


def test_scale_covariance():
    """x' = s x, a' = s a, c_o' = s^2 c_o, interface nuggets s^2: S' = s S, same ids."""
    st = _rand_stack(8)
    s = 0.37
    x = np.random.default_rng(1).uniform(0, 1000, (40, 3))
    r1 = ok.compute_stack(st, dict(range=A_, c_o=C_), x)
    st2 = dict(st, sp_xyz=st["sp_xyz"] * s, ori_xyz=st["ori_xyz"] * s, sp_nugget=st["sp_nugget"] * s * s)
    r2 = ok.compute_stack(st2, dict(range=A_ * s, c_o=C_ * s * s), x * s)
    np.testing.assert_allclose(r2["scalar"], s * r1["scalar"], rtol=1e-7, atol=1e-7)
    np.testing.assert_array_equal(r1["ids"], r2["ids"])


def test_segment_rules():
    t = np.array([10.0, 5.0, 0.0])
    np.testing.assert_array_equal(ok.segment(np.array([11.0, 10.0, 7.0, 2.0, -1.0]), t), [1, 1, 2, 3, 4])
    np.testing.assert_array_equal(ok.segment(np.array([11.0, 7.0, -1.0]), t[::-1]), [4, 3, 1])


def test_grid_points_ordering():
    """Cell centres, ij ravel with z fastest (regular_grid.py:58-71,191-201)."""
    g = ok.grid_points([0, 10, 0, 20, 0, 40], [2, 2, 4])
    assert g.shape == (16, 3)
    np.testing.assert_allclose(g[0], [2.5, 5.0, 5.0])
    np.testing.assert_allclose(g[1], [2.5, 5.0, 15.0])
    np.testing.assert_allclose(g[4], [2.5, 15.0, 5.0])
    np.testing.assert_allclose(g[8], [7.5, 5.0, 5.0])


def test_erosion_combination():
    """A younger stack owns every point above its last surface."""
    young = _rand_stack(9, n_surf=1, per=8)
    young["sp_xyz"][:, 2] = 800 + young["sp_xyz"][:, 2] * 0.01
    old = _rand_stack(10, n_surf=2, per=8)
    for s in (young, old):
        s["relation"] = "erode"
    x = np.array([[500, 500, 950.0], [500, 500, 600.0], [500, 500, 50.0]])
    res = ok.compute_model([young, old], dict(range=A_, c_o=C_), x)
    assert res["lith_block"][0] == 1
    assert res["lith_block"][2] == 4
    assert res["lith_block"][1] == 3


def test_config1_fixture_reproduces():
    """The committed config-1 fixture is what the oracle computes now."""
    from gempy_amd import synthetic
    d = np.load(os.path.join(GOLD, "config1_oracle.npz"))
    m = synthetic.config1()
    st = m.stacks[0].as_oracle_dict()
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], m.options.kernel_options.range, m.options.kernel_options.c_o)
    np.testing.assert_allclose(A, d["A"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(ok.solve_system(A, b), d["weights"], rtol=1e-9)


def test_octree_oracle_planar_closed_form():
    """One planar interface z = c: at every level exactly the voxels whose
    z-extent [z0, z1] straddles c (z0 <= c < z1) are refined, and the base
    resolution follows round(length / min length) * 2 (grid.py:128-145)."""
    np.testing.assert_array_equal(ok.octree_base_resolution([0, 2500, 0, 1000, 0, 1000]), [4, 2, 2])
    np.testing.assert_array_equal(ok.octree_base_resolution([0, 10, 0, 30, 0, 10]), [2, 6, 2])
    c = 0.37
    ext = [0, 1, 0, 1, 0, 1]
    lv = ok.octree_levels(lambda p: (p[:, 2] > c).astype(int)[None], ext, (3, 3, 3), 4)
    for l in lv[:-1]:
        res = np.array(l["resolution"])
        iz = l["voxels"] % res[2]
        z0, z1 = iz / res[2], (iz + 1) / res[2]
        np.testing.assert_array_equal(l["refine"], (z0 <= c) & (c < z1))
    assert lv[-1]["resolution"] == (24, 24, 24)
    assert len(lv[-1]["voxels"]) == 24 * 24 * 2          # the two child layers of the one refined layer
