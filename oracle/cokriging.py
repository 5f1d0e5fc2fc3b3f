"""FP64 NumPy restatement of GemPy's implicit co-kriging interpolator.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and there only as the checker of the HIP path.

What it restates
----------------
The reference hands every numeric step of ``gp.compute_model`` to the external
package ``gempy_engine`` (``gempy/API/compute_API.py:68-73``), which is not part
of the reference tree.  This module therefore restates the published method
(Lajaunie et al. 1997; Calcagno et al. 2008; de la Varga et al. 2019, cited at
``README.md:218-221``) with the engine's option names and defaults as recorded
in ``test/test_modules/test_geophysics/test_gravity.test_gravity.verify/
2-layers.approved.txt:150-181`` (range, c_o, uni_degree, i_res=4, gi_res=2,
kernel_function="cubic", sigmoid_slope).

Conventions (each is pinned by a test in ``tests/test_oracle_*.py``):

* Cubic covariance C(r) and its derivatives C'(r)/r, C''(r)  -- closed forms,
  pinned by finite differences and C(0)=c_o, -C''(0)=14 c_o/a^2.
* Unknown order per stack: [w_G (x comps of all orientations, then y, then z),
  w_I (one per *rest* point, surfaces contiguous, first point of every surface
  is its reference point), lambda (uni drift), lambda_F (fault drift)].
  Component-major gradient ordering follows the engine's universal-drift
  layout (U_G[:n,0], U_G[n:2n,1], ...).
* The system is the co-kriging system of the data vector (grad Z, gi_res *
  (Z(p_i) - Z(p_ref))).  With i_res = gi_res**2 (the recorded defaults 4 and 2)
  it is the self-consistent co-kriging of Z, and the evaluated field is
  S = gi_res * Z.  For planar data S is therefore exactly 2 * (n . x), which is
  what the recorded ``scalar_field_at_interface`` values 0.09 / -0.248333 of the
  2-layers model show (2-layers.approved.txt:29,46).
* Nuggets: orientation nugget on the G diagonal; interface nugget
  nu_i + nu_ref(i) on the I diagonal (variance of the increment noise).
* Surface thresholds: S at each surface's reference point.
* Ids: 1 + #{k : S < t_k} for a normal stack (thresholds decreasing with the
  stratigraphic order); basement = n_surfaces + 1; reversed stacks count
  S > t_k.  Above the top surface -> id 1 (test_custom_grid.py:62-65).
* Grid: cell centres ``extent_min + d/2 ... extent_max - d/2`` with ``ij``
  meshgrid ravel order, z fastest (regular_grid.py:58-71,191-201).
* Multi-stack: stacks ordered youngest first; an ERODE stack owns every point
  above its last surface; fault stacks produce a fault block (1 above / 2
  below the fault surface) and one drift column in every stack they offset.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "cov", "dcov_r", "d2cov", "split_ref_rest", "drift_functions",
    "build_system", "solve_system", "evaluate", "surface_thresholds",
    "segment", "grid_points", "compute_stack", "compute_model", "fault_columns_used",
    "octree_base_resolution", "octree_levels",
]


# --------------------------------------------------------------------------
# A.1  cubic covariance (kernel_function "cubic", 2-layers.approved.txt:157)
# --------------------------------------------------------------------------
def cov(r, a, c_o):
    """C(r) = c_o (1 - 7 rho^2 + 35/4 rho^3 - 7/2 rho^5 + 3/4 rho^7), rho=r/a<1."""
    rho = np.asarray(r, dtype=np.float64) / a
    val = c_o * (1.0 - 7.0 * rho**2 + 35.0 / 4.0 * rho**3 - 3.5 * rho**5 + 0.75 * rho**7)
    return np.where(rho < 1.0, val, 0.0)


def dcov_r(r, a, c_o):
    """C'(r)/r = c_o/a^2 (-14 + 105/4 rho - 35/2 rho^3 + 21/4 rho^5)."""
    rho = np.asarray(r, dtype=np.float64) / a
    val = c_o / a**2 * (-14.0 + 105.0 / 4.0 * rho - 17.5 * rho**3 + 5.25 * rho**5)
    return np.where(rho < 1.0, val, 0.0)


def d2cov(r, a, c_o):
    """C''(r) = c_o/a^2 (-14 + 105/2 rho - 70 rho^3 + 63/2 rho^5)."""
    rho = np.asarray(r, dtype=np.float64) / a
    val = c_o / a**2 * (-14.0 + 52.5 * rho - 70.0 * rho**3 + 31.5 * rho**5)
    return np.where(rho < 1.0, val, 0.0)


def _d2_minus_d1(r, a, c_o):
    """(C'' - C'/r) = 105/4 c_o/a^2 rho (1 - rho^2)^2 (exact simplification)."""
    rho = np.asarray(r, dtype=np.float64) / a
    val = 105.0 / 4.0 * c_o / a**2 * rho * (1.0 - rho**2) ** 2
    return np.where(rho < 1.0, val, 0.0)


# --------------------------------------------------------------------------
# data split: reference point = first point of each surface
# --------------------------------------------------------------------------
def split_ref_rest(sp_xyz, counts, nugget=None):
    """Return (ref_xyz (S,3), rest_xyz (R,3), rest_surf (R,), ref_idx (S,), rest_idx (R,))."""
    sp_xyz = np.asarray(sp_xyz, dtype=np.float64)
    counts = np.asarray(counts, dtype=np.int64)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    ref_idx = starts
    mask = np.ones(len(sp_xyz), dtype=bool)
    mask[ref_idx] = False
    rest_idx = np.nonzero(mask)[0]
    surf_of_point = np.repeat(np.arange(len(counts)), counts)
    return sp_xyz[ref_idx], sp_xyz[rest_idx], surf_of_point[rest_idx], ref_idx, rest_idx


# --------------------------------------------------------------------------
# A.3  universal drift functions
# --------------------------------------------------------------------------
def drift_functions(x, degree):
    """f_l(x) (P,L) and grad f_l (P,L,3) for uni_degree 0, 1, 2.

    Order for degree 2: x, y, z, x^2, y^2, z^2, xy, xz, yz.
    """
    x = np.asarray(x, dtype=np.float64)
    P = x.shape[0]
    if degree == 0:
        return np.zeros((P, 0)), np.zeros((P, 0, 3))
    f = [x[:, 0], x[:, 1], x[:, 2]]
    one, zero = np.ones(P), np.zeros(P)
    g = [np.stack([one, zero, zero], 1), np.stack([zero, one, zero], 1),
         np.stack([zero, zero, one], 1)]
    if degree >= 2:
        X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
        f += [X * X, Y * Y, Z * Z, X * Y, X * Z, Y * Z]
        g += [np.stack([2 * X, zero, zero], 1), np.stack([zero, 2 * Y, zero], 1),
              np.stack([zero, zero, 2 * Z], 1), np.stack([Y, X, zero], 1),
              np.stack([Z, zero, X], 1), np.stack([zero, Z, Y], 1)]
    if degree > 2:
        raise ValueError("uni_degree must be 0, 1 or 2")
    return np.stack(f, 1), np.stack(g, 1)


def n_drift(degree):
    return {0: 0, 1: 3, 2: 9}[degree]


# --------------------------------------------------------------------------
# A.2 - A.4  system assembly
# --------------------------------------------------------------------------
def _pair(x, y):
    h = x[:, None, :] - y[None, :, :]
    r = np.sqrt(np.sum(h * h, axis=-1))
    return h, r


def _chunk_rows(n_cols, budget=2_000_000):
    """Rows per chunk so that a (rows, n_cols, 3) temporary stays near `budget` entries."""
    return max(1, budget // max(1, n_cols))


def build_system(sp_xyz, counts, sp_nugget, ori_xyz, ori_grad, ori_nugget,
                 a, c_o, uni_degree=1, i_res=4.0, gi_res=2.0,
                 fault_sp=None, out=None):
    """Co-kriging matrix A (N,N) and right-hand side b (N,) of one stack.

    fault_sp: optional (n_sp, F) fault-block values at the surface points.
    The covariance blocks are filled in row chunks (temporaries of a few
    million entries), so a config-3 system (N = 34,993) needs only the
    matrix itself.  out: optional preallocated (N, N) float64 array.
    """
    ref, rest, rsurf, ref_idx, rest_idx = split_ref_rest(sp_xyz, counts)
    sp_nugget = np.asarray(sp_nugget, dtype=np.float64)
    q = np.asarray(ori_xyz, dtype=np.float64).reshape(-1, 3)
    G = np.asarray(ori_grad, dtype=np.float64).reshape(-1, 3)
    nug_o = np.asarray(ori_nugget, dtype=np.float64).reshape(-1)
    n_o, R = q.shape[0], rest.shape[0]
    L = n_drift(uni_degree)
    F = 0 if fault_sp is None else np.asarray(fault_sp).shape[1]
    NG = 3 * n_o
    N = NG + R + L + F
    if out is None:
        A = np.zeros((N, N))
    else:
        A = out
        A[...] = 0.0
    refp = ref[rsurf]                     # reference point of every rest point

    # --- GG block: -[d_uv C'/r + hh_uv (C'' - C'/r)], h = q_g - q_g'
    if n_o:
        cs = _chunk_rows(n_o)
        for g0 in range(0, n_o, cs):
            g1 = min(n_o, g0 + cs)
            h, r = _pair(q[g0:g1], q)
            d1 = dcov_r(r, a, c_o)
            dd = _d2_minus_d1(r, a, c_o)
            with np.errstate(invalid="ignore", divide="ignore"):
                hh = np.where(r[..., None] > 0, h / np.where(r > 0, r, 1.0)[..., None], 0.0)
            for u in range(3):
                for v in range(3):
                    blk = -(dd * hh[..., u] * hh[..., v] + (d1 if u == v else 0.0))
                    A[u * n_o + g0:u * n_o + g1, v * n_o:(v + 1) * n_o] = blk
        A[np.arange(NG), np.arange(NG)] += np.tile(nug_o, 3)

    # --- GI block: gi_res [ (q-p_i)_u C'/r(|q-p_i|) - (q-p_ref)_u C'/r(|q-p_ref|) ]
    if n_o and R:
        cs = _chunk_rows(R)
        for g0 in range(0, n_o, cs):
            g1 = min(n_o, g0 + cs)
            h1, r1 = _pair(q[g0:g1], rest)
            h2, r2 = _pair(q[g0:g1], refp)
            k1 = dcov_r(r1, a, c_o)
            k2 = dcov_r(r2, a, c_o)
            for u in range(3):
                blk = gi_res * (h1[..., u] * k1 - h2[..., u] * k2)
                A[u * n_o + g0:u * n_o + g1, NG:NG + R] = blk
                A[NG:NG + R, u * n_o + g0:u * n_o + g1] = blk.T

    # --- II block: i_res [C(pi,pj) - C(pi,p0j) - C(p0i,pj) + C(p0i,p0j)]
    if R:
        cs = _chunk_rows(R)
        for i0 in range(0, R, cs):
            i1 = min(R, i0 + cs)
            _, r_rr = _pair(rest[i0:i1], rest)
            blk = cov(r_rr, a, c_o)
            _, r_rf = _pair(rest[i0:i1], refp)
            blk -= cov(r_rf, a, c_o)
            _, r_fr = _pair(refp[i0:i1], rest)
            blk -= cov(r_fr, a, c_o)
            _, r_ff = _pair(refp[i0:i1], refp)
            blk += cov(r_ff, a, c_o)
            A[NG + i0:NG + i1, NG:NG + R] = i_res * blk
        nug = sp_nugget[rest_idx] + sp_nugget[ref_idx][rsurf]
        A[NG + np.arange(R), NG + np.arange(R)] += nug

    # --- drift blocks
    if L:
        _, gq = drift_functions(q, uni_degree)          # (n_o, L, 3)
        for u in range(3):
            A[u * n_o:(u + 1) * n_o, NG + R:NG + R + L] = gq[:, :, u]
        f_rest, _ = drift_functions(rest, uni_degree)
        f_ref, _ = drift_functions(refp, uni_degree)
        UI = gi_res * (f_rest - f_ref)
        A[NG:NG + R, NG + R:NG + R + L] = UI
        A[NG + R:NG + R + L, :NG + R] = A[:NG + R, NG + R:NG + R + L].T
    if F:
        fs = np.asarray(fault_sp, dtype=np.float64)
        UF = gi_res * (fs[rest_idx] - fs[ref_idx][rsurf])
        A[NG:NG + R, NG + R + L:] = UF
        A[NG + R + L:, NG:NG + R] = UF.T

    b = np.zeros(N)
    b[:NG] = G.T.reshape(-1)
    return A, b


def solve_system(A, b, overwrite=False):
    """Dense LU solve (numpy.linalg.solve): the reference's default direct solver
    (kernel_solver 1, 2-layers.approved.txt:158).  overwrite: factor A in place
    (LAPACK gesv through scipy), for systems too large for a second copy."""
    if overwrite:
        import scipy.linalg
        return scipy.linalg.solve(A, b, overwrite_a=True, check_finite=False)
    return np.linalg.solve(A, b)


# --------------------------------------------------------------------------
# A.5  evaluation of S and grad S
# --------------------------------------------------------------------------
def evaluate(x, sp_xyz, counts, ori_xyz, weights, a, c_o, uni_degree=1,
             i_res=4.0, gi_res=2.0, fault_x=None, gradient=False,
             chunk=None):
    """Scalar field S(x) (P,) and optionally grad S (P,3).

    fault_x: optional (P, F) fault-block values at x.
    """
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    ref, rest, rsurf, _, _ = split_ref_rest(sp_xyz, counts)
    q = np.asarray(ori_xyz, dtype=np.float64).reshape(-1, 3)
    n_o, R, S = q.shape[0], rest.shape[0], ref.shape[0]
    L = n_drift(uni_degree)
    NG = 3 * n_o
    w = np.asarray(weights, dtype=np.float64)
    wG = w[:NG].reshape(3, n_o).T                       # (n_o, 3)
    wI = w[NG:NG + R]
    lam = w[NG + R:NG + R + L]
    lamF = w[NG + R + L:]
    # ref terms collapse per surface: W_s = sum_{i in s} w_i
    Wref = np.zeros(S)
    np.add.at(Wref, rsurf, wI)
    P = x.shape[0]
    out = np.zeros(P)
    dout = np.zeros((P, 3)) if gradient else None
    nsrc = max(1, n_o + R + S)
    if chunk is None:
        chunk = max(1, min(500_000, 4_000_000 // nsrc))
    for s0 in range(0, P, chunk):
        xs = x[s0:s0 + chunk]
        acc = np.zeros(xs.shape[0])
        dacc = np.zeros((xs.shape[0], 3)) if gradient else None
        if n_o:
            h, r = _pair(xs, q)                          # h = x - q
            d1 = dcov_r(r, a, c_o)
            # (q - x)_u C'/r = -h_u C'/r
            acc += gi_res * np.einsum("pgu,gu->p", -h * d1[..., None], wG)
            if gradient:
                dd = _d2_minus_d1(r, a, c_o)
                with np.errstate(invalid="ignore", divide="ignore"):
                    hh = np.where(r[..., None] > 0,
                                  h / np.where(r > 0, r, 1.0)[..., None], 0.0)
                # d/dx_v [(q-x)_u C'/r] = -[d_uv C'/r + hh_u hh_v (C''-C'/r)]
                t = np.einsum("pgu,gu->pg", hh, wG)          # (P, n_o)
                dacc += gi_res * (-(d1[..., None] * wG[None]).sum(1)
                                  - np.einsum("pg,pgv->pv", t * dd, hh))
        if R:
            h, r = _pair(xs, rest)
            acc += i_res * cov(r, a, c_o) @ wI
            if gradient:
                dacc += i_res * np.einsum("pi,piv->pv", dcov_r(r, a, c_o) * wI, h)
        if S:
            h, r = _pair(xs, ref)
            acc -= i_res * cov(r, a, c_o) @ Wref
            if gradient:
                dacc -= i_res * np.einsum("ps,psv->pv", dcov_r(r, a, c_o) * Wref, h)
        if L:
            f, gf = drift_functions(xs, uni_degree)
            acc += gi_res * f @ lam
            if gradient:
                dacc += gi_res * np.einsum("plv,l->pv", gf, lam)
        if len(lamF):
            fx = np.asarray(fault_x, dtype=np.float64).reshape(P, -1)[s0:s0 + chunk]
            acc += gi_res * fx @ lamF
        out[s0:s0 + chunk] = acc
        if gradient:
            dout[s0:s0 + chunk] = dacc
    return (out, dout) if gradient else out


def surface_thresholds(sp_xyz, counts, s_at_sp):
    """Threshold of every surface = S at its reference (first) point."""
    _, _, _, ref_idx, _ = split_ref_rest(sp_xyz, counts)
    return np.asarray(s_at_sp)[ref_idx]


def segment(s, thresholds):
    """Integer unit id per point: 1 above the first surface ... S+1 (basement)."""
    s = np.asarray(s)
    t = np.asarray(thresholds)
    if len(t) >= 2 and t[0] < t[-1]:
        below = (s[:, None] > t[None, :])
    else:
        below = (s[:, None] < t[None, :])
    return 1 + below.sum(axis=1).astype(np.int32)


def grid_points(extent, resolution):
    """Regular-grid cell centres in ``ij`` ravel order (regular_grid.py:58-71,191-201)."""
    ext = np.asarray(extent, dtype=np.float64)
    res = np.asarray(resolution, dtype=np.int64)
    axes = []
    for d in range(3):
        lo, hi = ext[2 * d], ext[2 * d + 1]
        step = (hi - lo) / res[d]
        axes.append(lo + step / 2 + step * np.arange(res[d]))
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([c.ravel() for c in g], axis=1)


# --------------------------------------------------------------------------
# stack + model drivers
# --------------------------------------------------------------------------
def fault_columns_used(counts, fault_sp):
    """Indices of the fault columns whose block differs between some rest
    point and its reference point.  A fault that does not separate any two
    points of one surface leaves an all-zero drift column (a singular
    system); it carries no information about the stack and is dropped
    (its weight is 0)."""
    fs = np.asarray(fault_sp, dtype=np.float64)
    _, _, rsurf, ref_idx, rest_idx = split_ref_rest(np.zeros((fs.shape[0], 3)), counts)
    inc = fs[rest_idx] - fs[ref_idx][rsurf]
    return np.nonzero(np.any(inc != 0.0, axis=0))[0]


def compute_stack(stack, opts, x, fault_sp=None, fault_x=None, gradient=False):
    """Assemble, solve and evaluate one stack.  Returns a dict."""
    if fault_sp is not None:
        keep = fault_columns_used(stack["counts"], fault_sp)
        fault_sp = np.asarray(fault_sp)[:, keep] if len(keep) else None
        if fault_x is not None:
            fault_x = np.asarray(fault_x).reshape(np.asarray(x).reshape(-1, 3).shape[0], -1)[:, keep] \
                if len(keep) else None
    A, b = build_system(stack["sp_xyz"], stack["counts"], stack["sp_nugget"],
                        stack["ori_xyz"], stack["ori_grad"], stack["ori_nugget"],
                        opts["range"], opts["c_o"], opts.get("uni_degree", 1),
                        opts.get("i_res", 4.0), opts.get("gi_res", 2.0), fault_sp)
    w = solve_system(A, b)
    kw = dict(a=opts["range"], c_o=opts["c_o"], uni_degree=opts.get("uni_degree", 1),
              i_res=opts.get("i_res", 4.0), gi_res=opts.get("gi_res", 2.0))
    s_sp = evaluate(stack["sp_xyz"], stack["sp_xyz"], stack["counts"], stack["ori_xyz"],
                    w, fault_x=fault_sp, **kw)
    thr = surface_thresholds(stack["sp_xyz"], stack["counts"], s_sp)
    res = evaluate(x, stack["sp_xyz"], stack["counts"], stack["ori_xyz"], w,
                   fault_x=fault_x, gradient=gradient, **kw)
    s, ds = (res if gradient else (res, None))
    return dict(A=A, b=b, weights=w, s_at_sp=s_sp, thresholds=thr, scalar=s,
                gradient=ds, ids=segment(s, thr))


def finite_fault_values(ids, pts, ff):
    """Drift values of a fault block at pts: the ids, or for a finite fault
    (ff: rotation R, pivot, center, radii, k) (id - 1) f(R^T (x - pivot) + pivot)
    with the ellipsoid function of
    gempy/modules/custom_implicit_functions/ellipsoid_implicit_function.py:21-35."""
    if ff is None:
        return ids
    p = np.asarray(ff["pivot"], dtype=np.float64)
    xr = (np.asarray(pts, dtype=np.float64).reshape(-1, 3) - p) @ np.asarray(ff["rotation"]) + p
    k = np.asarray(ff["k"], dtype=np.float64)
    v = (((xr - np.asarray(ff["center"])) / (np.asarray(ff["radii"]) * k)) ** 2).sum(1) - 1.0
    with np.errstate(over="ignore"):            # exp -> inf gives f = 1
        f = 1.0 - 1.0 / (1.0 + np.exp(-v * np.prod(k)))
    return (ids - 1.0) * f


def compute_model(stacks, opts, x, gradient=False):
    """Multi-stack model.

    stacks: list (youngest first) of dicts with keys sp_xyz, counts, sp_nugget,
    ori_xyz, ori_grad, ori_nugget, relation ('erode' | 'onlap' | 'fault'), faults (list
    of indices of fault stacks that offset it; they must come earlier).
    Returns dict with per-stack results, lith_block (global ids over the
    non-fault stacks, basement last) and fault_blocks.
    """
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    out = []
    fault_vals_x, fault_vals_sp_fn = {}, {}
    for k, st in enumerate(stacks):
        fl = st.get("faults", [])
        fsp = fx = None
        if fl:
            fsp = np.stack([fault_vals_sp_fn[j](st["sp_xyz"]) for j in fl], 1)
            fx = np.stack([fault_vals_x[j] for j in fl], 1)
        r = compute_stack(st, opts, x, fsp, fx, gradient)
        out.append(r)
        if st.get("relation") == "fault":
            kw = dict(a=opts["range"], c_o=opts["c_o"], uni_degree=opts.get("uni_degree", 1),
                      i_res=opts.get("i_res", 4.0), gi_res=opts.get("gi_res", 2.0))

            ff = st.get("finite_fault")

            def block_fn(pts, st=st, r=r, kw=kw, fl=fl, ff=ff):
                # a fault offset by older faults needs their blocks at pts too
                fpts = (np.stack([fault_vals_sp_fn[j](pts) for j in fl], 1) if fl else None)
                s = evaluate(pts, st["sp_xyz"], st["counts"], st["ori_xyz"], r["weights"],
                             fault_x=fpts, **kw)
                return finite_fault_values(segment(s, r["thresholds"]).astype(np.float64), pts, ff)
            fault_vals_sp_fn[k] = block_fn
            fault_vals_x[k] = finite_fault_values(r["ids"].astype(np.float64), x, ff)
    # lithology: youngest non-fault stack whose domain (above its last surface)
    # contains the point; an onlapping stack ('onlap') only above the first
    # surface of the next older non-fault stack, whose top it does not erode
    lith = np.zeros(x.shape[0], dtype=np.int32)
    assigned = np.zeros(x.shape[0], dtype=bool)
    layer = [k for k, st in enumerate(stacks) if st.get("relation") != "fault"]
    masks = {}
    next_id = 1
    for pos, k in enumerate(layer):
        st, r = stacks[k], out[k]
        ns = len(st["counts"])
        inside = r["ids"] <= ns
        if st.get("relation") == "onlap" and pos + 1 < len(layer):
            inside &= out[layer[pos + 1]]["ids"] == 1
        take = inside & ~assigned
        lith[take] = r["ids"][take] + next_id - 1
        masks[k] = take
        assigned |= take
        next_id += ns
    lith[~assigned] = next_id                   # basement
    if layer:                                   # held by the oldest stack's mask
        masks[layer[-1]] = masks[layer[-1]] | ~assigned
    faults = [r["ids"] for st, r in zip(stacks, out) if st.get("relation") == "fault"]
    mask_matrix = [masks.get(k, np.ones(x.shape[0], dtype=bool)) for k in range(len(stacks))]
    return dict(stacks=out, lith_block=lith, fault_blocks=faults, basement_id=next_id,
                mask_matrix=mask_matrix)


# --------------------------------------------------------------------------
# octree levels (the refinement the GPU path implements; the engine's own
# criterion -- octree_error_threshold, fault.approved.txt:210-214 -- is not in
# the reference tree: this restates the library's rule, refine a voxel when any
# id field differs among its centre and 8 corners)
# --------------------------------------------------------------------------
def octree_base_resolution(extent):
    """round(length / min length) * 2 per axis (gempy/core/data/grid.py:128-145)."""
    e = np.asarray(extent, dtype=np.float64)
    lengths = np.array([e[1] - e[0], e[3] - e[2], e[5] - e[4]])
    return (np.round(lengths / lengths.min()).astype(int) * 2)


def octree_levels(fields_at, extent, base_resolution, levels):
    """Level by level from the base lattice: probe every active voxel at its
    centre and 8 corners (corner c at offsets ((c>>2)&1, (c>>1)&1, c&1)),
    refine when any field differs among the 9 probes, the 8 children of each
    refined voxel (same corner order, resolution doubled) in voxel order form
    the next level; the last level is probed at the centres only.
    fields_at(points (n, 3)) -> int array (n_fields, n).  Returns per level a
    dict(resolution, voxels, ids (fields at the centres), refine)."""
    e = np.asarray(extent, dtype=np.float64)
    res = np.asarray(base_resolution, dtype=np.int64)
    vox = np.arange(int(np.prod(res)), dtype=np.int64)
    out = []
    for lvl in range(levels):
        last = lvl == levels - 1
        h = np.array([(e[1] - e[0]) / res[0], (e[3] - e[2]) / res[1], (e[5] - e[4]) / res[2]])
        ix, iy, iz = vox // (res[1] * res[2]), (vox // res[2]) % res[1], vox % res[2]
        base = np.stack([ix, iy, iz], 1).astype(np.float64)
        offs = [np.array([0.5, 0.5, 0.5])] if last else \
            [np.array([0.5, 0.5, 0.5])] + [np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1], float) for c in range(8)]
        pts = np.stack([e[[0, 2, 4]] + (base + o) * h for o in offs], 1).reshape(-1, 3)
        f = np.asarray(fields_at(pts)).reshape(-1, len(vox), len(offs))
        if last:
            out.append(dict(resolution=tuple(int(r) for r in res), voxels=vox, ids=f[:, :, 0], refine=None))
            break
        refine = np.any(f[:, :, 1:] != f[:, :, :1], axis=(0, 2))
        out.append(dict(resolution=tuple(int(r) for r in res), voxels=vox, ids=f[:, :, 0], refine=refine))
        r = vox[refine]
        rx, ry, rz = r // (res[1] * res[2]), (r // res[2]) % res[1], r % res[2]
        nres = 2 * res
        kids = [((2 * rx + ((c >> 2) & 1)) * nres[1] + (2 * ry + ((c >> 1) & 1))) * nres[2] + (2 * rz + (c & 1))
                for c in range(8)]
        vox = np.stack(kids, 1).reshape(-1)
        res = nres
    return out
