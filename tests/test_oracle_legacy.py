"""The oracle against the reference's legacy GemPy-v2 FP64 fault models
(examples/data/gempy_models/Tutorial_ch1-9a_Fault_relations, -9b; fixture
written by tests/golden/make_legacy_fixtures.py).

Tutorial 9a: a planar fault (9 points, 1 orientation) offsetting the older
of two layer series (3 + 2 surfaces), the younger series eroding.  Pinned:
* the fault's scalar field equals v2's up to v2's additive constant (slope 1),
  to the fixture's float32 storage precision;
* the fault block equals v2's rounded block everywhere;
* the lith block equals v2's (unit ids shifted by one, basement 0 in v2)
  outside the +-1e-6 band around the thresholds (inside it: cells lying
  exactly on a planar surface, whose tie v2 resolves the other way).
Tutorial 9b: the younger of its two faults is pinned the same way.  Its older
fault is offset by the younger one; v2 stores that block with other ids and
its field differs from the oracle's (whose fault-on-fault convention -- the
younger block as a drift column of the older fault -- is unpinned by v2).
"""
import os

import numpy as np

from oracle import cokriging as ok

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def legacy_model(tag):
    """Stacks (oracle dicts), options, grid points and the fixture arrays."""
    d = np.load(os.path.join(GOLD, "legacy_fault_models.npz"))
    g = {k[len(tag) + 1:]: d[k] for k in d.files if k.startswith(tag + "_")}
    stacks = []
    for k in range(int(g["n_stacks"])):
        sp = g[f"sp_xyz_{k}"]
        q = g[f"ori_xyz_{k}"]
        stacks.append(dict(sp_xyz=sp, counts=[int(c) for c in g[f"counts_{k}"]],
                           sp_nugget=np.full(len(sp), float(g["nugget_scalar"])), ori_xyz=q,
                           ori_grad=g[f"ori_grad_{k}"], ori_nugget=np.full(len(q), float(g["nugget_grad"])),
                           relation="fault" if bool(g[f"is_fault_{k}"]) else "erode",
                           faults=[int(f) for f in g[f"faults_{k}"]]))
    opts = dict(range=float(g["range"]), c_o=float(g["c_o"]))
    X = ok.grid_points(g["extent_r"], g["resolution"])
    return stacks, opts, X, g


def band_of(S, thr, tol=1e-6):
    rng = S.max() - S.min()
    return np.min(np.abs(S[:, None] - np.asarray(thr)[None, :]), 1) <= tol * rng


def check_fault_field(S, v2):
    """v2 = S + b (slope 1): returns the slope and the max residual / range."""
    M = np.column_stack([S, np.ones_like(S)])
    coef, *_ = np.linalg.lstsq(M, v2, rcond=None)
    res = np.abs(v2 - M @ coef).max() / np.ptp(v2)
    return coef[0], res


def test_tutorial_9a_fault_field_blocks_and_lith():
    stacks, opts, X, g = legacy_model("m9a")
    out = ok.compute_model(stacks, opts, X)
    idx = g["sample_index"]
    slope, res = check_fault_field(out["stacks"][0]["scalar"][idx], g["fault_scalar"][0])
    assert abs(slope - 1.0) < 2e-6 and res < 3e-6, (slope, res)
    fband = band_of(out["stacks"][0]["scalar"], out["stacks"][0]["thresholds"])
    fb = out["fault_blocks"][0]
    assert np.array_equal(fb[~fband], g["fault_block"][0][~fband])
    band = np.zeros(len(X), dtype=bool)
    for r in out["stacks"]:
        band |= band_of(r["scalar"], r["thresholds"])
    lith = out["lith_block"]
    basement = out["basement_id"]
    v2_ids = np.where(lith == basement, 0, lith + 1)
    mism = (v2_ids != g["lith_block"]) & ~band
    assert mism.sum() == 0
    assert len(np.unique(lith)) == 6                   # every unit is present
    assert (v2_ids != g["lith_block"]).sum() < 0.01 * len(X)


def test_tutorial_9b_younger_fault():
    stacks, opts, X, g = legacy_model("m9b")
    out = ok.compute_model(stacks, opts, X)
    idx = g["sample_index"]
    slope, res = check_fault_field(out["stacks"][0]["scalar"][idx], g["fault_scalar"][0])
    assert abs(slope - 1.0) < 2e-6 and res < 3e-6, (slope, res)
    fband = band_of(out["stacks"][0]["scalar"], out["stacks"][0]["thresholds"])
    assert np.array_equal(out["fault_blocks"][0][~fband], g["fault_block"][0][~fband])
    # no two points of the older fault's surface lie on different sides of the
    # younger fault: its fault column is dropped (weight 0) and the model solves
    st1 = stacks[1]
    n1 = 3 * len(st1["ori_xyz"]) + len(st1["sp_xyz"]) - len(st1["counts"]) + 3
    assert len(out["stacks"][1]["weights"]) == n1
