"""The .gempy loader (CPU).

* The reference's Greenstone.gempy decodes to exactly the points and
  orientations of the same model's legacy v2 tables (text CSV, an independent
  encoding), grouped by structural group in file order, in the file's input
  frame (position -(721180.0625, 6893494.75, -3951.36), scale 1/97123).
* A file written here in the same layout (zip of header.json + raw
  SurfacePointsTable / OrientationsTable records) round-trips, including
  fault relations given as group names and as special cases.
* Error behaviour of save_load.py:121-132: wrong extension -> ValueError,
  missing file -> FileNotFoundError.
* The loaded model solves with the oracle (thresholds are finite and ordered).
"""
import csv
import io
import json
import os
import zipfile

import numpy as np
import pytest

from gempy_amd.data import StackRelationType
from gempy_amd.io import ORI_DTYPE, SP_DTYPE, load_gempy
from oracle import cokriging as ok

GS = os.path.join(os.path.dirname(__file__), "golden", "greenstone")


def _csv(name):
    with open(os.path.join(GS, f"Greenstone_{name}.csv"), newline="") as f:
        return list(csv.DictReader(f))


def test_greenstone_matches_legacy_tables():
    lm = load_gempy(os.path.join(GS, "Greenstone.gempy"))
    m = lm.model
    assert lm.name == "Greenstone" and lm.is_dirty
    assert [s.name for s in m.stacks] == ["EarlyGranite_Series", "BIF_Series", "SimpleMafic_Series"]
    assert lm.element_names == [["EarlyGranite"], ["SimpleMafic2", "SimpleBIF"], ["SimpleMafic1"]]
    assert [s.counts for s in m.stacks] == [[27], [11, 15], [17]]
    assert [s.n_orientations for s in m.stacks] == [26, 2, 13]
    assert all(s.relation == StackRelationType.ERODE and s.faults == [] for s in m.stacks)
    np.testing.assert_allclose(lm.transform.scale, 1.0 / 97123.0, rtol=1e-6)
    k = m.options.kernel_options
    assert (k.range, k.c_o, k.uni_degree) == (1.7, 10.0, 1)
    # every point of the legacy table, per element, in the input frame
    sp = _csv("surface_points")
    for st, names in zip(m.stacks, lm.element_names):
        back = lm.transform.apply_inverse(st.sp_xyz)
        off = 0
        for name, c in zip(names, st.counts):
            rows = np.array([[float(r["X"]), float(r["Y"]), float(r["Z"])] for r in sp if r["surface"] == name])
            got = back[off:off + c]
            off += c
            assert len(rows) == c
            # same set of points (order within an element may differ)
            a = got[np.lexsort(got.T[::-1])]
            b = rows[np.lexsort(rows.T[::-1])]
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-6)
    ori = _csv("orientations")
    all_q = np.concatenate([lm.transform.apply_inverse(s.ori_xyz) for s in m.stacks])
    all_g = np.concatenate([s.ori_grad for s in m.stacks])
    rows = np.array([[float(r[c]) for c in ("X", "Y", "Z", "G_x", "G_y", "G_z")] for r in ori])
    mine = np.concatenate([all_q, all_g], 1)
    np.testing.assert_allclose(mine[np.lexsort(mine.T[::-1])], rows[np.lexsort(rows.T[::-1])], rtol=0, atol=1e-6)
    # the dense box is the octree grid's extent in the input frame
    assert lm.octree_extent == [696000.0, 747000.0, 6863000.0, 6930000.0, -20000.0, 200.0]
    # the recorded interface values are kept (the file is marked dirty: they
    # belong to an earlier state of the data)
    assert len(lm.scalar_field_at_interface[1]) == 2


def test_greenstone_solves_with_the_oracle():
    lm = load_gempy(os.path.join(GS, "Greenstone.gempy"))
    m = lm.model
    X = m.grid.values()[::97]
    res = ok.compute_model([s.as_oracle_dict() for s in m.stacks], m.options.as_dict(), X)
    for r in res["stacks"]:
        assert np.all(np.isfinite(r["thresholds"]))
        assert np.linalg.norm(r["A"] @ r["weights"] - r["b"]) <= 1e-9 * np.linalg.norm(r["b"])
    assert len(np.unique(res["lith_block"])) >= 3


def _write_gempy(groups, sp_rows, ori_rows, transform=None, extent=None):
    sp = np.array(sp_rows, dtype=SP_DTYPE)
    ori = np.array(ori_rows, dtype=ORI_DTYPE)
    header = {"meta": {"name": "tiny"},
              "structural_frame": {"structural_groups": groups, "is_dirty": False,
                                   "binary_meta_data": {"sp_binary_length": sp.nbytes,
                                                        "ori_binary_length": ori.nbytes}},
              "grid": {"_octree_grid": None,
                       "_dense_grid": {"resolution": [4, 3, 2], "extent": extent or [0, 10, 0, 10, 0, 10]},
                       "binary_meta_data": {"custom_grid_binary_length": 24, "topography_binary_length": 0}},
              "input_transform": transform or {"position": [-5, -5, -5], "rotation": [0, 0, 0],
                                               "scale": [0.05, 0.05, 0.05]},
              "_interpolation_options": {"kernel_options": {"range": 1.7, "c_o": 10.0, "uni_degree": 1},
                                         "evaluation_options": {"_number_octree_levels": 1}}}
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w", compression=zipfile.ZIP_DEFLATED) as zf:
        zf.writestr("header.json", json.dumps(header))
        zf.writestr("input.bin", sp.tobytes() + ori.tobytes())
        zf.writestr("grid.bin", np.array([1.0, 2.0, 3.0]).tobytes())
    return buf.getvalue()


def _el(name, eid):
    return {"name": name, "_id": eid, "is_active": True}


@pytest.mark.parametrize("relation,expect", [(None, [[], [0], [0]]), (3, [[], [0], [0]]), (2, [[], [], []]),
                                             (1, [[], [], [0]]), (["g2"], [[], [], [0]])])
def test_round_trip_and_fault_relations(relation, expect):
    groups = [{"name": "fault", "structural_relation": 3, "fault_relations": relation, "elements": [_el("f", 7)]},
              {"name": "g1", "structural_relation": 3, "fault_relations": 2, "elements": [_el("f2", 8)]},
              {"name": "g2", "structural_relation": 1, "fault_relations": None,
               "elements": [_el("a", 11), _el("b", 12)]}]
    sp = [(1, 2, 3, 7, 1e-5), (2, 2, 3, 7, 1e-5), (5, 5, 5, 8, 1e-5), (6, 5, 5, 8, 1e-5),
          (0, 0, 8, 11, 2e-5), (3, 4, 8, 11, 2e-5), (0, 0, 4, 12, 2e-5), (4, 4, 4, 12, 2e-5), (9, 1, 4, 12, 2e-5)]
    ori = [(1, 1, 1, 1, 0, 0, 7, 0.01), (5, 5, 5, 0, 1, 0, 8, 0.01), (2, 2, 6, 0, 0, 1, 11, 0.02)]
    lm = load_gempy(_write_gempy(groups, sp, ori))
    m = lm.model
    assert [s.faults for s in m.stacks] == expect
    assert [s.relation for s in m.stacks] == [StackRelationType.FAULT, StackRelationType.FAULT,
                                              StackRelationType.ERODE]
    st = m.stacks[2]
    assert st.counts == [2, 3]
    np.testing.assert_allclose(st.sp_xyz[0], (np.array([0, 0, 8.0]) - 5) * 0.05)
    np.testing.assert_allclose(st.sp_nugget, 2e-5)
    np.testing.assert_allclose(st.ori_grad, [[0, 0, 1]])
    np.testing.assert_allclose(st.ori_nugget, [0.02])
    np.testing.assert_allclose(m.grid.extent, [-0.25, 0.25] * 3)
    assert m.grid.resolution == (4, 3, 2)
    np.testing.assert_allclose(m.custom_grid, [[(1 - 5) * 0.05, (2 - 5) * 0.05, (3 - 5) * 0.05]])


def test_errors(tmp_path):
    with pytest.raises(ValueError):
        load_gempy(str(tmp_path / "model.zip"))
    with pytest.raises(FileNotFoundError):
        load_gempy(str(tmp_path / "missing.gempy"))
    groups = [{"name": "g", "structural_relation": 1, "elements": [_el("a", 1)]}]
    data = bytearray(_write_gempy(groups, [(0, 0, 0, 1, 1e-5)], []))
    p = tmp_path / "ok.gempy"
    p.write_bytes(bytes(data))
    assert load_gempy(str(p)).model.stacks[0].counts == [1]


COMB = os.path.join(os.path.dirname(__file__), "golden", "combination")


def combination_model(resolution=(40, 20, 20)):
    """The reference's COMBINATION example (examples_generator.py:243-288):
    a fault offsetting two erosive series."""
    from gempy_amd.io import read_csv_model
    return read_csv_model(os.path.join(COMB, "model7_surface_points.csv"), os.path.join(COMB, "model7_orientations.csv"),
                          {"Fault_Series": ("fault",), "Strat_Series1": ("rock3",),
                           "Strat_Series2": ("rock2", "rock1")},
                          [0, 2500, 0, 1000, 0, 1000], resolution,
                          relations={"Fault_Series": StackRelationType.FAULT},
                          fault_relations=[[0, 1, 1], [0, 0, 0], [0, 0, 0]], name="combination")


# vertex counts of the reference's marching-cubes test (test_marching_cubes.py:44-47)
COMBINATION_VERTICES = [600, 860, 1256, 1680]


def test_csv_model_and_pole_vectors():
    from gempy_amd.io import pole_vectors
    np.testing.assert_allclose(pole_vectors([90, 0, 30], [0, 90, 90], [1, 1, -1]),
                               [[0, 0, 1], [0, 1, 0], [-0.5, -np.sqrt(3) / 2, 0]], atol=1e-15)
    lm = combination_model()
    m = lm.model
    assert [s.counts for s in m.stacks] == [[3], [15], [39, 45]]
    assert [s.n_orientations for s in m.stacks] == [1, 1, 6]
    assert [s.faults for s in m.stacks] == [[], [0], [0]]
    # Transform.from_input_points of the 102 points and 8 orientations
    np.testing.assert_allclose(lm.transform.position, [-1150, -500, -525])
    np.testing.assert_allclose(lm.transform.scale, 1 / 4600)


def test_combination_mesh_vertex_counts_with_the_oracle():
    """The oracle's lattice-edge vertex sets of the COMBINATION model on the
    reference test's 40x20x20 dense grid: per surface at its threshold, the
    non-fault stacks within their masks (the oldest holding the basement),
    a cube meshed when its far corner is in the mask.  The fault and rock3
    counts equal the reference's exactly; rock2 / rock1 come within 0.5%
    (1260 / 1688 against 1256 / 1680: the reference's skimage 'lewiner'
    extraction, not available here, drops a few vertices this rule keeps)."""
    from oracle import meshing as om
    m = combination_model().model
    g = m.grid
    res = ok.compute_model([s.as_oracle_dict() for s in m.stacks], m.options.as_dict(), g.values())
    counts = []
    for k, (st, r) in enumerate(zip(m.stacks, res["stacks"])):
        mask = None if st.relation == StackRelationType.FAULT else res["mask_matrix"][k].reshape(g.resolution)
        for t in r["thresholds"]:
            counts.append(len(om.crossing_vertices(r["scalar"].reshape(g.resolution), t, g.origin, g.spacing, mask)))
    assert counts[:2] == COMBINATION_VERTICES[:2]
    for c, ref in zip(counts[2:], COMBINATION_VERTICES[2:]):
        assert abs(c - ref) <= 0.005 * ref, (counts, COMBINATION_VERTICES)
