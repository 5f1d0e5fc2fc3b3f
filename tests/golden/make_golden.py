"""Write the golden fixtures under tests/golden/ (run in the build container,
where the reference tree is readable; the fixtures travel, the reference does not).

* two_layers.npz  -- inputs of test/test_modules/test_geophysics/test_gravity.py:11-60
  mapped with the recorded input_transform (2-layers.approved.txt: position
  (-6, -0, -2.51), scale (1/12, 0.16778523489932887, 1/12)), options range 1.7,
  c_o 10 (2-layers.approved.txt:151-152); expected scalar_field_at_interface
  0.09000000000000002 and -0.2483333333333333 (2-layers.approved.txt:29,46).
* anticline.npz   -- ANTICLINE example (gempy/API/examples_generator.py:165-195):
  examples/data/input_data/jan_models/model2_{surface_points,orientations}.csv,
  stack ('rock2', 'rock1'), recorded transform position (-500, -500, -510),
  scale 0.0005 (test_custom_grid.test_custom_grid.verify/fold.approved.txt),
  default nuggets 2e-5 / 0.01 (surface_points.py:11, orientations.py:11),
  pole vectors from azimuth/dip/polarity (io_API.py:96-101); expected ids at
  the 8 box corners [3,3,3,3,1,1,1,1] (test_custom_grid.py:50-65).
* config1_oracle.npz -- the oracle's own FP64 outputs for config 1 (regression
  data for the GPU parity tests; pinned by the tests above, not independent).
* transforms.npz -- raw input points and the recorded input_transform
  (position, scale) of the 2-layers, anticline and 2.5-D (Model1) models.
* greenstone/ -- the Greenstone model's .gempy file and its legacy v2 tables.
* combination/ -- the COMBINATION example's input tables
  (examples/data/input_data/jan_models/model7_{surface_points,orientations}.csv,
  gempy/API/examples_generator.py:243-288); the reference's marching-cubes test
  records its mesh vertex counts 600 / 860 / 1256 / 1680 on a 40x20x20 dense grid
  (test/test_modules/test_marching_cubes.py:14-47).
"""
from __future__ import annotations

import csv
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"


def _read_csv(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def two_layers():
    pos = np.array([-6.0, -0.0, -2.51])
    scale = np.array([0.08333333333333333, 0.16778523489932887, 0.08333333333333333])
    sp_raw = np.array([[3, 0, 3.05], [9, 0, 3.05], [3, 0, 1.02], [9, 0, 1.02]], dtype=float)
    ori_raw = np.array([[6, 0, 4]], dtype=float)
    np.savez(os.path.join(HERE, "two_layers.npz"),
             sp_xyz=(sp_raw + pos) * scale, counts=np.array([2, 2]), sp_nugget=np.full(4, 2e-5),
             ori_xyz=(ori_raw + pos) * scale, ori_grad=np.array([[0.0, 0.0, 1.0]]),
             ori_nugget=np.array([0.01]), range=1.7, c_o=10.0,
             expected_thresholds=np.array([0.09000000000000002, -0.2483333333333333]))


def anticline():
    base = os.path.join(REF, "examples/data/input_data/jan_models")
    sp = _read_csv(os.path.join(base, "model2_surface_points.csv"))
    ori = _read_csv(os.path.join(base, "model2_orientations.csv"))
    pos, s = np.array([-500.0, -500.0, -510.0]), 0.0005
    pts, counts = [], []
    for name in ("rock2", "rock1"):
        rows = [[float(r["X"]), float(r["Y"]), float(r["Z"])] for r in sp if r["formation"] == name]
        pts += rows
        counts.append(len(rows))
    sp_xyz = (np.array(pts) + pos) * s
    xyz = np.array([[float(r["X"]), float(r["Y"]), float(r["Z"])] for r in ori])
    dip = np.array([float(r["dip"]) for r in ori])
    az = np.array([float(r["azimuth"]) for r in ori])
    pol = np.array([float(r["polarity"]) for r in ori])
    g = np.stack([np.sin(np.deg2rad(dip)) * np.sin(np.deg2rad(az)) * pol,
                  np.sin(np.deg2rad(dip)) * np.cos(np.deg2rad(az)) * pol,
                  np.cos(np.deg2rad(dip)) * pol], 1)
    corners = np.array([[0, 0, 0], [1000, 0, 0], [0, 1000, 0], [1000, 1000, 0],
                        [0, 0, 1000], [1000, 0, 1000], [0, 1000, 1000], [1000, 1000, 1000]], float)
    np.savez(os.path.join(HERE, "anticline.npz"),
             sp_xyz=sp_xyz, counts=np.array(counts), sp_nugget=np.full(len(sp_xyz), 2e-5),
             ori_xyz=(xyz + pos) * s, ori_grad=g, ori_nugget=np.full(len(g), 0.01),
             range=1.7, c_o=10.0, corners=(corners + pos) * s,
             expected_ids=np.array([3, 3, 3, 3, 1, 1, 1, 1]))


def config1_oracle():
    sys.path.insert(0, ROOT)
    from gempy_amd import synthetic
    from oracle import cokriging as ok
    m = synthetic.config1()
    st = m.stacks[0].as_oracle_dict()
    r = ok.compute_stack(st, m.options.as_dict(), m.grid.values(), gradient=True)
    np.savez_compressed(os.path.join(HERE, "config1_oracle.npz"), weights=r["weights"],
                        thresholds=r["thresholds"], scalar=r["scalar"], ids=r["ids"],
                        A=r["A"], b=r["b"])


def _approved_transform(rel):
    import json
    text = open(os.path.join(REF, rel)).read()
    d = json.loads(text[text.index("{"):])

    def find(o):
        if isinstance(o, dict):
            if "input_transform" in o:
                return o["input_transform"]
            for v in o.values():
                r = find(v)
                if r is not None:
                    return r
        if isinstance(o, list):
            for v in o:
                r = find(v)
                if r is not None:
                    return r
        return None
    t = find(d)
    return np.array(t["position"], dtype=float), np.array(t["scale"], dtype=float)


def transforms():
    """Raw input points and the recorded input_transform of three reference models:
    2-layers (test_gravity.py:11-60; 2-layers.approved.txt), the anticline
    (jan_models/model2_*.csv; test_custom_grid...fold.approved.txt) and the 2.5-D
    model's first element (examples_generator.py:33-60; Model1.approved.txt)."""
    out = {}
    # 2-layers
    sp = np.array([[3, 0, 3.05], [9, 0, 3.05], [3, 0, 1.02], [9, 0, 1.02]], dtype=float)
    ori = np.array([[6, 0, 4]], dtype=float)
    pos, sc = _approved_transform("test/test_modules/test_geophysics/test_gravity.test_gravity.verify/"
                                  "2-layers.approved.txt")
    out.update(two_layers_sp=sp, two_layers_ori=ori, two_layers_position=pos, two_layers_scale=sc)
    # anticline
    base = os.path.join(REF, "examples/data/input_data/jan_models")
    sp = np.array([[float(r["X"]), float(r["Y"]), float(r["Z"])]
                   for r in _read_csv(os.path.join(base, "model2_surface_points.csv"))])
    ori = np.array([[float(r["X"]), float(r["Y"]), float(r["Z"])]
                    for r in _read_csv(os.path.join(base, "model2_orientations.csv"))])
    pos, sc = _approved_transform("test/test_modules/test_grids/test_custom_grid.test_custom_grid.verify/"
                                  "fold.approved.txt")
    out.update(fold_sp=sp, fold_ori=ori, fold_position=pos, fold_scale=sc)
    # 2.5-D model, first element only (update_transform is called right after it)
    sp = np.array([[223, 0.01, -94], [458, 0, -197], [612, 0, -14]], dtype=float)
    ori = np.array([[350, 1, -300]], dtype=float)
    pos, sc = _approved_transform("test/test_modules/test_grids/test_grids_sections.test_topography_II.verify/"
                                  "Model1.approved.txt")
    out.update(model1_sp=sp, model1_ori=ori, model1_position=pos, model1_scale=sc)
    np.savez(os.path.join(HERE, "transforms.npz"), **out)


def greenstone():
    """The reference's Greenstone model as data: the .gempy file
    (examples/data/gempy_models/Greenstone.gempy) and the same model's legacy
    v2 surface-point / orientation tables (examples/data/gempy_models/
    Greenstone/Greenstone_{surface_points,orientations}.csv)."""
    import shutil
    dst = os.path.join(HERE, "greenstone")
    os.makedirs(dst, exist_ok=True)
    src = os.path.join(REF, "examples/data/gempy_models")
    shutil.copyfile(os.path.join(src, "Greenstone.gempy"), os.path.join(dst, "Greenstone.gempy"))
    for name in ("surface_points", "orientations"):
        shutil.copyfile(os.path.join(src, "Greenstone", f"Greenstone_{name}.csv"),
                        os.path.join(dst, f"Greenstone_{name}.csv"))


def _jan_model(model, folder):
    import shutil
    dst = os.path.join(HERE, folder)
    os.makedirs(dst, exist_ok=True)
    src = os.path.join(REF, "examples/data/input_data/jan_models")
    for name in ("surface_points", "orientations"):
        shutil.copyfile(os.path.join(src, f"{model}_{name}.csv"), os.path.join(dst, f"{model}_{name}.csv"))


def combination():
    """The COMBINATION example's surface-point and orientation tables."""
    _jan_model("model7", "combination")


def one_fault():
    """The ONE_FAULT example's tables (examples_generator.py:198-240)."""
    _jan_model("model5", "one_fault")


if __name__ == "__main__":
    two_layers()
    anticline()
    config1_oracle()
    transforms()
    greenstone()
    combination()
    one_fault()
    print("fixtures written to", HERE)
