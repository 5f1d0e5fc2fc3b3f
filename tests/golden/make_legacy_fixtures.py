"""Write tests/golden/legacy_fault_models.npz from the reference's legacy
GemPy-v2 (FP64) fault models (run in the build container, where the reference
tree is readable; the fixture travels, the reference does not).

Source: examples/data/gempy_models/Tutorial_ch1-9a_Fault_relations/ and
Tutorial_ch1-9b_Fault_relations/ (files *_surface_points.csv,
*_orientations.csv, *_kriging_data.csv:2, *_rescaling_data.csv:2,
*_extent.npy, *_resolution.npy, *_scalar_field_faults.npy,
*_fault_blocks.npy, *_lith_block.npy).  Inputs are taken in the rescaled
frame the files record (columns X_r, Y_r, Z_r: (X - centers)/1600 + 0.5001);
the kriging record gives range 1732.05 (rescaled: /1600), C_o 71428.57
(unscaled), nuggets 1e-6 (interface) / 0.01 (orientation), drift degree 1.

Stored per model: the stacks (youngest first: faults, then layer series),
their fault relations, the 50^3 grid in the rescaled frame, and the recorded
outputs:
* the fault scalar fields at a seeded sample of 20,000 cells (v2 stored them
  with an additive constant: compared up to a constant, scale 1);
* the fault blocks (v2 stores sigmoid-smoothed values; rounded to ids);
* 9a only: the lith block (v2 unit ids, rounded; 0 = basement).
"""
from __future__ import annotations

import csv
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/examples/data/gempy_models"
N_SAMPLE = 20_000


def _read(base, name):
    with open(f"{base}_{name}.csv", newline="") as f:
        return list(csv.DictReader(f))


def _model(tag, series, surfaces, faults):
    base = os.path.join(REF, f"Tutorial_ch1-{tag}_Fault_relations", f"Tutorial_ch1-{tag}_Fault_relations")
    sp, ori = _read(base, "surface_points"), _read(base, "orientations")
    kr = _read(base, "kriging_data")[0]
    rs = _read(base, "rescaling_data")[0]
    resc = float(rs["rescaling factor"])
    center = np.array([float(v) for v in rs["centers"].strip("[]").split()])
    out = {"range": float(kr["range"]) / resc, "c_o": float(kr["$C_o$"]),
           "nugget_grad": float(kr["nugget grad"]), "nugget_scalar": float(kr["nugget scalar"]),
           "rescale": resc, "center": center, "n_stacks": len(series)}
    for k, s in enumerate(series):
        pts, counts = [], []
        for name in surfaces[s]:
            rows = [r for r in sp if r["surface"] == name]
            pts += [[float(r["X_r"]), float(r["Y_r"]), float(r["Z_r"])] for r in rows]
            counts.append(len(rows))
        o = [r for r in ori if r["series"] == s]
        out[f"sp_xyz_{k}"] = np.array(pts)
        out[f"counts_{k}"] = np.array(counts)
        out[f"ori_xyz_{k}"] = np.array([[float(r["X_r"]), float(r["Y_r"]), float(r["Z_r"])] for r in o])
        out[f"ori_grad_{k}"] = np.array([[float(r["G_x"]), float(r["G_y"]), float(r["G_z"])] for r in o])
        out[f"faults_{k}"] = np.array(faults[s], dtype=np.int64)
        out[f"is_fault_{k}"] = s.startswith("fault")
    ext = np.load(f"{base}_extent.npy")
    res = np.load(f"{base}_resolution.npy")
    out["extent_r"] = np.array([(ext[2 * d + e] - center[d]) / resc + 0.5001 for d in range(3) for e in range(2)])
    out["resolution"] = res.astype(np.int64)
    n = int(np.prod(res))
    idx = np.sort(np.random.default_rng(9).choice(n, N_SAMPLE, replace=False))
    out["sample_index"] = idx
    sf = np.load(f"{base}_scalar_field_faults.npy")
    fb = np.load(f"{base}_fault_blocks.npy")
    out["fault_scalar"] = sf[:, idx]
    out["fault_block"] = np.rint(fb).astype(np.int8)
    return out


def main():
    m9a = _model("9a", ["fault_series1", "series2", "series1"],
                 {"fault_series1": ["fault1"], "series2": ["rock5", "rock4"],
                  "series1": ["rock3", "rock2", "rock1"]},
                 {"fault_series1": [], "series2": [], "series1": [0]})
    base = os.path.join(REF, "Tutorial_ch1-9a_Fault_relations", "Tutorial_ch1-9a_Fault_relations")
    lb = np.load(f"{base}_lith_block.npy")
    m9a["lith_block"] = np.rint(lb[:int(np.prod(m9a["resolution"]))]).astype(np.int8)
    m9b = _model("9b", ["fault_series2", "fault_series1", "series1"],
                 {"fault_series2": ["fault2"], "fault_series1": ["fault1"],
                  "series1": ["rock4", "rock3", "rock2", "rock1"]},
                 {"fault_series2": [], "fault_series1": [0], "series1": [0, 1]})
    out = {f"m9a_{k}": v for k, v in m9a.items()}
    out.update({f"m9b_{k}": v for k, v in m9b.items()})
    np.savez_compressed(os.path.join(HERE, "legacy_fault_models.npz"), **out)
    print("wrote", os.path.join(HERE, "legacy_fault_models.npz"))


if __name__ == "__main__":
    main()
