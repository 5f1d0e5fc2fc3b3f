"""Loader of the reference's ``.gempy`` model files.

Format (gempy/modules/serialization/save_load.py:138-189): a zip archive of
``header.json`` (the model record: structural frame, grid, interpolation
options, input transform), ``input.bin`` (the surface-point table followed by
the orientation table, raw little-endian records of
``SurfacePointsTable.dt`` = X, Y, Z f8, id i4, nugget f8 and
``OrientationsTable.dt`` = X, Y, Z, G_x, G_y, G_z f8, id i4, nugget f8,
gempy/core/data/surface_points.py:25, orientations.py:24; their byte lengths in
``structural_frame.binary_meta_data``) and ``grid.bin`` (custom-grid then
topography values, f8; lengths in ``grid.binary_meta_data``).

Only data are read (json and numpy.frombuffer); nothing in the file is
executed.  The model is mapped to this library's input records the way the
reference feeds its engine (_engine_factory.py:14-58): points and orientations
in the file's input frame (``input_transform``), groups in file order
(youngest first), relations from ``structural_relation`` and the fault
relations (``fault_relations``: a special case or the names of the offset
groups, structural_frame.py:243-273, 590-598).
"""
from __future__ import annotations

import io as _io
import json
import os
import zipfile
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from .adapters import Transform
from .data import (EvaluationOptions, GeoModelInput, InterpolationOptions, KernelOptions, RegularGrid, SolverForm,
                   StackInput, StackRelationType)

SP_DTYPE = np.dtype([("X", "<f8"), ("Y", "<f8"), ("Z", "<f8"), ("id", "<i4"), ("nugget", "<f8")])
ORI_DTYPE = np.dtype([("X", "<f8"), ("Y", "<f8"), ("Z", "<f8"), ("G_x", "<f8"), ("G_y", "<f8"), ("G_z", "<f8"),
                      ("id", "<i4"), ("nugget", "<f8")])
# FaultsRelationSpecialCase values (enum.auto order, structural_group.py:13-16)
OFFSET_FORMATIONS, OFFSET_NONE, OFFSET_ALL = 1, 2, 3


@dataclass
class LoadedModel:
    """A .gempy file: the engine input and what the file records besides."""
    model: GeoModelInput
    name: str
    transform: Transform
    element_names: List[List[str]]
    scalar_field_at_interface: List[List[Optional[float]]]     # recorded per element (stale if is_dirty)
    is_dirty: bool
    octree_extent: Optional[List[float]] = None
    octree_resolution: Optional[List[int]] = None
    octree_levels: int = 1
    header: Dict = field(default_factory=dict)


def _fault_matrix(groups: List[dict]) -> np.ndarray:
    """fault_relations of the groups as a boolean matrix M[i, j]: group i
    offsets group j (j after i in file order)."""
    n = len(groups)
    names = [g.get("name") for g in groups]
    m = np.zeros((n, n), dtype=bool)
    for i, g in enumerate(groups):
        if int(g.get("structural_relation", 1)) != int(StackRelationType.FAULT):
            continue
        fr = g.get("fault_relations", OFFSET_ALL)
        if fr is None or fr == OFFSET_ALL:
            m[i, i + 1:] = True
        elif fr == OFFSET_NONE:
            pass
        elif fr == OFFSET_FORMATIONS:
            for j in range(i + 1, n):
                m[i, j] = int(groups[j].get("structural_relation", 1)) != int(StackRelationType.FAULT)
        elif isinstance(fr, list):
            for name in fr:
                j = names.index(name)
                if j <= i:
                    raise ValueError(f"fault group {names[i]!r} cannot offset the older group {name!r}")
                m[i, j] = True
        else:
            raise ValueError(f"group {names[i]!r}: invalid fault relation {fr!r}")
    return m


def _element_id(e: dict) -> int:
    for key in ("_id", "id"):
        if key in e and e[key] is not None:
            return int(e[key])
    nm = e.get("surface_points", {}).get("name_id_map", {})
    if e.get("name") in nm:
        return int(nm[e["name"]])
    raise ValueError(f"element {e.get('name')!r} has no id")


def load_gempy(path_or_bytes, solver: SolverForm = SolverForm.FP32_REFINE) -> LoadedModel:
    """Read a .gempy file (path or bytes) into a LoadedModel."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        data = bytes(path_or_bytes)
    else:
        path = os.fspath(path_or_bytes)
        if not path.lower().endswith(".gempy"):
            raise ValueError(f"Invalid file extension: {os.path.splitext(path)[1]!r}. Expected: '.gempy'")
        if not os.path.exists(path):
            raise FileNotFoundError(f"File not found: {path}")
        with open(path, "rb") as f:
            data = f.read()
    with zipfile.ZipFile(_io.BytesIO(data), "r") as zf:
        header = json.loads(zf.read("header.json").decode("utf-8"))
        input_raw = zf.read("input.bin")
        grid_raw = zf.read("grid.bin")
    sf = header["structural_frame"]
    meta = sf.get("binary_meta_data", {})
    n_sp_bytes = int(meta.get("sp_binary_length", 0))
    n_ori_bytes = int(meta.get("ori_binary_length", len(input_raw) - n_sp_bytes))
    if n_sp_bytes % SP_DTYPE.itemsize or n_ori_bytes % ORI_DTYPE.itemsize or \
            n_sp_bytes + n_ori_bytes > len(input_raw):
        raise ValueError("input.bin does not hold whole surface-point / orientation records")
    sp = np.frombuffer(input_raw[:n_sp_bytes], dtype=SP_DTYPE)
    ori = np.frombuffer(input_raw[n_sp_bytes:n_sp_bytes + n_ori_bytes], dtype=ORI_DTYPE)
    t = header.get("input_transform") or {}
    tr = Transform(position=t.get("position", [0, 0, 0]), scale=t.get("scale", [1, 1, 1]),
                   rotation=t.get("rotation", [0, 0, 0]))
    groups = sf["structural_groups"]
    fm = _fault_matrix(groups)
    stacks, enames, recorded = [], [], []
    for gi, g in enumerate(groups):
        pts, counts, nug, q, G, qn, names, rec = [], [], [], [], [], [], [], []
        for e in g["elements"]:
            if not e.get("is_active", True):
                continue
            eid = _element_id(e)
            s = sp[sp["id"] == eid]
            if len(s):
                pts.append(np.stack([s["X"], s["Y"], s["Z"]], 1))
                nug.append(np.asarray(s["nugget"], dtype=np.float64))
                counts.append(len(s))
                names.append(e.get("name"))
                rec.append(e.get("scalar_field_at_interface"))
            o = ori[ori["id"] == eid]
            if len(o):
                q.append(np.stack([o["X"], o["Y"], o["Z"]], 1))
                G.append(np.stack([o["G_x"], o["G_y"], o["G_z"]], 1))
                qn.append(np.asarray(o["nugget"], dtype=np.float64))
        P = tr.apply(np.concatenate(pts)) if pts else np.zeros((0, 3))
        Q = tr.apply(np.concatenate(q)) if q else np.zeros((0, 3))
        GG = tr.transform_gradient(np.concatenate(G)) if G else np.zeros((0, 3))
        rel = int(g.get("structural_relation", 1))
        stacks.append(StackInput(
            name=g.get("name", f"group{gi}"), sp_xyz=P, counts=counts,
            sp_nugget=np.concatenate(nug) if nug else np.zeros(0), ori_xyz=Q, ori_grad=GG,
            ori_nugget=np.concatenate(qn) if qn else np.zeros(0),
            relation=StackRelationType.ERODE if rel == int(StackRelationType.BASEMENT) else StackRelationType(rel),
            surface_names=names, faults=[int(i) for i in np.nonzero(fm[:gi, gi])[0]]))
        enames.append(names)
        recorded.append(rec)
    io_opts = header.get("_interpolation_options") or header.get("interpolation_options") or {}
    ko = io_opts.get("kernel_options", {})
    eo = io_opts.get("evaluation_options", {})
    kopts = KernelOptions(range=float(ko.get("range", 1.7)), c_o=float(ko.get("c_o", 10.0)),
                          uni_degree=int(ko.get("uni_degree", 1)), i_res=float(ko.get("i_res", 4.0)),
                          gi_res=float(ko.get("gi_res", 2.0)))
    levels = int(eo.get("_number_octree_levels", eo.get("number_octree_levels", 1)) or 1)
    opts = InterpolationOptions(kopts, EvaluationOptions(number_octree_levels=1,
                                                         compute_scalar_gradient=bool(
                                                             eo.get("compute_scalar_gradient", False))),
                                sigmoid_slope=float(io_opts.get("sigmoid_slope", 5e6)), solver=solver)
    gh = header.get("grid", {}) or {}
    grid, oct_ext, oct_res = None, None, None
    dense = gh.get("_dense_grid")
    octree = gh.get("_octree_grid")
    if octree:
        oct_ext, oct_res = list(octree["extent"]), [int(r) for r in octree["resolution"]]
    box = dense or octree
    if box:
        ext = np.asarray(box["extent"], dtype=np.float64)
        lo = tr.apply(ext[[0, 2, 4]][None])[0]
        hi = tr.apply(ext[[1, 3, 5]][None])[0]
        grid = RegularGrid([lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]], [int(r) for r in box["resolution"]])
    custom = None
    gmeta = gh.get("binary_meta_data", {}) or {}
    n_custom = int(gmeta.get("custom_grid_binary_length", 0))
    if n_custom:
        custom = tr.apply(np.frombuffer(grid_raw[:n_custom], dtype="<f8").reshape(-1, 3))
    model = GeoModelInput(stacks=stacks, options=opts, grid=grid,
                          name=(header.get("meta") or {}).get("name", "model"), custom_grid=custom)
    return LoadedModel(model=model, name=model.name, transform=tr, element_names=enames,
                       scalar_field_at_interface=recorded, is_dirty=bool(sf.get("is_dirty", False)),
                       octree_extent=oct_ext, octree_resolution=oct_res, octree_levels=levels, header=header)


# --------------------------------------------------------------------------
# CSV tables (gempy/API/io_API.py: the ImporterHelper path of create_geomodel)
# --------------------------------------------------------------------------
_COLUMNS = {"X": ("X", "x"), "Y": ("Y", "y"), "Z": ("Z", "z"), "azimuth": ("azimuth", "Azimuth"),
            "dip": ("dip", "Dip"), "polarity": ("polarity", "Polarity"),
            "formation": ("formation", "Formation", "surface"), "G_x": ("G_x", "gradient_x"),
            "G_y": ("G_y", "gradient_y"), "G_z": ("G_z", "gradient_z")}
DEFAULT_SP_NUGGET, DEFAULT_ORI_NUGGET = 2e-5, 0.01      # surface_points.py:11, orientations.py:11


def _read_table(path_or_text) -> Dict[str, list]:
    import csv
    if isinstance(path_or_text, str) and "\n" not in path_or_text:
        with open(path_or_text, newline="") as f:
            rows = list(csv.DictReader(f))
    else:
        rows = list(csv.DictReader(_io.StringIO(path_or_text)))
    out: Dict[str, list] = {}
    for std, names in _COLUMNS.items():
        for n in names:
            if rows and n in rows[0]:
                out[std] = [r[n] for r in rows]
                break
    return out


def pole_vectors(azimuth, dip, polarity) -> np.ndarray:
    """Gradient (pole) vectors of orientations given as azimuth / dip /
    polarity in degrees (io_API.py:96-101, manipulate_points.py:324-337)."""
    az, dp, pol = (np.deg2rad(np.asarray(azimuth, dtype=np.float64)), np.deg2rad(np.asarray(dip, dtype=np.float64)),
                   np.asarray(polarity, dtype=np.float64))
    return np.stack([np.sin(dp) * np.sin(az) * pol, np.sin(dp) * np.cos(az) * pol, np.cos(dp) * pol], 1)


def read_csv_model(surface_points, orientations, mapping: Dict[str, tuple], extent, resolution,
                   relations: Optional[Dict[str, StackRelationType]] = None,
                   fault_relations: Optional[np.ndarray] = None, transform: Optional[Transform] = None,
                   options: Optional[InterpolationOptions] = None, name: str = "model") -> LoadedModel:
    """A model from surface-point and orientation tables (CSV path or text;
    columns X, Y, Z, formation and G_x/G_y/G_z or azimuth/dip/polarity) and a
    group -> elements mapping (map_stack_to_surfaces; groups youngest first,
    elements in the given order).  Points keep the table order per element
    (the first is the element's reference point); nuggets take the reference's
    defaults; the transform defaults to Transform.from_input_points of all
    points and orientations; relations default to ERODE; fault_relations[i, j]
    (bool) says that group i offsets group j."""
    sp = _read_table(surface_points)
    ori = _read_table(orientations)
    xyz = np.array([sp["X"], sp["Y"], sp["Z"]], dtype=np.float64).T
    fm = np.asarray(sp["formation"])
    qx = np.array([ori["X"], ori["Y"], ori["Z"]], dtype=np.float64).T
    qf = np.asarray(ori["formation"])
    if "G_x" in ori:
        grads = np.array([ori["G_x"], ori["G_y"], ori["G_z"]], dtype=np.float64).T
    else:
        grads = pole_vectors(ori["azimuth"], ori["dip"], ori["polarity"])
    tr = transform or Transform.from_input_points(xyz, qx)
    groups = list(mapping.items())
    n = len(groups)
    fr = np.zeros((n, n), dtype=bool) if fault_relations is None else np.asarray(fault_relations, dtype=bool)
    stacks, enames = [], []
    for gi, (gname, elements) in enumerate(groups):
        elements = (elements,) if isinstance(elements, str) else tuple(elements)
        P, counts, names, Q, G = [], [], [], [], []
        for e in elements:
            sel = fm == e
            if sel.any():
                P.append(xyz[sel])
                counts.append(int(sel.sum()))
                names.append(e)
            qs = qf == e
            if qs.any():
                Q.append(qx[qs])
                G.append(grads[qs])
        rel = (relations or {}).get(gname, StackRelationType.ERODE)
        Pc = np.concatenate(P) if P else np.zeros((0, 3))
        Qc = np.concatenate(Q) if Q else np.zeros((0, 3))
        stacks.append(StackInput(name=gname, sp_xyz=tr.apply(Pc), counts=counts,
                                 sp_nugget=np.full(len(Pc), DEFAULT_SP_NUGGET), ori_xyz=tr.apply(Qc),
                                 ori_grad=tr.transform_gradient(np.concatenate(G)) if G else np.zeros((0, 3)),
                                 ori_nugget=np.full(len(Qc), DEFAULT_ORI_NUGGET), relation=rel, surface_names=names,
                                 faults=[int(i) for i in np.nonzero(fr[:gi, gi])[0]]))
        enames.append(names)
    ext = np.asarray(extent, dtype=np.float64)
    lo = tr.apply(ext[[0, 2, 4]][None])[0]
    hi = tr.apply(ext[[1, 3, 5]][None])[0]
    grid = RegularGrid([lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]], [int(r) for r in resolution])
    model = GeoModelInput(stacks=stacks, options=options or InterpolationOptions(), grid=grid, name=name)
    return LoadedModel(model=model, name=name, transform=tr, element_names=enames,
                       scalar_field_at_interface=[[None] * len(e) for e in enames], is_dirty=True)
