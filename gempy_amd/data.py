"""Input and output records of the interpolator.

The names follow the records the reference passes across its engine boundary
(``gempy/modules/data_manipulation/_engine_factory.py:27-58``:
``SurfacePoints(sp_coords, nugget_effect_scalar)``,
``Orientations(dip_positions, dip_gradients, nugget_effect_grad)``, the engine
grid, ``unit_values``) and the option records serialized in
``2-layers.approved.txt:149-182`` (``kernel_options``, ``evaluation_options``,
``sigmoid_slope``).  One :class:`StackInput` holds the data of one structural
group (``structural_frame.py:305-317``: points per surface, orientations per
stack, relation type).
"""
from __future__ import annotations

import copy
import enum
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np


class StackRelationType(enum.IntEnum):
    """Relation of a structural group to the older ones (values as serialized in
    the reference's approved JSON: ERODE = 1, FAULT = 3)."""
    ERODE = 1
    ONLAP = 2
    FAULT = 3
    BASEMENT = 4


class SolverForm(str, enum.Enum):
    """How the co-kriging system is solved on the device."""
    FP64_DIRECT = "fp64_direct"        # FP64 assembly + FP64 blocked LU (MFMA), the reference's solve class
    FP64_CHOL = "fp64_chol"            # FP64 Cholesky of the covariance block + Schur complement on the drift
    FP32_REFINE = "fp32_refine"        # FP32 Cholesky (MFMA) + FP64 GMRES refinement, FP64_CHOL fallback
    GMRES_LEARNED = "gmres_learned"    # matrix-free FP32 GMRES + learned preconditioner


@dataclass
class KernelOptions:
    """``kernel_options`` record (2-layers.approved.txt:150-162)."""
    range: float = 1.7
    c_o: float = 10.0
    uni_degree: int = 1
    i_res: float = 4.0
    gi_res: float = 2.0
    number_dimensions: int = 3
    kernel_function: str = "cubic"
    kernel_solver: int = 1
    compute_condition_number: bool = False
    condition_number: Optional[float] = None      # set by compute_model when compute_condition_number


@dataclass
class EvaluationOptions:
    """``evaluation_options`` record (2-layers.approved.txt:163-176)."""
    number_octree_levels: int = 1
    evaluation_chunk_size: int = 500_000
    compute_scalar_gradient: bool = False


@dataclass
class InterpolationOptions:
    kernel_options: KernelOptions = field(default_factory=KernelOptions)
    evaluation_options: EvaluationOptions = field(default_factory=EvaluationOptions)
    sigmoid_slope: float = 5_000_000.0
    solver: SolverForm = SolverForm.FP32_REFINE

    def __init__(self, kernel_options: Optional[KernelOptions] = None,
                 evaluation_options: Optional[EvaluationOptions] = None,
                 sigmoid_slope: float = 5_000_000.0, solver: SolverForm = SolverForm.FP32_REFINE,
                 number_octree_levels: Optional[int] = None,
                 compute_scalar_gradient: Optional[bool] = None):
        self.kernel_options = kernel_options or KernelOptions()
        self.evaluation_options = evaluation_options or EvaluationOptions()
        if number_octree_levels is not None:
            self.evaluation_options.number_octree_levels = int(number_octree_levels)
        if compute_scalar_gradient is not None:
            self.evaluation_options.compute_scalar_gradient = bool(compute_scalar_gradient)
        self.sigmoid_slope = sigmoid_slope
        self.solver = SolverForm(solver)

    def as_dict(self):
        k = self.kernel_options
        return {"range": k.range, "c_o": k.c_o, "uni_degree": k.uni_degree,
                "i_res": k.i_res, "gi_res": k.gi_res}


@dataclass
class RegularGrid:
    """Cell-centred lattice: ``extent`` (xmin, xmax, ymin, ymax, zmin, zmax) and
    ``resolution`` (nx, ny, nz); ravel order ``ij`` with z fastest
    (regular_grid.py:58-71,191-201)."""
    extent: Sequence[float]
    resolution: Sequence[int]

    def __post_init__(self):
        self.extent = tuple(float(v) for v in self.extent)
        self.resolution = tuple(int(v) for v in self.resolution)

    @property
    def n_points(self) -> int:
        nx, ny, nz = self.resolution
        return nx * ny * nz

    @property
    def spacing(self):
        e, r = self.extent, self.resolution
        return tuple((e[2 * d + 1] - e[2 * d]) / r[d] for d in range(3))

    @property
    def origin(self):
        """Coordinates of the first cell centre."""
        e, dx = self.extent, self.spacing
        return tuple(e[2 * d] + dx[d] / 2 for d in range(3))

    @classmethod
    def octree(cls, extent: Sequence[float], octree_levels: int, base_resolution: Optional[Sequence[int]] = None,
               legacy: bool = False) -> "RegularGrid":
        """Finest lattice of an octree: base_resolution * 2^(levels-1), the
        base resolution defaulting to round(length / min length) * 2 per axis
        ([2, 2, 2] with legacy) (gempy/core/data/grid.py:128-145,
        grid_modules/regular_grid.py:36-43)."""
        e = np.asarray(extent, dtype=np.float64).reshape(6)
        if legacy:
            base = np.array([2, 2, 2])
        elif base_resolution is None:
            lengths = np.array([e[1] - e[0], e[3] - e[2], e[5] - e[4]])
            base = np.round(lengths / lengths.min()).astype(int) * 2
        else:
            base = np.asarray(base_resolution, dtype=int)
        return cls(e, base * 2 ** (int(octree_levels) - 1))

    def values(self) -> np.ndarray:
        o, dx, r = self.origin, self.spacing, self.resolution
        axes = [o[d] + dx[d] * np.arange(r[d]) for d in range(3)]
        g = np.meshgrid(*axes, indexing="ij")
        return np.stack([c.ravel() for c in g], axis=1)


@dataclass
class StackInput:
    """Data of one structural group.

    ``sp_xyz`` rows are grouped by surface (``counts`` points each, surfaces in
    stratigraphic order, youngest first); the first point of every surface is
    its reference point.  ``faults`` lists the indices (in the model's stack
    list) of the fault stacks whose block enters this stack as a drift column.
    """
    name: str
    sp_xyz: np.ndarray
    counts: List[int]
    sp_nugget: np.ndarray
    ori_xyz: np.ndarray
    ori_grad: np.ndarray
    ori_nugget: np.ndarray
    relation: StackRelationType = StackRelationType.ERODE
    surface_names: List[str] = field(default_factory=list)
    faults: List[int] = field(default_factory=list)
    # fault stacks only: gempy_amd.faults.FiniteFaultData (tapered offset)
    finite_fault: Optional[object] = None

    def __post_init__(self):
        self.sp_xyz = np.ascontiguousarray(np.asarray(self.sp_xyz, dtype=np.float64).reshape(-1, 3))
        self.counts = [int(c) for c in self.counts]
        self.sp_nugget = np.ascontiguousarray(np.asarray(self.sp_nugget, dtype=np.float64).reshape(-1))
        self.ori_xyz = np.ascontiguousarray(np.asarray(self.ori_xyz, dtype=np.float64).reshape(-1, 3))
        self.ori_grad = np.ascontiguousarray(np.asarray(self.ori_grad, dtype=np.float64).reshape(-1, 3))
        self.ori_nugget = np.ascontiguousarray(np.asarray(self.ori_nugget, dtype=np.float64).reshape(-1))
        self.relation = StackRelationType(self.relation)
        self.faults = [int(f) for f in self.faults]

    @property
    def n_surfaces(self) -> int:
        return len(self.counts)

    @property
    def n_orientations(self) -> int:
        return self.ori_xyz.shape[0]

    @property
    def n_rest(self) -> int:
        return self.sp_xyz.shape[0] - len(self.counts)

    def copy(self) -> "StackInput":
        return copy.deepcopy(self)

    def as_oracle_dict(self):
        return {"sp_xyz": self.sp_xyz, "counts": self.counts, "sp_nugget": self.sp_nugget,
                "ori_xyz": self.ori_xyz, "ori_grad": self.ori_grad,
                "ori_nugget": self.ori_nugget,
                "relation": {StackRelationType.FAULT: "fault", StackRelationType.ONLAP: "onlap"}.get(
                    self.relation, "erode"),
                "faults": list(self.faults),
                "finite_fault": self.finite_fault.as_oracle_dict() if self.finite_fault is not None else None}


@dataclass
class GeoModelInput:
    """Everything ``compute_model`` needs: stacks (youngest first), options, grid."""
    stacks: List[StackInput]
    options: InterpolationOptions
    grid: Optional[RegularGrid] = None
    name: str = "model"
    custom_grid: Optional[np.ndarray] = None
    # previous solution's weights per stack: the initial guess of the
    # iterative refinement (the reference passes them on,
    # gempy/modules/data_manipulation/_engine_factory.py:45-48)
    weights: Optional[List[np.ndarray]] = None
    # gravity input (gempy_amd.gravity.GeophysicsInput: centred grid, tz,
    # densities); Solutions.gravity holds the result
    geophysics: Optional[object] = None

    def system_size(self, k: int = 0) -> int:
        st = self.stacks[k]
        L = {0: 0, 1: 3, 2: 9}[self.options.kernel_options.uni_degree]
        return 3 * st.n_orientations + st.n_rest + L + len(st.faults)


@dataclass
class ExportedFields:
    """Fields of one stack at one set of points (the engine's exported_fields:
    scalar_field, gradient components, scalar_field_at_surface_points)."""
    scalar_field: Optional[object] = None
    gx_field: Optional[object] = None
    gy_field: Optional[object] = None
    gz_field: Optional[object] = None
    scalar_field_at_surface_points: Optional[np.ndarray] = None


@dataclass
class StackSolution:
    """Per-stack output (the engine's ``InterpOutput``: weights, exported fields,
    ids block, smooth block; read at test_example_models_I.py:20-21,
    test_cg_solver.py:48-50, gp3_to_gp2_output.py:58-78, marching_cubes.py:36-46).
    Arrays are device tensors unless stated."""
    weights: np.ndarray
    scalar_field_at_surface_points: np.ndarray
    thresholds: np.ndarray
    scalar_field: Optional[np.ndarray] = None
    gradient: Optional[np.ndarray] = None
    ids_block: Optional[np.ndarray] = None
    solver_info: dict = field(default_factory=dict)
    mask: Optional[object] = None              # points this stack holds in the lith block
    sigmoid_slope: float = 5_000_000.0
    _block: Optional[object] = None

    @property
    def exported_fields(self) -> ExportedFields:
        g = self.gradient
        return ExportedFields(scalar_field=self.scalar_field,
                              gx_field=None if g is None else g[:, 0], gy_field=None if g is None else g[:, 1],
                              gz_field=None if g is None else g[:, 2],
                              scalar_field_at_surface_points=np.asarray(self.thresholds))

    # the dense-grid and the generic accessors of the reference read the same fields here
    exported_fields_dense_grid = exported_fields

    @property
    def scalar_fields(self) -> "StackSolution":
        return self

    @property
    def block(self):
        """Smooth ids (logistic steps of slope sigmoid_slope at the thresholds)."""
        if self._block is None and self.scalar_field is not None:
            from .engine import block_values
            self._block = block_values(self.scalar_field, self.thresholds, self.sigmoid_slope)
        return self._block

    @property
    def values_block(self):
        b = self.block
        return None if b is None else b.reshape(1, -1)


@dataclass
class OctreeLevel:
    """Points evaluated at one octree level: voxel indices at that level's
    resolution, ids at the voxel centres, the mask of voxels refined next and
    every stack's output at the voxel centres (``outputs``)."""
    resolution: tuple
    voxel_index: np.ndarray
    ids: np.ndarray
    scalar: np.ndarray
    refine: np.ndarray
    outputs: List[StackSolution] = field(default_factory=list)

    @property
    def last_output_center(self) -> Optional[StackSolution]:
        return self.outputs[-1] if self.outputs else None


@dataclass
class RawArrays:
    """Dense-grid arrays as host numpy arrays (``raw_arrays``: lith_block,
    scalar_field_matrix, block_matrix, mask_matrix, custom;
    compute_API.py:114, test_subduction.py:87)."""
    lith_block: Optional[np.ndarray] = None
    fault_block: Optional[np.ndarray] = None
    scalar_field_matrix: Optional[np.ndarray] = None
    block_matrix: Optional[np.ndarray] = None
    mask_matrix: Optional[np.ndarray] = None
    custom: Optional[np.ndarray] = None
    vertices: Optional[list] = None
    edges: Optional[list] = None

    @property
    def mask_matrix_squeezed(self):
        return self.mask_matrix


class _Root:
    def __init__(self, outputs):
        self.outputs = outputs


@dataclass
class Solutions:
    """Model output (``Solutions`` contract consumed at geo_model.py:100-127)."""
    stacks: List[StackSolution]
    lith_block: Optional[np.ndarray] = None
    fault_blocks: List[np.ndarray] = field(default_factory=list)
    scalar_field_matrix: Optional[np.ndarray] = None
    octrees_output: List[OctreeLevel] = field(default_factory=list)
    custom: Optional[np.ndarray] = None
    timings: dict = field(default_factory=dict)
    dc_meshes: Optional[list] = None
    gravity: Optional[np.ndarray] = None
    _raw: Optional[RawArrays] = None

    @property
    def root_output(self):
        """Per-stack outputs whose weights seed the next solve (_engine_factory.py:45-48)."""
        return _Root(self.stacks)

    @property
    def scalar_field_at_surface_points(self) -> np.ndarray:
        """Threshold of every surface of every stack, in order (one per element
        but the basement; geo_model.py:106-108)."""
        return np.concatenate([np.asarray(s.thresholds, dtype=np.float64) for s in self.stacks]) \
            if self.stacks else np.zeros(0)

    @property
    def _ordered_elements(self) -> List[np.ndarray]:
        """Per stack, its surfaces ordered by their thresholds from the top
        (geo_model.py:123-127)."""
        out = []
        for s in self.stacks:
            t = np.asarray(s.thresholds, dtype=np.float64)
            rev = len(t) >= 2 and t[0] < t[-1]
            out.append(np.argsort(t if rev else -t, kind="stable"))
        return out

    @property
    def block_solution_type(self) -> str:
        return "octree" if self.octrees_output else "dense"

    @property
    def raw_arrays(self) -> RawArrays:
        if self._raw is None:
            host = lambda t: None if t is None else t.detach().cpu().numpy()  # noqa: E731
            r = RawArrays()
            if self.lith_block is not None:
                r.lith_block = host(self.lith_block).astype(np.float64)
            if self.fault_blocks:
                r.fault_block = np.stack([host(f) for f in self.fault_blocks]).astype(np.float64)
            if self.scalar_field_matrix is not None:
                r.scalar_field_matrix = host(self.scalar_field_matrix)
            if self.stacks and all(s.scalar_field is not None for s in self.stacks):
                r.block_matrix = np.stack([host(s.block) for s in self.stacks])
            if self.stacks and all(s.mask is not None for s in self.stacks):
                r.mask_matrix = np.stack([host(s.mask).astype(bool) for s in self.stacks])
            if self.custom is not None:
                r.custom = host(self.custom).astype(np.float64)
            if self.dc_meshes is not None:
                r.vertices = [m.vertices for m in self.dc_meshes]
                r.edges = [m.edges for m in self.dc_meshes]
            self._raw = r
        return self._raw
