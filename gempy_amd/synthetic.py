"""Seeded synthetic geological models used by the benchmark and the tests.

Every generator draws from ``numpy.random.default_rng(seed)`` in a fixed order
and returns a :class:`gempy_amd.data.GeoModelInput`.  Sizes default to the
workloads the benchmark is quoted on and can be scaled down for parity tests.

Parameter convention (legacy GemPy v2 kriging records, e.g.
``examples/data/gempy_models/Tutorial_ch1-9a_Fault_relations/
Tutorial_ch1-9a_Fault_relations_kriging_data.csv:2``): range = sqrt(3)*extent,
c_o = range**2/14/3, interface nugget 1e-6, orientation nugget 0.01.
"""
from __future__ import annotations

import numpy as np

from .data import (GeoModelInput, InterpolationOptions, KernelOptions, RegularGrid,
                   StackInput, StackRelationType)

EXTENT = (0.0, 1000.0, 0.0, 1000.0, 0.0, 1000.0)
RANGE = float(np.sqrt(3.0) * 1000.0)
C_O = RANGE ** 2 / 14.0 / 3.0
SP_NUGGET = 1e-6
ORI_NUGGET = 0.01


def _options(uni_degree=1, **kw):
    return InterpolationOptions(
        kernel_options=KernelOptions(range=RANGE, c_o=C_O, uni_degree=uni_degree), **kw)


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _tilt(rng, n, sigma_deg):
    """Rotate unit normals n by a random angle ~ N(0, sigma) about a random axis."""
    if sigma_deg <= 0:
        return n
    ang = np.deg2rad(rng.normal(0.0, sigma_deg, size=n.shape[0]))
    rnd = rng.normal(size=n.shape)
    perp = rnd - np.sum(rnd * n, axis=1, keepdims=True) * n
    perp = _unit(perp)
    return _unit(np.cos(ang)[:, None] * n + np.sin(ang)[:, None] * perp)


def config1(seed: int = 1, resolution=(50, 50, 50)) -> GeoModelInput:
    """Plain method, small: 3 conformable surfaces z0 = 700, 500, 300 (top first),
    8 points each with z = z0 + 40 sin(2 pi x / 1000); 6 orientations on the
    z0 = 500 surface with its analytic normal; grid 50^3."""
    rng = np.random.default_rng(seed)
    xyz, counts = [], []
    for z0 in (700.0, 500.0, 300.0):
        x = rng.uniform(0, 1000, 8)
        y = rng.uniform(0, 1000, 8)
        z = z0 + 40.0 * np.sin(2 * np.pi * x / 1000.0)
        xyz.append(np.stack([x, y, z], 1))
        counts.append(8)
    xo = rng.uniform(0, 1000, 6)
    yo = rng.uniform(0, 1000, 6)
    zo = 500.0 + 40.0 * np.sin(2 * np.pi * xo / 1000.0)
    g = np.stack([-40.0 * (2 * np.pi / 1000.0) * np.cos(2 * np.pi * xo / 1000.0),
                  np.zeros(6), np.ones(6)], 1)
    sp = np.concatenate(xyz)
    stack = StackInput(
        name="strat", sp_xyz=sp, counts=counts, sp_nugget=np.full(len(sp), SP_NUGGET),
        ori_xyz=np.stack([xo, yo, zo], 1), ori_grad=_unit(g),
        ori_nugget=np.full(6, ORI_NUGGET), relation=StackRelationType.ERODE,
        surface_names=["surface_700", "surface_500", "surface_300"])
    return GeoModelInput(stacks=[stack], options=_options(),
                         grid=RegularGrid(EXTENT, resolution), name="config1")


def _fold_stack(rng, n_surfaces, n_points, n_ori, z0s, amp, lam, noise, name, x_phase=0.0):
    """One folded stack: z = z0 + amp sin(2 pi x/lam) cos(2 pi y/lam) + N(0, noise)."""
    counts = [n_points // n_surfaces + (1 if k < n_points % n_surfaces else 0)
              for k in range(n_surfaces)]
    xyz = []
    for z0, c in zip(z0s, counts):
        x = rng.uniform(0, 1000, c)
        y = rng.uniform(0, 1000, c)
        z = z0 + amp * np.sin(2 * np.pi * (x + x_phase) / lam) * np.cos(2 * np.pi * y / lam)
        z = z + rng.normal(0.0, noise, c) if noise > 0 else z
        xyz.append(np.stack([x, y, z], 1))
    k = rng.integers(0, n_surfaces, n_ori)
    xo = rng.uniform(0, 1000, n_ori)
    yo = rng.uniform(0, 1000, n_ori)
    sx, cy = np.sin(2 * np.pi * (xo + x_phase) / lam), np.cos(2 * np.pi * yo / lam)
    zo = np.asarray(z0s)[k] + amp * sx * cy
    dzdx = amp * (2 * np.pi / lam) * np.cos(2 * np.pi * (xo + x_phase) / lam) * cy
    dzdy = -amp * (2 * np.pi / lam) * sx * np.sin(2 * np.pi * yo / lam)
    g = _unit(np.stack([-dzdx, -dzdy, np.ones(n_ori)], 1))
    sp = np.concatenate(xyz)
    return StackInput(
        name=name, sp_xyz=sp, counts=counts, sp_nugget=np.full(len(sp), SP_NUGGET),
        ori_xyz=np.stack([xo, yo, zo], 1), ori_grad=g, ori_nugget=np.full(n_ori, ORI_NUGGET),
        relation=StackRelationType.ERODE,
        surface_names=[f"{name}_{i}" for i in range(n_surfaces)])


def config3(seed: int = 3, n_points: int = 20_000, n_ori: int = 5_000, n_surfaces: int = 10,
            resolution=(200, 200, 200)) -> GeoModelInput:
    """Precision form: one folded stack, 10 surfaces x 2,000 points,
    z = z0 + 150 sin(2 pi x/600) cos(2 pi y/600) + N(0, 5 m), 5,000 orientations
    with analytic normals; N = 3*5000 + 19990 + 3 = 34,993; grid 200^3."""
    rng = np.random.default_rng(seed)
    z0s = np.linspace(800.0, 200.0, n_surfaces)
    st = _fold_stack(rng, n_surfaces, n_points, n_ori, z0s, 150.0, 600.0, 5.0, "fold")
    return GeoModelInput(stacks=[st], options=_options(), grid=RegularGrid(EXTENT, resolution),
                         name="config3")


def config4(seed: int = 4, n_points: int = 100_000, n_ori: int = 10_000, n_surfaces: int = 20,
            resolution=(256, 256, 256)) -> GeoModelInput:
    """Learned-preconditioner form: one folded stack of 20 surfaces with a random
    fold wavelength lam ~ U(300, 1200) m; N = 3 n_ori + n_points - 20 + 3."""
    rng = np.random.default_rng(seed)
    lam = rng.uniform(300.0, 1200.0)
    z0s = np.linspace(850.0, 150.0, n_surfaces)
    st = _fold_stack(rng, n_surfaces, n_points, n_ori, z0s, 100.0, lam, 5.0, "fold")
    return GeoModelInput(stacks=[st], options=_options(), grid=RegularGrid(EXTENT, resolution),
                         name="config4")


def training_system(seed: int, n_target: int) -> GeoModelInput:
    """One training system for the learned preconditioner (seeds 1000-2999):
    a config-4-style fold with size N ~ n_target."""
    rng = np.random.default_rng(seed)
    n_surf = int(rng.integers(4, 21))
    n_ori = max(4, n_target // 8)
    n_pts = max(2 * n_surf, n_target - 3 * n_ori)
    lam = rng.uniform(300.0, 1200.0)
    z0s = np.linspace(850.0, 150.0, n_surf)
    st = _fold_stack(rng, n_surf, n_pts, n_ori, z0s, 100.0, lam, 5.0, "fold")
    return GeoModelInput(stacks=[st], options=_options(), grid=RegularGrid(EXTENT, (8, 8, 8)),
                         name=f"train{seed}")


# -------------------------------------------------------------------- config 2
FAULT_POINT = np.array([500.0, 500.0, 500.0])
FAULT_NORMAL = np.array([np.sin(np.deg2rad(60.0)), 0.0, np.cos(np.deg2rad(60.0))])
FAULT_THROW = -60.0


def _sinusoid_params(rng):
    amp = rng.uniform(20.0, 80.0, 3)
    kx = rng.uniform(-1.0, 1.0, 3) / 700.0
    ky = rng.uniform(-1.0, 1.0, 3) / 700.0
    ph = rng.uniform(0, 2 * np.pi, 3)
    return amp, kx, ky, ph


def _sinusoid(params, x, y):
    amp, kx, ky, ph = params
    arg = 2 * np.pi * (np.outer(x, kx) + np.outer(y, ky)) + ph
    z = (amp * np.sin(arg)).sum(1)
    dzdx = (amp * 2 * np.pi * kx * np.cos(arg)).sum(1)
    dzdy = (amp * 2 * np.pi * ky * np.cos(arg)).sum(1)
    return z, dzdx, dzdy


def _layer_stack(rng, name, z0s, n_points, n_ori, faulted, tilt_deg=2.0):
    params = _sinusoid_params(rng)
    ns = len(z0s)
    counts = [n_points // ns + (1 if k < n_points % ns else 0) for k in range(ns)]
    xyz = []
    for z0, c in zip(z0s, counts):
        x = rng.uniform(0, 1000, c)
        y = rng.uniform(0, 1000, c)
        z = z0 + _sinusoid(params, x, y)[0]
        if faulted:
            # hanging wall (positive side of the fault plane) moved down by the throw
            z = _offset(x, y, z)
        xyz.append(np.stack([x, y, z], 1))
    k = rng.integers(0, ns, n_ori)
    xo = rng.uniform(0, 1000, n_ori)
    yo = rng.uniform(0, 1000, n_ori)
    zs, dzdx, dzdy = _sinusoid(params, xo, yo)
    zo = np.asarray(z0s)[k] + zs
    if faulted:
        zo = _offset(xo, yo, zo)
    g = _tilt(rng, _unit(np.stack([-dzdx, -dzdy, np.ones(n_ori)], 1)), tilt_deg)
    sp = np.concatenate(xyz)
    return StackInput(
        name=name, sp_xyz=sp, counts=counts, sp_nugget=np.full(len(sp), SP_NUGGET),
        ori_xyz=np.stack([xo, yo, zo], 1), ori_grad=g, ori_nugget=np.full(n_ori, ORI_NUGGET),
        relation=StackRelationType.ERODE,
        surface_names=[f"{name}_{i}" for i in range(ns)])


def _offset(x, y, z):
    side = (np.stack([x, y, z], 1) - FAULT_POINT) @ FAULT_NORMAL
    return np.where(side > 0, z + FAULT_THROW, z)


def config2(seed: int = 2, scale: float = 1.0, resolution=(200, 200, 200),
            octree_levels: int = 4) -> GeoModelInput:
    """Plain method at scale: a planar fault (200 points / 40 orientations) and
    three stacks (youngest first: an erosive 2-surface stack, then two
    conformable 3-surface stacks, both offset by the fault) of 1,600 points /
    320 orientations each; 5,000 points and 1,000 orientations over 9
    surfaces; z = z0 + sum of 3 seeded sinusoids (amplitude U(20, 80) m);
    normals analytic + N(0, 2 deg) tilt.  ``scale`` shrinks the point counts."""
    rng = np.random.default_rng(seed)
    n_f, no_f = max(6, int(200 * scale)), max(2, int(40 * scale))
    n_s, no_s = max(12, int(1600 * scale)), max(3, int(320 * scale))
    # fault plane points: random (y, z), x from the plane equation
    yf = rng.uniform(0, 1000, n_f)
    zf = rng.uniform(50, 950, n_f)
    xf = FAULT_POINT[0] - (FAULT_NORMAL[2] / FAULT_NORMAL[0]) * (zf - FAULT_POINT[2])
    yo = rng.uniform(100, 900, no_f)
    zo = rng.uniform(100, 900, no_f)
    xo = FAULT_POINT[0] - (FAULT_NORMAL[2] / FAULT_NORMAL[0]) * (zo - FAULT_POINT[2])
    gf = _tilt(rng, np.tile(FAULT_NORMAL, (no_f, 1)), 2.0)
    fault = StackInput(
        name="fault", sp_xyz=np.stack([xf, yf, zf], 1), counts=[n_f],
        sp_nugget=np.full(n_f, SP_NUGGET), ori_xyz=np.stack([xo, yo, zo], 1), ori_grad=gf,
        ori_nugget=np.full(no_f, ORI_NUGGET), relation=StackRelationType.FAULT,
        surface_names=["fault_0"])
    # surfaces of every stack top first (youngest first), as the ids count them
    erosive = _layer_stack(rng, "erosive", [870.0, 800.0], n_s, no_s, faulted=False)
    upper = _layer_stack(rng, "upper", [680.0, 610.0, 540.0], n_s, no_s, faulted=True)
    lower = _layer_stack(rng, "lower", [380.0, 300.0, 220.0], n_s, no_s, faulted=True)
    upper.faults = [0]
    lower.faults = [0]
    opts = _options(number_octree_levels=octree_levels)
    return GeoModelInput(stacks=[fault, erosive, upper, lower], options=opts,
                         grid=RegularGrid(EXTENT, resolution), name="config2")


def config5(seed: int = 5, scale: float = 1.0, resolution=(128, 128, 128), n_obs: int = 10_000):
    """Differentiation form: config-2 geometry on a 128^3 grid; the misfit is
    taken on the 'lower' stack (offset by the fault, whose block is held
    fixed) at 10,000 observation points ~ U(box) against the field of the same
    stack with every surface z0 shifted by N(0, 10 m).
    Returns (model, index of the differentiated stack, perturbed copy of that
    stack, observation points)."""
    base = config2(seed=2, scale=scale, resolution=resolution)
    rng = np.random.default_rng(seed)
    k = 3
    lower = base.stacks[k]
    obs = rng.uniform(0, 1000, (n_obs, 3))
    shift = np.repeat(rng.normal(0, 10.0, len(lower.counts)), lower.counts)
    pert = lower.copy()
    pert.sp_xyz = lower.sp_xyz.copy()
    pert.sp_xyz[:, 2] += shift
    base.name = "config5"
    return base, k, pert, obs
