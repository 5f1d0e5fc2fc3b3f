"""Whole-model GPU tests through compute_model: small reference models with the
default engine configuration, multi-fault models against the FP64 oracle,
the learned solve form, and the error word of the sync-free triangular solves.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest
import torch

from oracle import cokriging as ok

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _two_layers_model(solver=None):
    from gempy_amd import GeoModelInput, InterpolationOptions, KernelOptions, StackInput
    d = np.load(os.path.join(GOLD, "two_layers.npz"))
    st = StackInput("Strat_Series", d["sp_xyz"], list(d["counts"]), d["sp_nugget"], d["ori_xyz"],
                    d["ori_grad"], d["ori_nugget"])
    kw = {} if solver is None else {"solver": solver}
    opts = InterpolationOptions(KernelOptions(range=float(d["range"]), c_o=float(d["c_o"])), **kw)
    return GeoModelInput([st], opts), d["expected_thresholds"]


@pytest.mark.parametrize("solver", [None, "gmres_learned", "fp64_direct"])
def test_compute_model_two_layers_default_config(gpu, solver):
    """The reference's 2-layer model (1 orientation, 2 rest points: fewer data
    points than a neighbourhood) through the default engine configuration;
    the recorded interface values 0.09 / -0.248333 (planar closed form)."""
    say(f"compute_model two_layers solver={solver}")
    from gempy_amd import GemPyEngineConfig, compute_model
    m, expected = _two_layers_model(solver)
    sol = compute_model(m, GemPyEngineConfig())
    say(f"two_layers {solver}: thresholds {sol.stacks[0].thresholds} info {sol.stacks[0].solver_info}")
    np.testing.assert_allclose(sol.stacks[0].thresholds, expected, rtol=0, atol=1e-7)


def _two_fault_model(scale=0.05, res=20):
    """config-2 geometry plus a second planar fault that the first offsets;
    both faults offset the two lower stacks."""
    from gempy_amd import synthetic
    from gempy_amd.data import StackInput, StackRelationType
    m = synthetic.config2(scale=scale, resolution=(res,) * 3, octree_levels=1)
    rng = np.random.default_rng(77)
    n_pts, n_ori = 12, 3
    nrm = np.array([0.0, np.sin(np.deg2rad(70.0)), np.cos(np.deg2rad(70.0))])
    x = rng.uniform(0, 1000, n_pts)
    z = rng.uniform(80, 920, n_pts)
    y = 420.0 - (nrm[2] / nrm[1]) * (z - 500.0)
    z = synthetic._offset(x, y, z)
    xo = rng.uniform(100, 900, n_ori)
    zo = rng.uniform(150, 850, n_ori)
    yo = 420.0 - (nrm[2] / nrm[1]) * (zo - 500.0)
    f2 = StackInput("fault2", np.stack([x, y, z], 1), [n_pts], np.full(n_pts, synthetic.SP_NUGGET),
                    np.stack([xo, yo, zo], 1), np.tile(nrm, (n_ori, 1)), np.full(n_ori, synthetic.ORI_NUGGET),
                    relation=StackRelationType.FAULT, faults=[0])
    fault, erosive, upper, lower = m.stacks
    upper.faults = [0, 1]
    lower.faults = [0, 1]
    m.stacks = [fault, f2, erosive, upper, lower]
    return m


def test_compute_model_two_faults_fault_on_fault_matches_oracle(gpu):
    """Two fault stacks, the second offset by the first (its own fault-drift
    column), both entering the two lower stacks: lith block and both fault
    blocks against the oracle, every stack's field to 1e-6 of its range."""
    say("compute_model two faults (fault on fault) vs oracle, 20^3")
    from gempy_amd import GemPyEngineConfig, compute_model
    m = _two_fault_model()
    X = m.grid.values()
    res = ok.compute_model([s.as_oracle_dict() for s in m.stacks], m.options.as_dict(), X)
    for dt in ("float64", "float32"):
        sol = compute_model(m, GemPyEngineConfig(dtype=dt))
        band = np.zeros(len(X), dtype=bool)
        for s_gpu, r in zip(sol.stacks, res["stacks"]):
            rng_s = r["scalar"].max() - r["scalar"].min()
            band |= np.min(np.abs(r["scalar"][:, None] - r["thresholds"][None, :]), 1) <= 1e-6 * rng_s
            err = np.abs(s_gpu.scalar_field.cpu().numpy() - r["scalar"]).max() / rng_s
            assert err <= 1e-6, err
        lith = sol.lith_block.cpu().numpy()
        mism = int(((lith != res["lith_block"]) & ~band).sum())
        say(f"two faults {dt}: lith mismatches outside band {mism}, band {int(band.sum())}, "
            f"forms {[s.solver_info['form'] for s in sol.stacks]}")
        assert mism == 0
        assert len(sol.fault_blocks) == 2
        for fb, rb in zip(sol.fault_blocks, res["fault_blocks"]):
            fbn = fb.cpu().numpy()
            assert int(((fbn != rb) & ~band).sum()) == 0
        # the second fault block really is offset by the first
        assert len(np.unique(res["fault_blocks"][1])) == 2


def test_learned_form_through_compute_model(gpu):
    """SolverForm.GMRES_LEARNED reached from compute_model: the network-
    preconditioned matrix-free GMRES attempt is recorded, and the field meets
    the 1e-6 bound against the oracle (directly or through the fallback)."""
    say("compute_model learned form, config1-like stack")
    from gempy_amd import GemPyEngineConfig, compute_model, synthetic
    from gempy_amd.data import SolverForm
    m = synthetic.config3(n_points=600, n_ori=150, resolution=(16, 16, 16))
    m.options.solver = SolverForm.GMRES_LEARNED
    sol = compute_model(m, GemPyEngineConfig())
    info = sol.stacks[0].solver_info
    say(f"learned form: {info}")
    assert "learned" in info and info["learned"]["iterations"] > 0
    st = m.stacks[0].as_oracle_dict()
    r = ok.compute_stack(st, m.options.as_dict(), m.grid.values())
    rng_s = r["scalar"].max() - r["scalar"].min()
    err = np.abs(sol.stacks[0].scalar_field.cpu().numpy() - r["scalar"]).max() / rng_s
    say(f"learned form: max|dS|/range {err:.3e}")
    assert err <= 1e-6


def test_triangular_solve_error_word_is_reported(gpu):
    """A wait that reached its spin limit (error word forced from the host)
    is returned as GK_ERR_HIP by the FP64 Cholesky form and by the public
    Cholesky solve, and the next calls run normally again."""
    say("trsv error word: FP64_CHOL solve and gk_cholesky_solve_f64")
    from gempy_amd import _lib, synthetic
    from gempy_amd.engine import DeviceStack, _stream, handle
    lib = _lib.load()
    m = synthetic.config3(n_points=400, n_ori=100, resolution=(8, 8, 8))
    ds = DeviceStack(m.stacks[0], m.options.kernel_options, [500.0, 500.0, 500.0])
    w0, _ = ds.solve(_lib.GK_SOLVE_FP64_CHOL)
    _lib.check("set_error", lib.gk_test_trsv_set_error(handle(0), 1, _stream()))
    with pytest.raises(_lib.GkError) as e:
        ds.solve(_lib.GK_SOLVE_FP64_CHOL)
    assert e.value.code == -2 and "spin limit" in str(e.value)
    w1, info = ds.solve(_lib.GK_SOLVE_FP64_CHOL)
    assert info["relres"] < 1e-8
    assert torch.allclose(w0, w1, rtol=0, atol=1e-9 * float(w0.abs().max()))
    # public solve on a small SPD matrix
    n = 512
    rng = np.random.default_rng(3)
    Bm = rng.standard_normal((n, n))
    A = Bm @ Bm.T / n + np.eye(n)
    tA = torch.tensor(A, dtype=torch.float64, device=gpu).contiguous()
    dinv = torch.zeros(_lib.GK_CHOL_PLANES * n * _lib.GK_BLOCK, dtype=torch.float64, device=gpu)
    info32 = C.c_int32(0)
    _lib.check("chol", lib.gk_cholesky_f64(handle(0), C.c_void_p(tA.data_ptr()), n, n, C.c_void_p(dinv.data_ptr()),
                                           C.byref(info32), _stream()))
    b = rng.standard_normal(n)
    tb = torch.tensor(b, dtype=torch.float64, device=gpu)
    _lib.check("solve", lib.gk_cholesky_solve_f64(handle(0), C.c_void_p(tA.data_ptr()), C.c_void_p(dinv.data_ptr()),
                                                  n, C.c_void_p(tb.data_ptr()), _stream()))
    _lib.check("set_error", lib.gk_test_trsv_set_error(handle(0), 1, _stream()))
    tb2 = torch.tensor(b, dtype=torch.float64, device=gpu)
    rc = lib.gk_cholesky_solve_f64(handle(0), C.c_void_p(tA.data_ptr()), C.c_void_p(dinv.data_ptr()), n,
                                   C.c_void_p(tb2.data_ptr()), _stream())
    assert rc == -2
    tb3 = torch.tensor(b, dtype=torch.float64, device=gpu)
    _lib.check("solve", lib.gk_cholesky_solve_f64(handle(0), C.c_void_p(tA.data_ptr()), C.c_void_p(dinv.data_ptr()),
                                                  n, C.c_void_p(tb3.data_ptr()), _stream()))
    x = np.linalg.solve(A, b)
    assert np.abs(tb3.cpu().numpy() - x).max() <= 1e-10 * np.abs(x).max()
    assert np.abs(tb.cpu().numpy() - x).max() <= 1e-10 * np.abs(x).max()


def test_kept_factorization_is_tied_to_its_system(gpu):
    """gk_solve_rhs refuses a system whose data live elsewhere (same sizes,
    another stack), instead of reusing the kept factor of the last solve."""
    say("kept factorization vs another system of the same sizes")
    from gempy_amd import _lib, synthetic
    from gempy_amd.engine import DeviceStack, _ptr, _stream
    lib = _lib.load()
    m = synthetic.config3(n_points=300, n_ori=60, resolution=(8, 8, 8))
    a = DeviceStack(m.stacks[0], m.options.kernel_options, [500.0, 500.0, 500.0])
    b = DeviceStack(m.stacks[0].copy(), m.options.kernel_options, [500.0, 500.0, 500.0])
    a.solve(_lib.GK_SOLVE_FP64_CHOL)
    rhs = torch.ones(a.N, dtype=torch.float64, device=gpu)
    x = torch.empty_like(rhs)
    info = _lib.GkSolveInfo()
    assert lib.gk_solve_rhs(a.h, C.byref(a.sys), _ptr(rhs), _ptr(x), C.byref(info), _stream()) == 0
    assert lib.gk_solve_rhs(b.h, C.byref(b.sys), _ptr(rhs), _ptr(x), C.byref(info), _stream()) == -1


@pytest.mark.parametrize("tag", ["m9a", "m9b"])
def test_legacy_fault_models_match_v2_and_oracle(gpu, tag):
    """The reference's legacy v2 fault models (Tutorial_ch1-9a / -9b) through
    compute_model: every stack's field against the oracle (1e-6 of its range),
    the youngest fault's field against v2's recorded one (up to v2's additive
    constant), its block against v2's, and (9a) the lith block against v2's
    outside the +-1e-6 band."""
    say(f"legacy fault model {tag} through compute_model")
    from test_oracle_legacy import band_of, check_fault_field, legacy_model
    from gempy_amd import GemPyEngineConfig, compute_model
    from gempy_amd.data import (GeoModelInput, InterpolationOptions, KernelOptions, RegularGrid, StackInput,
                                StackRelationType)
    stacks, opts, X, g = legacy_model(tag)
    res = ok.compute_model(stacks, opts, X)
    st_in = [StackInput(f"s{k}", s["sp_xyz"], s["counts"], s["sp_nugget"], s["ori_xyz"], s["ori_grad"],
                        s["ori_nugget"], relation=StackRelationType.FAULT if s["relation"] == "fault"
                        else StackRelationType.ERODE, faults=s["faults"]) for k, s in enumerate(stacks)]
    m = GeoModelInput(st_in, InterpolationOptions(KernelOptions(range=opts["range"], c_o=opts["c_o"])),
                      grid=RegularGrid(g["extent_r"], g["resolution"]))
    for dt in ("float64", "float32"):
        sol = compute_model(m, GemPyEngineConfig(dtype=dt))
        band = np.zeros(len(X), dtype=bool)
        for k, (s, r) in enumerate(zip(sol.stacks, res["stacks"])):
            S = s.scalar_field.cpu().numpy()
            err = np.abs(S - r["scalar"]).max() / np.ptp(r["scalar"])
            say(f"{tag} {dt} stack {k}: max|dS|/range vs oracle {err:.3e} info {s.solver_info}")
            assert err <= 1e-6
            band |= band_of(r["scalar"], r["thresholds"])
        S0 = sol.stacks[0].scalar_field.cpu().numpy()
        slope, resid = check_fault_field(S0[g["sample_index"]], g["fault_scalar"][0])
        say(f"{tag} {dt}: youngest fault vs v2: slope {slope:.8f} residual/range {resid:.3e}")
        assert abs(slope - 1.0) < 2e-6 and resid < 3e-6
        fb = sol.fault_blocks[0].cpu().numpy()
        assert np.array_equal(fb[~band], g["fault_block"][0][~band])
        if "lith_block" in g:
            lith = sol.lith_block.cpu().numpy()
            v2 = np.where(lith == lith.max(), 0, lith + 1)
            mism = int(((v2 != g["lith_block"]) & ~band).sum())
            say(f"{tag} {dt}: lith mismatches vs v2 outside band {mism} (band {int(band.sum())})")
            assert mism == 0


def test_warm_start_fewer_iterations_same_field(gpu):
    """The previous solution's weights as the initial guess (the reference
    passes them on, _engine_factory.py:45-48): a perturbed system solved in the
    FP32 form from the old weights needs fewer refinement iterations and gives
    the same field as the cold solve."""
    say("warm start: perturbed config3-style system (N ~ 6k)")
    from gempy_amd import GemPyEngineConfig, compute_model, synthetic
    m = synthetic.config3(n_points=3_000, n_ori=1_000, resolution=(24, 24, 24))
    sol0 = compute_model(m, GemPyEngineConfig(dtype="float32"))
    pert = synthetic.config3(n_points=3_000, n_ori=1_000, resolution=(24, 24, 24))
    rng = np.random.default_rng(12)
    st = pert.stacks[0]
    g = st.ori_grad + rng.normal(0, 1e-5, st.ori_grad.shape)         # a small update of the data
    st.ori_grad = g / np.linalg.norm(g, axis=1, keepdims=True)
    cold = compute_model(pert, GemPyEngineConfig(dtype="float32"))
    pert.weights = [s.weights.cpu().numpy() for s in sol0.root_output.outputs]
    warm = compute_model(pert, GemPyEngineConfig(dtype="float32"))
    ic, iw = cold.stacks[0].solver_info, warm.stacks[0].solver_info
    Sc = cold.stacks[0].scalar_field.cpu().numpy()
    Sw = warm.stacks[0].scalar_field.cpu().numpy()
    err = np.abs(Sc - Sw).max() / np.ptp(Sc)
    say(f"warm start: cold {ic['iterations']} iterations (relres {ic['relres']:.2e}), warm {iw['iterations']} "
        f"(relres {iw['relres']:.2e}, warm_start={iw['warm_start']}); max|dS|/range {err:.2e}")
    assert iw["warm_start"] and not ic["warm_start"]
    assert iw["iterations"] < ic["iterations"]
    assert iw["relres"] <= 1e-6 and err <= 1e-6


def test_solutions_view_matches_oracle(gpu):
    """raw_arrays / masks / smooth block / octree outputs of a multi-stack model
    with an onlapping stack against the oracle's lith block and mask matrix."""
    say("solutions view + onlap: config2 scale 0.05, 24^3, 2 octree levels")
    from gempy_amd import GemPyEngineConfig, compute_model, synthetic
    from gempy_amd.data import StackRelationType
    m = synthetic.config2(scale=0.05, resolution=(24, 24, 24), octree_levels=2)
    m.stacks[2].relation = StackRelationType.ONLAP          # 'upper' onlaps 'lower'
    X = m.grid.values()
    res = ok.compute_model([s.as_oracle_dict() for s in m.stacks], m.options.as_dict(), X)
    sol = compute_model(m, GemPyEngineConfig(dtype="float64"))
    raw = sol.raw_arrays
    band = np.zeros(len(X), dtype=bool)
    for r in res["stacks"]:
        band |= np.min(np.abs(r["scalar"][:, None] - r["thresholds"][None, :]), 1) <= 1e-6 * np.ptp(r["scalar"])
    mism = int(((raw.lith_block != res["lith_block"]) & ~band).sum())
    mmask = int(((raw.mask_matrix != np.stack(res["mask_matrix"])) & ~band[None]).sum())
    say(f"onlap: lith mismatches outside band {mism}, mask mismatches {mmask}, band {int(band.sum())}")
    assert mism == 0 and mmask == 0
    # the onlap rule changes the lith block against pure erosion
    ero = [dict(s.as_oracle_dict(), relation="erode" if s.relation == StackRelationType.ONLAP
                else s.as_oracle_dict()["relation"]) for s in m.stacks]
    res_e = ok.compute_model(ero, m.options.as_dict(), X)
    assert int((res_e["lith_block"] != res["lith_block"]).sum()) > 0
    assert raw.scalar_field_matrix.shape == (4, len(X)) and raw.fault_block.shape == (1, len(X))
    # smooth block: equals the ids away from the thresholds (steep slope)
    blk = raw.block_matrix
    for k, r in enumerate(res["stacks"]):
        far = np.min(np.abs(r["scalar"][:, None] - r["thresholds"][None, :]), 1) > 1e-3 * np.ptp(r["scalar"])
        np.testing.assert_allclose(blk[k][far], r["ids"][far], rtol=0, atol=1e-9)
    assert np.all(np.diff(np.sort(blk[1])) >= 0)
    # octree outputs: every stack at the voxel centres of each level
    assert len(sol.octrees_output) == 2
    lvl = sol.octrees_output[-1]
    assert len(lvl.outputs) == 4
    assert lvl.last_output_center is lvl.outputs[-1]
    dense_S = sol.stacks[3].scalar_field
    np.testing.assert_allclose(lvl.outputs[3].exported_fields.scalar_field.cpu().numpy(),
                               dense_S[lvl.voxel_index].cpu().numpy(), rtol=0, atol=1e-9 * float(dense_S.abs().max()))
    np.testing.assert_array_equal(sol.octrees_output[0].outputs[3].weights.cpu().numpy(),
                                  sol.stacks[3].weights.cpu().numpy())


def test_greenstone_gempy_file_matches_oracle(gpu):
    """The reference's Greenstone.gempy, loaded and computed on the 64^3 box in
    both forms, against the oracle on sampled cells (fields within 1e-6 of
    their range, lith ids outside the band)."""
    say("Greenstone.gempy through compute_model (64^3)")
    from gempy_amd import GemPyEngineConfig, compute_model
    from gempy_amd.io import load_gempy
    lm = load_gempy(os.path.join(GOLD, "greenstone", "Greenstone.gempy"))
    m = lm.model
    X = m.grid.values()
    idx = np.random.default_rng(0).choice(len(X), 6000, replace=False)
    res = ok.compute_model([s.as_oracle_dict() for s in m.stacks], m.options.as_dict(), X[idx])
    for dt in ("float64", "float32"):
        sol = compute_model(m, GemPyEngineConfig(dtype=dt))
        band = np.zeros(len(idx), dtype=bool)
        for k, (s, r) in enumerate(zip(sol.stacks, res["stacks"])):
            S = s.scalar_field.cpu().numpy()[idx]
            err = np.abs(S - r["scalar"]).max() / np.ptp(r["scalar"])
            say(f"greenstone {dt} stack {k}: max|dS|/range {err:.3e} thresholds {np.round(s.thresholds, 6)} "
                f"(file: {lm.scalar_field_at_interface[k]}) info {s.solver_info['form']}/{s.solver_info['iterations']}")
            assert err <= 1e-6
            band |= np.min(np.abs(r["scalar"][:, None] - r["thresholds"][None, :]), 1) <= 1e-6 * np.ptp(r["scalar"])
        lith = sol.lith_block.cpu().numpy()[idx]
        assert int(((lith != res["lith_block"]) & ~band).sum()) == 0


def test_octree_levels_match_oracle(gpu):
    """compute_model's octree levels (config 2 geometry, base 6^3, 3 levels)
    against the oracle's restatement of the refinement: the same active
    voxels, refinement masks and ids at the voxel centres at every level."""
    say("octree vs oracle: config2 scale 0.05, base 6^3, 3 levels")
    from gempy_amd import GemPyEngineConfig, compute_model, synthetic
    from gempy_amd.data import RegularGrid
    m = synthetic.config2(scale=0.05, resolution=(24, 24, 24), octree_levels=3)
    stacks = [s.as_oracle_dict() for s in m.stacks]

    def fields_at(pts):
        r = ok.compute_model(stacks, m.options.as_dict(), pts)
        return np.stack([r["lith_block"]] + r["fault_blocks"])
    ref = ok.octree_levels(fields_at, m.grid.extent, (6, 6, 6), 3)
    sol = compute_model(m, GemPyEngineConfig(dtype="float64"))
    octs = sol.octrees_output
    assert len(octs) == len(ref) == 3
    for lvl, (o, r) in enumerate(zip(octs, ref)):
        say(f"octree level {lvl}: res {o.resolution} voxels {o.voxel_index.numel()} (oracle {len(r['voxels'])})")
        assert o.resolution == r["resolution"]
        np.testing.assert_array_equal(o.voxel_index.cpu().numpy(), r["voxels"])
        np.testing.assert_array_equal(o.ids.cpu().numpy(), r["ids"][0])
        if r["refine"] is not None:
            np.testing.assert_array_equal(o.refine.cpu().numpy(), r["refine"])
    assert RegularGrid.octree((0, 1000, 0, 1000, 0, 1000), 3).resolution == (8, 8, 8)


def test_config4_scaled_learned_form_matches_oracle(gpu):
    """Config 4 (the learned-preconditioner workload) at 6,000 points / 600
    orientations (N = 7,783) through compute_model in the learned form: the
    network-preconditioned matrix-free GMRES attempt, then the direct solve,
    against the oracle's FP64 LU on sampled cells of the grid."""
    say("config4 scaled (N=7,783) learned form vs oracle")
    from gempy_amd import GemPyEngineConfig, compute_model, synthetic
    from gempy_amd.data import SolverForm
    m = synthetic.config4(n_points=6_000, n_ori=600, resolution=(32, 32, 32))
    m.options.solver = SolverForm.GMRES_LEARNED
    sol = compute_model(m, GemPyEngineConfig())
    info = sol.stacks[0].solver_info
    X = m.grid.values()
    idx = np.random.default_rng(4).choice(len(X), 5000, replace=False)
    r = ok.compute_stack(m.stacks[0].as_oracle_dict(), m.options.as_dict(), X[idx])
    S = sol.stacks[0].scalar_field.cpu().numpy()[idx]
    err = np.abs(S - r["scalar"]).max() / np.ptp(r["scalar"])
    band = np.min(np.abs(r["scalar"][:, None] - r["thresholds"][None, :]), 1) <= 1e-6 * np.ptp(r["scalar"])
    mism = int(((sol.lith_block.cpu().numpy()[idx] != r["ids"]) & ~band).sum())
    say(f"config4 scaled: N={len(r['b'])} max|dS|/range {err:.3e} id mismatches {mism} relres {info['relres']:.2e} "
        f"form {info['form']} fallback {info['fallback']} network attempt {info['learned']['iterations']} its "
        f"relres {info['learned']['relres']:.3e}")
    assert err <= 1e-6 and mism == 0 and info["relres"] <= 1e-6
