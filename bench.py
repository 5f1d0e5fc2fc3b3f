#!/usr/bin/env python
"""Benchmark of the implicit co-kriging hot path on one MI355X.

Workload (default): the precision-form model -- one folded stack of 20,000
interface points on 10 surfaces and 5,000 orientations (system N = 34,993),
evaluated on a 200^3 regular grid (8,000,000 points).  One step = assemble +
solve + surface thresholds + grid evaluation (scalar field + ids) with the
inputs already resident in HBM.  Three solve forms are timed on the same GPU:
the FP32 form (shifted Cholesky with split-fp16 MFMA updates, exact local
inverses, FP64 FGMRES refinement: the headline form), the FP64 Cholesky + Schur
form and the plain FP64 blocked LU (the FP64 direct form); the grid is
evaluated in FP64 in all
(FP32 pair sums cannot reach 1e-6 on this workload: the weighted terms
cancel over six decades).

Prints ONE JSON line.  `value` = grid points modelled per second, end to end,
in the FP32+refinement form.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "model s & grid Gpts/s at N=35k: FP32+refine/learned vs FP64 direct, err≤1e-6"
PEAK = {"fp64_valu": 78.6, "fp32_valu": 157.3, "fp32_mfma": 157.3, "fp64_mfma": 78.6, "fp16_mfma": 2500.0,
        "hbm": 8000.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="scale of point/orientation counts")
    ap.add_argument("--res", type=int, default=200, help="grid resolution per axis")
    ap.add_argument("--eval-precision", choices=["f64", "f32"], default="f64")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 direct form")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the multi-stack/octree and adjoint workloads")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (3 times), the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every box in this setup has one GPU and the co-kriging "
              "path is not sharded here.", file=sys.stderr)
        sys.exit(2)
    import ctypes as C
    import numpy as np
    import torch
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")      # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read_per_launch": 1 << 30,
                          "bytes_written_per_launch": 1 << 30, "launches": 3}), flush=True)
        return
    from gempy_amd import _lib, synthetic
    from gempy_amd.engine import DeviceStack
    lib = _lib.load()

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    n_pts, n_ori = int(round(20000 * args.scale)), int(round(5000 * args.scale))
    model = synthetic.config3(n_points=n_pts, n_ori=n_ori, resolution=(args.res,) * 3)
    st, grid, k = model.stacks[0], model.grid, model.options.kernel_options
    center = [(grid.extent[2 * d] + grid.extent[2 * d + 1]) / 2 for d in range(3)]
    ds = DeviceStack(st, k, center)
    N, P = ds.N, grid.n_points
    R, S, no = ds.sys.n_rest, ds.sys.n_surf, ds.sys.n_ori
    Nc = 3 * no + R
    eprec = _lib.GK_F64 if args.eval_precision == "f64" else _lib.GK_F32
    log(f"config3 N={N} P={P} R={R} S={S} n_ori={no} eval={args.eval_precision}")

    def step(form, record=None):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        w, info = ds.solve(form)
        ev[1].record()
        thr, _ = ds.surface_thresholds(w)
        ev[2].record()
        Sg, _, ids = ds.evaluate_lattice(grid, w, eprec, False, thr)
        ev[3].record()
        if record is not None:
            torch.cuda.synchronize()
            record.append(dict(solve=ev[0].elapsed_time(ev[1]), thr=ev[1].elapsed_time(ev[2]),
                               eval=ev[2].elapsed_time(ev[3]), info=info))
        return w, thr, Sg, ids

    def run(form, name, profile=False):
        for i in range(args.warmup):
            log(f"{name}: warmup {i + 1}/{args.warmup}")
            step(form)
        torch.cuda.synchronize()
        rec = []
        if profile:
            _lib.check("gk_profile", lib.gk_profile(ds.h, 1))
        t0 = time.perf_counter()
        for i in range(args.steps):
            out = step(form, rec)
            log(f"{name}: step {i + 1}/{args.steps} solve {rec[-1]['solve']:.1f} ms eval {rec[-1]['eval']:.1f} ms "
                f"info {rec[-1]['info']}")
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        if profile:
            prof = {}
            for kind in range(12):
                n, ms, work = C.c_int64(0), C.c_double(0), C.c_double(0)
                _lib.check("gk_profile_read", lib.gk_profile_read(ds.h, kind, C.byref(n), C.byref(ms), C.byref(work)))
                if n.value:
                    prof[kind] = (n.value, ms.value, work.value)
            _lib.check("gk_profile", lib.gk_profile(ds.h, 0))
            rec.append({"profile": prof})
        return t, rec, out

    t_fast, rec_fast, out_fast = run(_lib.GK_SOLVE_FP32_REFINE, "fp32_refine", profile=True)
    prof_fast = rec_fast.pop()["profile"]
    res = {}
    if not args.no_fp64:
        t_chol, rec_chol, out_chol = run(_lib.GK_SOLVE_FP64_CHOL, "fp64_chol")
        res.update(model_s_fp64_chol=t_chol, fp64_chol_factor_ms=float(np.mean([r["info"]["t_factor_ms"] for r in rec_chol])),
                   fp64_chol_solve_ms=float(np.mean([r["solve"] for r in rec_chol])))
        t_f64, rec_f64, out_f64 = run(_lib.GK_SOLVE_FP64_LU, "fp64_lu")
        s_fast, s_ref = out_fast[2].cpu().numpy(), out_f64[2].cpu().numpy()
        rng = float(s_ref.max() - s_ref.min())
        band = np.zeros(P, dtype=bool)
        thr_ref = out_f64[1]
        for t in thr_ref:
            band |= np.abs(s_ref - t) <= 1e-6 * rng
        ids_f, ids_r = out_fast[3].cpu().numpy(), out_f64[3].cpu().numpy()
        res.update(model_s_fp64=t_f64, speedup_vs_fp64=t_f64 / t_fast,
                   err_vs_fp64_max_rel=float(np.abs(s_fast - s_ref).max() / rng),
                   err_vs_fp64_rms_rel=float(np.sqrt(np.mean((s_fast - s_ref) ** 2)) / rng),
                   ids_mismatch_outside_band=int(((ids_f != ids_r) & ~band).sum()),
                   ids_band_voxels=int(band.sum()),
                   fp64_solve_ms=float(np.mean([r["solve"] for r in rec_f64])),
                   fp64_eval_ms=float(np.mean([r["eval"] for r in rec_f64])),
                   fp64_lu_factor_ms=float(np.mean([r["info"]["t_factor_ms"] for r in rec_f64])))

    # error of the FP32 form against the FP64 oracle: the committed fixture holds
    # the oracle's S and ids at 65,536 points of this 200^3 lattice
    gold = os.path.join(ROOT, "tests", "golden", "config3_oracle.npz")
    if args.scale == 1.0 and args.res == 200 and os.path.exists(gold):
        d = np.load(gold)
        gi = torch.as_tensor(d["sample_index"], device=dev)
        So, thr_o = d["scalar"], d["thresholds"]
        rng_o = float(So.max() - So.min())
        s_g = out_fast[2][gi].cpu().numpy()
        ids_g = out_fast[3][gi].cpu().numpy()
        band_o = np.zeros(len(So), dtype=bool)
        for t in thr_o:
            band_o |= np.abs(So - t) <= 1e-6 * rng_o
        res["err_vs_oracle"] = {
            "max_rel": float(np.abs(s_g - So).max() / rng_o),
            "rms_rel": float(np.sqrt(np.mean((s_g - So) ** 2)) / rng_o),
            "thresholds_max_rel": float(np.abs(np.asarray(out_fast[1]) - thr_o).max() / rng_o),
            "ids_mismatch_outside_band": int(((ids_g != d["ids"]) & ~band_o).sum()),
            "points": int(len(So)),
            "fixture": "tests/golden/config3_oracle.npz (FP64 NumPy oracle solve of the same system)"}
    # FP32 assembly alone (gk_assemble): the equilibrated covariance block, all
    # lda x lda entries written (both triangles and the identity padding)
    from gempy_amd.engine import _stream
    A32, _, d32 = ds.assemble(dtype=_lib.GK_F32)
    ta = []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _lib.check("gk_assemble", lib.gk_assemble(ds.h, C.byref(ds.sys), _lib.GK_F32, C.c_void_p(A32.data_ptr()),
                                                  A32.shape[0], None, C.c_void_p(d32.data_ptr()), _stream()))
        e1.record()
        torch.cuda.synchronize()
        ta.append(e0.elapsed_time(e1))
    bytes_a = 4.0 * float(A32.shape[0]) ** 2
    res["assembly"] = {"ms": float(np.median(ta)), "bytes_written": bytes_a,
                       "GB_s": bytes_a / (np.median(ta) * 1e-3) / 1e9, "dtype": "f32 (full padded block)"}
    del A32, d32

    solve_ms = float(np.mean([r["solve"] for r in rec_fast]))
    eval_ms = float(np.mean([r["eval"] for r in rec_fast]))
    factor_ms = float(np.mean([r["info"]["t_factor_ms"] for r in rec_fast]))
    # grid-evaluation kernel: 21 flop per (point, interface/reference source)
    # pair and 24 per (point, orientation) pair, FMA = 2.
    flop_eval = P * (21.0 * (R + S) + 24.0 * no)
    ach_eval = flop_eval / (eval_ms * 1e-3) / 1e12
    peak_eval = PEAK["fp64_valu"] if eprec == _lib.GK_F64 else PEAK["fp32_valu"]
    # per-kernel figures of the timed steps (HIP events around each launch)
    kernels = {}
    for kind, (n, ms, work) in prof_fast.items():
        if kind < 8:
            name = "gemm_kernel<%s, %s, %s>" % ("double" if kind >= 4 else "float", "true" if kind & 2 else "false",
                                                "true" if kind & 1 else "false")
            peak = PEAK["fp64_mfma"] if kind >= 4 else PEAK["fp32_mfma"]
            ach, unit, bound = work / (ms * 1e-3) / 1e12, "TFLOP/s", "mfma"
        elif kind >= 10:
            # split-fp16 rank-256 update: three fp16 MFMA products per product;
            # achieved counts the fp16 flops executed (3x the algorithmic ones)
            name = "syrk_h3_kernel<%s>" % ("true" if kind == 11 else "false")
            peak, ach, unit, bound = PEAK["fp16_mfma"], 3 * work / (ms * 1e-3) / 1e12, "TFLOP/s", "mfma"
        else:
            name = "trsv_fwd_kernel+trsv_bwd_kernel<%s>" % ("double" if kind == 9 else "float")
            peak, ach, unit, bound = PEAK["hbm"], work / (ms * 1e-3) / 1e9, "GB/s", "hbm"
        kernels[name] = {"launches": n, "total_ms": ms, "avg_ms": ms / n, "work_per_launch": work / n,
                         "achieved": ach, "peak": peak, "unit": unit, "bound": bound, "frac": ach / peak}
        if kind >= 10:
            kernels[name]["algorithmic_tflops"] = work / (ms * 1e-3) / 1e12
    kernels["eval_kernel<double, true, false, false, 8>" if eprec == _lib.GK_F64 else "eval_kernel (FP32 lattice)"] = {
        "launches": args.steps, "total_ms": eval_ms * args.steps, "avg_ms": eval_ms, "work_per_launch": flop_eval,
        "achieved": ach_eval, "peak": peak_eval, "unit": "TFLOP/s", "bound": "valu", "frac": ach_eval / peak_eval}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    traffic = _traffic(dom)
    out = {
        "metric": METRIC, "value": P / t_fast / 1e9, "unit": "Gpts/s (model: assemble+solve+evaluate)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_fast * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 factor + f64 refine/eval" if eprec == _lib.GK_F64 else "f32",
        "data": "synthetic (seeded config-3 generator)",
        "config": {"workload": "config3 precision form: 1 stack, 10 surfaces, N=%d, grid %d^3" % (N, args.res),
                   "N": N, "grid_points": P, "interface_points": n_pts, "orientations": n_ori},
        "model_s": t_fast, "grid_gpts_s": P / (eval_ms * 1e-3) / 1e9, "solve_ms": solve_ms,
        "eval_ms": eval_ms, "factor_ms": factor_ms,
        "gmres_iterations": rec_fast[-1]["info"]["iterations"], "relres": rec_fast[-1]["info"]["relres"],
        "fallback": rec_fast[-1]["info"]["fallback"], "shift": rec_fast[-1]["info"]["shift"],
        "roofline": {"bound": d["bound"], "kernel": dom, "achieved": d["achieved"], "peak": d["peak"],
                     "unit": d["unit"], "frac": d["frac"], "traffic": traffic,
                     "launches_per_step": d["launches"] / args.steps, "avg_launch_ms": d["avg_ms"]},
        "kernels": kernels,
    }
    out.update(res)
    if not args.no_extras:
        from gempy_amd import workloads
        log("multi-stack + fault + 4 octree levels, 200^3 (FP32 form)")
        out["multistack_octree"] = workloads.multistack_octree(steps=1, warmup=1)
        log(f"  {out['multistack_octree']}")
        log("adjoint of a 10,000-observation misfit (faulted stack, 128^3 grid)")
        out["adjoint"] = workloads.adjoint_misfit(steps=1, warmup=1)
        log(f"  {out['adjoint']}")
        from gempy_amd import learned
        if os.path.exists(learned.WEIGHTS):
            log("learned form: config-4 system (N~130k: budgeted network GMRES, then the direct solve, 256^3 grid) "
                "+ 35k judge")
            out["learned"] = workloads.learned_gmres(eval_grid=True)
            log(f"  {out['learned']}")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(model, out_fast[0].cpu().numpy(), args.cpu_seconds, N, P)
    print(json.dumps(out), flush=True)


def _traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-kernel
    calibration), or None when no summary for it is committed.  Format: table
    rows '| kernel | ... | traffic_bytes_per_launch |'."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 2 and cells[0] == kernel:
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def cpu_baseline(model, weights, budget_s, N, P):
    """FP64 NumPy oracle on the host cores, on a bounded sample of the same
    workload, extrapolated: solve time of a sub-sampled system scaled by
    (N/Ns)^3 and evaluation time of a sample of grid points scaled by P/Ps."""
    import numpy as np
    from gempy_amd import synthetic
    from oracle import cokriging as ok
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        cores = 1
    k = model.options.kernel_options
    st = model.stacks[0]
    # solve sample
    sub = synthetic.config3(n_points=2400, n_ori=600, resolution=(8, 8, 8)).stacks[0]
    t0 = time.perf_counter()
    A, b = ok.build_system(sub.sp_xyz, sub.counts, sub.sp_nugget, sub.ori_xyz, sub.ori_grad, sub.ori_nugget,
                           k.range, k.c_o)
    ok.solve_system(A, b)
    t_solve_s = time.perf_counter() - t0
    Ns = len(b)
    # evaluation sample with the full system's weights
    X = model.grid.values()[:: max(1, P // 20000)][:20000]
    t0 = time.perf_counter()
    ok.evaluate(X[:2000], st.sp_xyz, st.counts, st.ori_xyz, weights, k.range, k.c_o)
    t_probe = time.perf_counter() - t0
    n_eval = int(min(len(X), max(2000, 2000 * (0.6 * budget_s) / max(t_probe, 1e-3))))
    t0 = time.perf_counter()
    ok.evaluate(X[:n_eval], st.sp_xyz, st.counts, st.ori_xyz, weights, k.range, k.c_o)
    t_eval_s = time.perf_counter() - t0
    t_model = t_solve_s * (N / Ns) ** 3 + t_eval_s * P / n_eval
    return {"value": P / t_model / 1e9, "unit": "Gpts/s (model, extrapolated)", "cores": cores,
            "kind": "port",
            "sample": f"FP64 NumPy oracle: assemble+solve of a N={Ns} system ({t_solve_s:.2f} s, x(N/Ns)^3) "
                      f"+ evaluation of {n_eval} grid points with the full N={N} weights "
                      f"({t_eval_s:.2f} s, xP/{n_eval}); extrapolated model time {t_model:.0f} s; "
                      f"{cores} BLAS threads for the solve, elementwise NumPy single-threaded"}


if __name__ == "__main__":
    main()
