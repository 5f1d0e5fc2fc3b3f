"""Write the oracle's FP64 fixtures of the full-size workloads (run in the build
container; ~10 min and ~25 GB of host memory for config 3):

* config3_oracle.npz -- the precision-form model (one folded stack, 20,000
  interface points on 10 surfaces, 5,000 orientations, N = 34,993): the FP64
  weights of the dense LU solve (numpy/LAPACK gesv, the reference's
  kernel_solver class), the surface thresholds, and S and ids at a seeded
  sample of 65,536 cells of the 200^3 lattice (indices in ij ravel order,
  z fastest).
* config2_oracle.npz -- the multi-stack model at full scale (a fault stack and
  three layer stacks, 5,000 points / 1,000 orientations, per-stack N ~ 2.6k):
  per-stack weights and thresholds, S of every stack, the lith block and the
  fault block at a seeded sample of 65,536 cells of the 200^3 lattice.

The oracle restates the method (oracle/cokriging.py); these files are its
outputs at sizes the GPU tests cannot recompute on the CPU in seconds.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

N_SAMPLE = 65_536


def lattice_sample(grid, n, seed):
    idx = np.sort(np.random.default_rng(seed).choice(grid.n_points, n, replace=False)).astype(np.int64)
    return idx, grid.values()[idx]


def config3():
    from gempy_amd import synthetic
    from oracle import cokriging as ok
    m = synthetic.config3()
    st, o = m.stacks[0].as_oracle_dict(), m.options.as_dict()
    t0 = time.time()
    A, b = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], o["range"], o["c_o"], 1, o["i_res"], o["gi_res"])
    print(f"config3: assembled N={len(b)} in {time.time() - t0:.1f} s", flush=True)
    # residual of the LU solution needs A afterwards: keep the matvec pieces by
    # solving a copy-free system and re-assembling for the residual
    t0 = time.time()
    w = ok.solve_system(A, b, overwrite=True)
    print(f"config3: LU solve in {time.time() - t0:.1f} s", flush=True)
    del A
    A, _ = ok.build_system(st["sp_xyz"], st["counts"], st["sp_nugget"], st["ori_xyz"], st["ori_grad"],
                           st["ori_nugget"], o["range"], o["c_o"], 1, o["i_res"], o["gi_res"])
    relres = float(np.linalg.norm(b - A @ w) / np.linalg.norm(b))
    del A
    print(f"config3: relres {relres:.3e}", flush=True)
    kw = dict(a=o["range"], c_o=o["c_o"], uni_degree=1, i_res=o["i_res"], gi_res=o["gi_res"])
    t0 = time.time()
    s_sp = ok.evaluate(st["sp_xyz"], st["sp_xyz"], st["counts"], st["ori_xyz"], w, **kw)
    thr = ok.surface_thresholds(st["sp_xyz"], st["counts"], s_sp)
    idx, X = lattice_sample(m.grid, N_SAMPLE, 303)
    S = ok.evaluate(X, st["sp_xyz"], st["counts"], st["ori_xyz"], w, **kw)
    ids = ok.segment(S, thr)
    print(f"config3: evaluated {len(X)} lattice cells in {time.time() - t0:.1f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "config3_oracle.npz"), weights=w, thresholds=thr,
                        sample_index=idx, scalar=S, ids=ids.astype(np.int8), relres=relres)


def config2():
    from gempy_amd import synthetic
    from oracle import cokriging as ok
    m = synthetic.config2()
    idx, X = lattice_sample(m.grid, N_SAMPLE, 202)
    t0 = time.time()
    res = ok.compute_model([s.as_oracle_dict() for s in m.stacks], m.options.as_dict(), X)
    print(f"config2: {len(m.stacks)} stacks, N {[len(r['b']) for r in res['stacks']]} "
          f"in {time.time() - t0:.1f} s", flush=True)
    out = dict(sample_index=idx, lith_block=res["lith_block"].astype(np.int8),
               fault_block=res["fault_blocks"][0].astype(np.int8))
    for k, r in enumerate(res["stacks"]):
        out[f"weights_{k}"] = r["weights"]
        out[f"thresholds_{k}"] = r["thresholds"]
        out[f"scalar_{k}"] = r["scalar"]
    np.savez_compressed(os.path.join(HERE, "config2_oracle.npz"), **out)


if __name__ == "__main__":
    which = sys.argv[1:] or ["config2", "config3"]
    for name in which:
        globals()[name]()
    print("fixtures written to", HERE)
