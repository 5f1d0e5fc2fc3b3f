"""gk_eval_noalloc (no device allocation; file-scope scratch array) against
gk_eval through the C ABI: two calls in a row, the second with a smaller
input, explicit points (split partials in the scratch) and the lattice, with
gradient and ids.  Results must be identical (same kernels, same split)."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _desc(L, kind, n, xyz, S, G, ids, thr, grid=None):
    d = L.GkEvalDesc()
    d.kind = kind
    d.precision = L.GK_F64
    d.n_points = n
    d.xyz = C.c_void_p(xyz.data_ptr()) if xyz is not None else None
    if grid is not None:
        for k in range(3):
            d.origin[k], d.spacing[k], d.res[k] = grid.origin[k], grid.spacing[k], grid.resolution[k]
    d.scalar = C.c_void_p(S.data_ptr())
    d.gradient = C.c_void_p(G.data_ptr())
    d.ids = C.c_void_p(ids.data_ptr())
    d.n_thresholds = len(thr)
    for k, v in enumerate(thr):
        d.thresholds[k] = float(v)
    return d


def _run(fn_name, ds, w, kind, n, xyz, thr, grid=None):
    from gempy_amd import _lib as L
    from gempy_amd.engine import _stream
    S = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    G = torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda")
    ids = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    d = _desc(L, kind, n, xyz, S, G, ids, thr, grid)
    fn = getattr(L.load(), fn_name)
    L.check(fn_name, fn(ds.h, C.byref(ds.sys), C.c_void_p(w.data_ptr()), C.byref(d), _stream()))
    torch.cuda.synchronize()
    return S.cpu().numpy(), G.cpu().numpy(), ids.cpu().numpy()


@pytest.mark.parametrize("kind", ["explicit", "lattice"])
def test_eval_noalloc_matches_eval_twice(gpu, kind):
    from gempy_amd import _lib as L
    from gempy_amd import synthetic
    from gempy_amd.data import RegularGrid
    from gempy_amd.engine import DeviceStack
    for call, (n_pts, n_ori, n_eval) in enumerate([(3000, 800, 20_000), (600, 150, 3_000)]):
        say(f"eval_noalloc {kind} call {call + 1}: {n_pts} points, {n_ori} orientations, {n_eval} evaluation points")
        m = synthetic.config3(n_points=n_pts, n_ori=n_ori, resolution=(8, 8, 8))
        ds = DeviceStack(m.stacks[0], m.options.kernel_options, [500.0, 500.0, 500.0])
        w, _ = ds.solve(L.GK_SOLVE_FP64_CHOL)
        thr, _ = ds.surface_thresholds(w)
        if kind == "explicit":
            X = torch.tensor(np.random.default_rng(call).uniform(0, 1000, (n_eval, 3)), device="cuda")
            args = (L.GK_POINTS_EXPLICIT, n_eval, X, thr)
            grid = None
        else:
            r = int(round(n_eval ** (1 / 3)))
            grid = RegularGrid((0, 1000, 0, 1000, 0, 1000), (r, r + 1, r + 2))
            args = (L.GK_POINTS_LATTICE, grid.n_points, None, thr)
        a = _run("gk_eval", ds, w, *args, grid=grid)
        b = _run("gk_eval_noalloc", ds, w, *args, grid=grid)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), "gk_eval_noalloc differs from gk_eval"
        assert np.isfinite(b[0]).all() and (b[2] >= 1).all()
        say(f"eval_noalloc {kind} call {call + 1}: identical ({len(b[0])} points)")


def test_eval_noalloc_rejects_oversized_system(gpu):
    """A system whose packed sources exceed the scratch array is refused with
    a clear error (checked on the host before any launch)."""
    from gempy_amd import _lib as L
    from gempy_amd.engine import _stream, handle
    n_ori = L.GK_EVAL_SCRATCH_BYTES // 64 + 1         # orientations alone overflow it
    s = L.GkSystem()
    s.n_ori, s.n_rest, s.n_surf, s.uni_degree, s.n_fault = n_ori, 0, 0, 1, 0
    s.range, s.c_o, s.i_res, s.gi_res = 1.0, 1.0, 4.0, 2.0
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = C.c_void_p(dummy.data_ptr())
    s.ori_xyz = s.ori_grad = s.ori_nugget = p
    S = torch.zeros(4, dtype=torch.float64, device="cuda")
    d = L.GkEvalDesc()
    d.kind, d.precision, d.n_points = L.GK_POINTS_EXPLICIT, L.GK_F64, 1
    d.xyz, d.scalar = p, C.c_void_p(S.data_ptr())
    rc = L.load().gk_eval_noalloc(handle(0), C.byref(s), p, C.byref(d), _stream())
    assert rc == -1
    assert b"scratch" in L.load().gk_last_error()
