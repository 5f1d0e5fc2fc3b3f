"""The grid search for the FP32 form's local-inverse neighbourhoods
(gk_test_knn_local) against brute force in numpy: for every data point the
distances of its 32 returned neighbours (itself first) equal the 32 smallest
distances, on clustered sheet-like data (config-3 surfaces) and on a
multi-stack model's small stack."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _check(st, ko, center):
    from gempy_amd import _lib
    from gempy_amd.engine import DeviceStack, _ptr, _stream
    lib = _lib.load()
    K = _lib.GK_LOCAL_NEIGHBOURS
    ds = DeviceStack(st, ko, center)
    n = ds.sys.n_ori + ds.sys.n_rest
    idx = torch.empty(n * K, dtype=torch.int32, device="cuda")
    hk = torch.empty(n, dtype=torch.float32, device="cuda")
    _lib.check("gk_test_knn_local", lib.gk_test_knn_local(ds.h, C.byref(ds.sys), _ptr(idx), _ptr(hk), _stream()))
    idx = idx.cpu().numpy().reshape(n, K)
    # data points: orientations, then rest points (every surface's points after its first)
    rest = np.concatenate([st.sp_xyz[o + 1:o + c] for o, c in zip(np.cumsum([0] + st.counts[:-1]), st.counts)])
    X = (np.concatenate([st.ori_xyz, rest]) - np.asarray(center)) / ko.range
    assert idx.min() >= 0 and idx.max() < n
    D = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    exact = np.sort(D, axis=1)[:, :K]
    got = np.take_along_axis(D, idx.astype(np.int64), axis=1)
    assert np.all(idx[:, 0] == np.arange(n)) or np.all(got[:, 0] == 0)
    err = np.abs(np.sort(got, axis=1) - exact).max()
    say(f"neighbours: n {n}, K {K}, max |d_got - d_exact| {err:.2e}, max d {exact[:, -1].max():.3e}")
    assert err <= 1e-5 * exact[:, -1].max()
    for row in idx[:50]:
        assert len(set(row.tolist())) == K


def test_neighbours_on_surfaces(gpu):
    from gempy_amd import synthetic
    say("grid neighbours: config3 scaled (1500 points, 400 orientations)")
    m = synthetic.config3(n_points=1500, n_ori=400, resolution=(8, 8, 8))
    _check(m.stacks[0], m.options.kernel_options, [500.0, 500.0, 500.0])


def test_neighbours_small_stack(gpu):
    from gempy_amd import synthetic
    say("grid neighbours: config2 scale 0.1 stacks")
    m = synthetic.config2(scale=0.1, resolution=(8, 8, 8), octree_levels=1)
    for st in m.stacks:
        if st.n_orientations + sum(st.counts) - len(st.counts) >= 32:
            _check(st, m.options.kernel_options, [500.0, 500.0, 500.0])
