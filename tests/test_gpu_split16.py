"""The split-fp16 rank-256 and paired rank-512 updates of the FP32
factorization (gk_test_syrk_split16: three fp16 MFMA products per product
over hi / lo parts) against the FP64 product: its error stays within a small multiple of
an FP32 GEMM's on the same data for unit-scale entries (factor rows of an
equilibrated block), within an absolute 1e-12 for entries below fp16's
normal range, on full and lower-triangular launches."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _run(A, B, C0, lower):
    """C0 - A B^T; A, B of 256 columns (one panel pair) or 512 (two pairs,
    one rank-512 pass)."""
    from gempy_amd import _lib
    from gempy_amd.engine import _ptr, _stream, handle
    lib = _lib.load()
    two = A.shape[1] == 512
    dev = lambda X: torch.tensor(np.ascontiguousarray(X), dtype=torch.float32, device="cuda")  # noqa: E731
    dA, dB = dev(A[:, :256]), dev(B[:, :256])
    dA2, dB2 = (dev(A[:, 256:]), dev(B[:, 256:])) if two else (None, None)
    dC = torch.tensor(C0, dtype=torch.float32, device="cuda")
    m, n = A.shape[0], B.shape[0]
    _lib.check("gk_test_syrk_split16", lib.gk_test_syrk_split16(
        handle(0), m, n, _ptr(dA), _ptr(dB), _ptr(dA2) if two else None, _ptr(dB2) if two else None, _ptr(dC), n,
        int(lower), _stream()))
    return dC.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("m,n,scale,lower,k", [(512, 384, 1.0, False, 256), (640, 640, 1.0, True, 256),
                                               (384, 256, 1e-4, False, 256), (256, 256, 3e-7, True, 256),
                                               (512, 256, 1.0, False, 512), (768, 768, 1.0, True, 512)])
def test_split16_update_matches_fp64(gpu, m, n, scale, lower, k):
    say(f"split16 update: m {m} n {n} k {k} scale {scale} lower {lower}")
    rng = np.random.default_rng(m + n + k)
    A = (rng.uniform(-1, 1, (m, k)) * scale).astype(np.float32)
    B = A[:n].copy() if lower else (rng.uniform(-1, 1, (n, k)) * scale).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, n)).astype(np.float32) * np.float32(k * scale * scale)
    got = _run(A, B, C0, lower)
    ref = C0.astype(np.float64) - A.astype(np.float64) @ B.astype(np.float64).T
    f32 = (C0 - A @ B.T).astype(np.float64)
    mask = np.tril(np.ones((m, n), dtype=bool)) if lower else np.ones((m, n), dtype=bool)
    if lower:     # tiles above the diagonal are not touched
        blk = np.add.outer(np.arange(m) // 128, -(np.arange(n) // 128)) < 0
        assert np.array_equal(got[blk], C0.astype(np.float64)[blk])
    err = np.abs(got - ref)[mask].max()
    e32 = np.abs(f32 - ref)[mask].max()
    say(f"split16 update: max err {err:.3e}, FP32 GEMM {e32:.3e}, |ref| max {np.abs(ref).max():.3e}")
    # Below fp16's normal range (|x| < 6e-5) the split keeps absolute, not
    # relative, accuracy: the dropped lo*lo term then reaches ~1e-13 per sum,
    # far below the FP32 resolution (6e-8) of the unit-diagonal blocks the
    # factorization updates.
    assert err <= 4 * e32 + 1e-6 * np.abs(ref).max() + 1e-12


@pytest.mark.parametrize("k", [256, 512])
def test_split16_persistent_lower_update_twice(gpu, k):
    """A lower update with more tiles (666 at m = 4608) than the persistent
    grid holds: workgroups draw tiles from per-XCD counters, which the last
    draw resets.  Two launches in a row give the same bits, and every tile on
    or below the diagonal is updated (the counters were back at 0)."""
    m = 4608
    say(f"split16 persistent lower update: m {m}, k {k}, twice")
    rng = np.random.default_rng(11)
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (m, m)).astype(np.float32) * np.float32(k)
    g1 = _run(A, A, C0, True)
    g2 = _run(A, A, C0, True)
    assert np.array_equal(g1, g2)
    ref = C0.astype(np.float64) - A.astype(np.float64) @ A.astype(np.float64).T
    f32 = (C0 - A @ A.T).astype(np.float64)
    blk = np.add.outer(np.arange(m) // 128, -(np.arange(m) // 128)) < 0
    assert np.array_equal(g1[blk], C0.astype(np.float64)[blk])
    err = np.abs(g1 - ref)[~blk].max()
    e32 = np.abs(f32 - ref)[~blk].max()
    say(f"split16 persistent: max err {err:.3e}, FP32 GEMM {e32:.3e}")
    assert err <= 4 * e32 + 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5, 8])
def test_split16_factorization_pairs(gpu, nb):
    """The FP32 form's factorization (panels in pairs, rank-512 updates, blocked
    panels, the inverse's lower-left quadrants formed at the end) at an odd and
    even number of 256-panels, padded: L L^T against the unit-diagonal SPD
    matrix, every diagonal block's stored inverse against L_kk, and the
    triangular solves through the chain planes against numpy."""
    from gempy_amd import _lib
    from gempy_amd.engine import _ptr, _stream, handle
    lib = _lib.load()
    lda = 256 * nb
    n = lda - 37
    say(f"split16 factorization: nb {nb}, n {n}")
    rng = np.random.default_rng(100 + nb)
    M = rng.standard_normal((n, n))
    A = M @ M.T / n + 0.5 * np.eye(n)
    d = np.sqrt(np.diag(A))
    A = A / d[:, None] / d[None, :]
    tA = torch.zeros((lda, lda), dtype=torch.float32, device="cuda")
    tA[:n, :n] = torch.tensor(A, dtype=torch.float32)
    dinv = torch.empty((_lib.GK_CHOL_PLANES, lda, 256), dtype=torch.float32, device="cuda")
    info = C.c_int32(-1)
    _lib.check("gk_test_cholesky_f32_split16", lib.gk_test_cholesky_f32_split16(
        handle(0), _ptr(tA), n, lda, _ptr(dinv), C.byref(info), _stream()))
    assert info.value == 0
    Lf = torch.tril(tA).double().cpu().numpy()
    err = np.abs(Lf[:n, :n] @ Lf[:n, :n].T - A).max()
    say(f"  max |L L^T - A| {err:.3e}")
    assert err <= 5e-5
    D = dinv[0].double().cpu().numpy()
    for k in range(nb):
        blk = slice(256 * k, 256 * (k + 1))
        e = np.abs(D[blk] @ Lf[blk, blk] - np.eye(256)).max()
        assert e <= 1e-4, (k, e)
    b = rng.standard_normal(n)
    tb = torch.zeros(lda, dtype=torch.float32, device="cuda")
    tb[:n] = torch.tensor(b, dtype=torch.float32)
    _lib.check("gk_cholesky_solve_f32", lib.gk_cholesky_solve_f32(handle(0), _ptr(tA), _ptr(dinv), lda, _ptr(tb),
                                                                  _stream()))
    x = np.linalg.solve(A, b)
    np.testing.assert_allclose(tb[:n].double().cpu().numpy(), x, rtol=0, atol=1e-4 * np.abs(x).max())
