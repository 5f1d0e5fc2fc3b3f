"""FP64 torch restatement of the learned preconditioner's network and local
SPAI loss -- TEST INFRASTRUCTURE ONLY (checks the HIP forward/backward and
loss kernels of gempy_amd/csrc/gk_learned.hip).

Network: features (8) -> tanh(W1 x + b1) (32) -> tanh(W2 h + b2) (32) ->
W3 h + b3 (9), plus the identity on self pairs (feature 6).  Loss of point P:
|| M_P H_loc - E_P ||_F^2 over the valid rows of P (3 for an orientation, 1 for
an interface point), M_P the (3 x 3K) row block assembled from the K outputs.
"""
from __future__ import annotations

import torch

F, H, O = 8, 32, 9


def unpack(p):
    i = 0
    out = []
    for shp in [(H, F), (H,), (H, H), (H,), (O, H), (O,)]:
        n = 1
        for s in shp:
            n *= s
        out.append(p[i:i + n].reshape(shp))
        i += n
    return out


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = torch.tanh(x @ W1.T + b1)
    h2 = torch.tanh(h1 @ W2.T + b2)
    y = h2 @ W3.T + b3
    eye = torch.zeros(O, dtype=x.dtype)
    eye[[0, 4, 8]] = 1.0
    return y + (x[:, 6:7] > 0.5).to(x.dtype) * eye


def spai_loss(y, Hloc, rows, self_pos, K):
    """y (nq*K, 9), Hloc (nq, 3K, 3K), rows (nq,) valid row count, self_pos (nq,)."""
    nq = Hloc.shape[0]
    Yb = y.reshape(nq, K, 3, 3)                      # [b, j, r, c]
    M = Yb.permute(0, 2, 1, 3).reshape(nq, 3, 3 * K)  # [b, r, 3j + c]
    R = M @ Hloc
    E = torch.zeros_like(R)
    for b in range(nq):
        for r in range(3):
            E[b, r, 3 * int(self_pos[b]) + r] = 1.0
    R = R - E
    mask = (torch.arange(3)[None, :] < rows[:, None]).to(R.dtype)[:, :, None]
    return ((R * mask) ** 2).sum(dim=(1, 2))
