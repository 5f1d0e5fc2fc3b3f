"""Gravity forward restatement -- TEST INFRASTRUCTURE ONLY (only tests/ use it).

* ``prism_tz``: vertical attraction of homogeneous rectangular prisms (unit
  density) at the origin, from the closed form over the 8 corners
  (Nagy 1966; Plouff 1976): sum over corners of
  -mu (x ln(y + r) + y ln(x + r) - z atan(x y / (z r))) with the corner signs
  mu = (1, -1, -1, 1, -1, 1, 1, -1) in the corner order (x1..x2) x (y1..y2) x
  (z1..z2), z fastest; G = 6.674e-3 (mGal with lengths in m and densities in
  g/cm^3).  Pinned by cubature of G (-z) / r^3 over the prism and by the
  Bouguer-plate limit 2 pi G t.
* ``centered_kernel``: the centred grid around a device that the reference's
  engine builds (``CenteredGrid(centers, resolution, radius)``, set at
  gempy/API/grid_API.py:99-108; the engine's source is not in the reference
  tree): restated from GemPy v2's irregular-grid kernel -- x and y cells at
  +-geomspace(0.01, 1, res/2) * radius and 0, z cells at -(concat(0,
  geomspace(0.01, 1, res)) + 0.005) * radius, cell half-sizes from the
  odd-reflected spacings.  Parity unpinned: for the reference's 2-layers test
  (test_gravity.py:67-89) this kernel gives -1520.89 where the record holds
  -1624.1714.
"""
from __future__ import annotations

import numpy as np

G_UGAL = 6.674e-3
MU = np.array([1, -1, -1, 1, -1, 1, 1, -1], dtype=np.float64)


def prism_tz(x1, x2, y1, y2, z1, z2, G=G_UGAL):
    x = np.stack([np.atleast_1d(x1), np.atleast_1d(x2)], 1)
    y = np.stack([np.atleast_1d(y1), np.atleast_1d(y2)], 1)
    z = np.stack([np.atleast_1d(z1), np.atleast_1d(z2)], 1)
    X = np.repeat(x, 4, axis=1)
    Y = np.tile(np.repeat(y, 2, axis=1), (1, 2))
    Z = np.tile(z, (1, 4))
    R = np.sqrt(X * X + Y * Y + Z * Z)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = X * np.log(Y + R) + Y * np.log(X + R) - Z * np.arctan(X * Y / (Z * R))
    return G * np.sum(-MU * t, axis=1)


def centered_kernel(resolution, radius):
    """Cell centres (n, 3) relative to the device, and the half-sizes towards
    lower and higher coordinates (n, 3) each (signed as the spacings are)."""
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (3,))
    g2, d = [], []
    for k in range(3):
        if k == 2:
            g = np.geomspace(0.01, 1, int(resolution[k]))
            gk = (np.concatenate(([0.0], g)) + 0.005) * -radius[k]
        else:
            g = np.geomspace(0.01, 1, int(resolution[k] / 2))
            gk = np.concatenate((-g[::-1], [0.0], g)) * radius[k]
        g2.append(gk)
        d.append(np.diff(np.pad(gk, 1, "reflect", reflect_type="odd")))
    flat = lambda M: np.stack([m.ravel() for m in M], 1)   # noqa: E731
    centres = flat(np.meshgrid(*g2))
    lo = flat(np.meshgrid(d[0][:-1] / 2, d[1][:-1] / 2, d[2][:-1] / 2))
    hi = flat(np.meshgrid(d[0][1:] / 2, d[1][1:] / 2, d[2][1:] / 2))
    return centres, lo, hi


def kernel_tz(centres, lo, hi, G=G_UGAL):
    c = np.asarray(centres)
    return prism_tz(c[:, 0] - lo[:, 0], c[:, 0] + hi[:, 0], c[:, 1] - lo[:, 1], c[:, 1] + hi[:, 1],
                    c[:, 2] - lo[:, 2], c[:, 2] + hi[:, 2], G)
