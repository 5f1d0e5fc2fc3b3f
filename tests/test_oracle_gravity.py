"""Pins of the gravity restatement (CPU): the prism closed form against
cubature and the Bouguer-plate limit; the centred kernel's geometry."""
import numpy as np

from oracle import gravity as og


def test_prism_tz_matches_cubature():
    G = og.G_UGAL
    for (x1, x2, y1, y2, z1, z2) in [(-3, 5, -2, 4, -9, -6), (1, 2, 1, 3, -4, -1), (-10, 10, -10, 10, -2, -1)]:
        n = 120
        xs = np.linspace(x1, x2, n + 1); ys = np.linspace(y1, y2, n + 1); zs = np.linspace(z1, z2, n + 1)
        c = lambda v: (v[:-1] + v[1:]) / 2          # noqa: E731
        X, Y, Z = np.meshgrid(c(xs), c(ys), c(zs), indexing="ij")
        dV = (x2 - x1) * (y2 - y1) * (z2 - z1) / n ** 3
        num = G * np.sum(-Z / (X * X + Y * Y + Z * Z) ** 1.5) * dV
        ana = og.prism_tz(x1, x2, y1, y2, z1, z2)[0]
        assert abs(ana - num) <= 2e-4 * abs(num), (ana, num)


def test_prism_tz_bouguer_limit():
    """A plate of thickness t below the observer, growing laterally, tends to
    2 pi G t (per unit density)."""
    t = 10.0
    for L, tol in [(1e4, 2e-3), (1e6, 2e-5)]:
        v = og.prism_tz(-L, L, -L, L, -5.0 - t, -5.0)[0]
        assert abs(v / (2 * np.pi * og.G_UGAL * t) - 1) < tol


def test_centered_kernel_geometry():
    c, lo, hi = og.centered_kernel([10, 10, 100], 16000.0)
    assert c.shape == (11 * 11 * 101, 3)
    assert np.all(c[:, 2] < 0)                        # all cells below the device
    # adjacent cells tile the axes: a cell's upper face is the next cell's lower face
    xs = np.unique(c[:, 0])
    assert len(xs) == 11 and np.isclose(xs[5], 0.0)
