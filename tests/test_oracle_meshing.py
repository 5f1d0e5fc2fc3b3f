"""Pins of the meshing oracle (CPU): the crossing-edge vertex set and the
mesh measures against closed forms."""
import numpy as np

from oracle import meshing as om


def test_crossing_vertices_of_a_plane():
    """Field x + 2y + 3z on a lattice: every vertex lies on the plane, and the
    count equals the number of lattice edges straddling it (brute force)."""
    n = (9, 8, 7)
    X, Y, Z = np.meshgrid(*[np.arange(k, dtype=float) for k in n], indexing="ij")
    F = X + 2 * Y + 3 * Z
    level = 10.3
    V = om.crossing_vertices(F, level, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    np.testing.assert_allclose(V @ [1, 2, 3], level, rtol=0, atol=1e-12)
    count = 0
    for x in range(n[0]):
        for y in range(n[1]):
            for z in range(n[2]):
                for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    q = (x + d[0], y + d[1], z + d[2])
                    if all(q[i] < n[i] for i in range(3)):
                        count += (F[x, y, z] > level) != (F[q] > level)
    assert len(V) == count
    # a mask removes exactly the edges of the masked cubes: a cube counts when
    # its far corner (+1, +1, +1) is set, so mask[:4] = False keeps cubes x >= 3
    mask = np.ones(n, dtype=bool)
    mask[:4] = False
    Vm = om.crossing_vertices(F, level, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), mask)
    assert 0 < len(Vm) < len(V) and np.all(Vm[:, 0] >= 3 - 1e-12) and np.any(Vm[:, 0] < 4)
    keep = V[V[:, 0] >= 3 - 1e-12]
    assert len(Vm) == len(keep)


def test_mesh_checks_of_a_unit_cube():
    V = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], dtype=float)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    F = np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))])
    r = om.mesh_checks(V, F)
    assert r["unpaired"] == 0 and r["euler"] == 2
    assert abs(r["area"] - 6.0) < 1e-12 and abs(abs(r["volume"]) - 1.0) < 1e-12
    r2 = om.mesh_checks(V, F[:-1])
    assert r2["unpaired"] > 0
