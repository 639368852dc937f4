"""Sphere and cylinder rasterisation, pinned independently of the code that
does it: a cell belongs to a shape when its integer position (i, j, k)
lies inside it.  The cells a shape covers are counted through the field
energy (a dielectric of eps 2 in vacuum, E = x^ everywhere: sum eps|E|^2 =
cells + cells inside), and must equal a brute-force count over the whole
grid (catches a clipped or shifted bounding box, 0- vs 1-based positions)
and converge to the shape's volume as it grows (relative error below
3 / r^1.5, the lattice-point bound of a smooth convex body)."""
import math

import numpy as np
import pytest


def _count(backend, shape, n):
    import fdtd_amd

    spec = {"grid": {"nx": n, "ny": n, "nz": n, "ds": 1e-9, "courant": 0.5, "periodic": [1, 1, 1]},
            "materials": [{"type": "dielectric", "region": shape, "eps": 2.0}], "nsteps": 0}
    if backend == "oracle":
        import oracle

        sim, _, _ = oracle.make_simulation(spec)
    else:
        sim, _, _ = fdtd_amd.make_simulation(spec)
    sim.commit()
    sim.set_field(0, np.ones((n, n, n)))
    we, _ = sim.energy()
    sim.close()
    return we - n ** 3


def _brute(shape, n):
    i, j, k = np.meshgrid(*(np.arange(1, n + 1, dtype=np.float64),) * 3, indexing="ij")
    p = (i, j, k)
    c = shape["center"]
    if shape["kind"] == "sphere":
        inside = sum((p[a] - c[a]) ** 2 for a in range(3)) <= shape["radius"] ** 2
    else:
        ax = shape["axis"]
        inside = (sum((p[a] - c[a]) ** 2 for a in range(3) if a != ax) <= shape["radius"] ** 2) & (
            np.abs(p[ax] - c[ax]) <= shape["half_length"])
    if "lo" in shape:
        for a in range(3):
            inside &= (p[a] >= shape["lo"][a]) & (p[a] <= shape["hi"][a])
    return int(inside.sum())


def _shapes(r):
    from fdtd_amd import cylinder, sphere

    n = int(2 * r + 8)
    c = (n + 1) / 2.0
    off = (0.31, -0.17, 0.43)  # off the lattice symmetry
    ctr = tuple(c + o for o in off)
    return n, [("sphere", sphere(ctr, r), 4.0 / 3.0 * math.pi * r ** 3),
               ("cylinder z", cylinder(ctr, r, 0.8 * r, 2), math.pi * r * r * 1.6 * r),
               ("cylinder x", cylinder(ctr, 0.7 * r, r, 0), math.pi * (0.7 * r) ** 2 * 2 * r)]


@pytest.mark.parametrize("r", [4.0, 9.5, 20.0, 33.0])
def test_oracle_rasteriser_counts_and_converges(r):
    n, shapes = _shapes(r)
    for name, shape, vol in shapes:
        got = _count("oracle", shape, n)
        assert got == _brute(shape, n), name
        assert abs(got - vol) / vol < 3.0 / r ** 1.5, (name, got, vol)


def test_oracle_rasteriser_clips_to_the_box_and_the_grid():
    """A sphere clipped by its lo/hi box, and one running over the grid's
    edge, cover exactly the brute-force cells."""
    from fdtd_amd import sphere

    n = 30
    a = sphere((15.2, 14.7, 16.1), 9.0, clip_lo=(1, 1, 12), clip_hi=(30, 15, 30))
    b = sphere((3.4, 26.8, 15.5), 8.0)
    for s in (a, b):
        assert _count("oracle", s, n) == _brute(s, n)


@pytest.mark.gpu
@pytest.mark.parametrize("r", [4.0, 20.0])
def test_library_rasteriser_counts_like_the_brute_force(r):
    """The HIP library's host-side rasteriser: the same counts."""
    from fdtd_amd import sphere

    n, shapes = _shapes(r)
    shapes.append(("clipped", sphere((n / 2 + 0.2, n / 2 - 0.3, n / 2), r, clip_lo=(1, 1, n // 2),
                                     clip_hi=(n, n // 2, n)), None))
    for name, shape, _ in shapes:
        assert _count("gpu", shape, n) == _brute(shape, n), name
