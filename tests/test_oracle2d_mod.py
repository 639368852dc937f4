"""The 2D oracle against the reference's own 2D modules (golden vectors of
tests/golden/gen_golden2d_mod.py: modules/src/2D graphene (CP4 with its
central-difference Raman term) and gold under a line source, and the 2D
Lorentz-Jp modules si, film gold and graphene3 1 nm, and the 2D y-slab PML
module mod2/BERpmly_HzExEy_mod.f90 with probes inside both slabs, each run
by a single-domain driver around the compiled modules and the reference's
2D Yee include).  Final fields and point traces."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT


def golden(name):
    return np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))


@pytest.mark.parametrize("name", ["ref2d_graphene_gold", "ref2d_ljp_si", "ref2d_ljp_film_gold",
                                  "ref2d_ljp_graphene3_1nm", "ref2d_pml"])
def test_oracle_matches_the_reference_2d_modules(name):
    """Fields and traces bit for bit: the oracle follows the modules'
    operations in order, built without multiply-add contraction as they are,
    with the complex quotient's real part in the form their build uses."""
    import oracle
    from fdtd_amd import sim2d

    g = golden(name)
    spec = json.loads(bytes(g["spec"]).decode())
    sim, _, pids = oracle.make_simulation2d(spec)
    sim.step(spec["nsteps"])
    for comp, key in ((sim2d.EX2, "Ex"), (sim2d.EY2, "Ey"), (sim2d.HZ2, "Hz"), (sim2d.DX2, "Dx"), (sim2d.DY2, "Dy")):
        assert np.array_equal(sim.field(comp), g[key]), (name, key)
    for p, pid in enumerate(pids):
        assert np.array_equal(sim.probe(pid), g["probes"][p]), (name, p)
    sim.close()
