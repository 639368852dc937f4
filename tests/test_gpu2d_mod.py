"""The 2D GPU path against the reference's own 2D modules (golden vectors
from tests/golden/gen_golden2d_mod.py): graphene CP4 + Raman with gold, and
the si, film gold and graphene3 1 nm Lorentz-Jp modules, and the 2D y-slab
PML module (mod2/BERpmly_HzExEy_mod.f90) around gold between two
dielectrics.  FP64, fields and traces within 1e-12 of their maxima."""
import json
import os

import numpy as np
import pytest

from conftest import check, ROOT, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["ref2d_graphene_gold", "ref2d_ljp_si", "ref2d_ljp_film_gold",
                                  "ref2d_ljp_graphene3_1nm", "ref2d_pml"])
def test_gpu_matches_the_reference_2d_modules(name):
    from fdtd_amd import sim2d

    g = np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))
    spec = json.loads(bytes(g["spec"]).decode())
    sim, _, pids = sim2d.make_simulation2d(spec, precision="fp64")
    sim.step(spec["nsteps"])
    for comp, key in ((sim2d.EX2, "Ex"), (sim2d.EY2, "Ey"), (sim2d.HZ2, "Hz")):
        check(f"test_gpu2d_mod:26 {(name, key)}", relerr(sim.field(comp), g[key]), 1e-12)
    for p, pid in enumerate(pids):
        check(f"test_gpu2d_mod:28 {(name, p)}", relerr(sim.probe(pid), g["probes"][p]), 1e-12)
    sim.close()
