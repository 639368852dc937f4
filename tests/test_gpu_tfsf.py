"""GPU TFSF plane wave through the C ABI: against the reference's own 1D
run (golden vectors, the full 54,747 steps) and against the oracle on the
orientation, engine, PML and precision cases.

Tolerances, normwise: FP64 traces and spectra against the reference 1e-12;
FP64 fields against the oracle 1e-12; FP32 fields 2e-4, FP32 traces 1e-3."""
import math

import numpy as np
import pytest

from conftest import check, relerr
from test_oracle_tfsf import ORIENT, tfsf_box

pytestmark = pytest.mark.gpu
PI = 4.0 * math.atan(1.0)


def test_gpu_tfsf_reproduces_the_reference_1d_run(golden_npz):
    import fdtd_amd
    from fdtd_amd import workloads

    z = golden_npz("ref1d_tfsf")
    spec = workloads.sim1_tfsf_shipped()
    sim, ids, pids = fdtd_amd.make_simulation(spec)
    sim.step(spec["nsteps"])
    rfl, inc = sim.probe(pids[0]), sim.probe(pids[1])
    assert len(inc) == int(z["nrows"])
    assert relerr(inc[::4, 1], z["inc"][:, 1]) < 1e-12
    assert relerr(inc[::4, 6], z["inc"][:, 2]) < 1e-12
    assert relerr(rfl[::4, 1], z["rfl"][:, 1]) < 1e-12
    assert relerr(rfl[::4, 6], z["rfl"][:, 2]) < 1e-12
    dt = spec["grid"]["courant"] * spec["grid"]["ds"] / workloads.C0
    syin = workloads.running_spectrum(inc, spec["freq"], dt, spec["grid"]["ds"], 1, 6)
    syr = workloads.running_spectrum(rfl, spec["freq"], dt, spec["grid"]["ds"], 1, 6)
    assert relerr(syin, z["syin"][:, 1]) < 1e-12
    assert relerr(syr, z["syr"][:, 1]) < 1e-12


def _both(spec, precision="fp64", engine="auto"):
    import fdtd_amd
    import oracle

    g, _, gp = fdtd_amd.make_simulation(spec, precision=precision, engine=engine)
    o, _, op = oracle.make_simulation(spec)
    g.step(spec["nsteps"])
    o.step(spec["nsteps"])
    return g, o, gp, op


def _field_err(g, o):
    return max(relerr(np.stack([g.field(q) for q in c]), np.stack([o.field(q) for q in c]))
               for c in ((0, 1, 2), (3, 4, 5)))


@pytest.mark.parametrize("axis,comp", ORIENT)
@pytest.mark.parametrize("engine", ["fused", "staged"])
def test_gpu_tfsf_orientations_vs_oracle(axis, comp, engine):
    spec = tfsf_box(axis, comp, True, nsteps=200)
    spec["tfsf_probes"] = [[0, 30]]
    g, o, gp, op = _both(spec, engine=engine)
    check(f"tfsf axis {axis} comp {comp} {engine} fields", _field_err(g, o), 1e-12)
    e_g, h_g = g.tfsf_line(0)
    e_o, h_o = o.tfsf_line(0)
    assert relerr(e_g, e_o) < 1e-12 and relerr(h_g, h_o) < 1e-12
    assert relerr(g.probe(gp[0]), o.probe(op[0])) < 1e-12


def _tfsf_through_pml(nsteps=240):
    """A y-propagating TFSF wave between x/z ALL_FACES PML slabs, with a
    dielectric block in the total-field region: the D and H terms cross the
    PML cells (split along y) and the block scatters into both regions."""
    from fdtd_amd import box

    spec = tfsf_box(1, 2, True, n_axis=80, nsteps=nsteps)
    spec["grid"].update({"nx": 24, "nz": 20, "periodic": [0, 0, 0]})
    spec["pml"] = {"mode": "all_faces", "npml": [6, 8, 6], "order": 3.0, "R0": 1e-8}
    spec["materials"] = [{"type": "dielectric", "region": box((1, 1, 1), (24, 80, 20)), "eps": 1.0},
                         {"type": "dielectric", "region": box((9, 38, 8), (16, 44, 13)), "eps": 4.0}]
    spec["probes"] = [[12, 20, 10], [12, 60, 10], [3, 30, 3]]
    return spec


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gpu_tfsf_through_all_faces_pml_vs_oracle(precision):
    spec = _tfsf_through_pml()
    g, o, gp, op = _both(spec, precision=precision, engine="fused")
    tol = 1e-12 if precision == "fp64" else 2e-4
    check(f"tfsf through pml {precision} fields", _field_err(g, o), tol)
    for a, b in zip(gp, op):
        assert relerr(g.probe(a), o.probe(b)) < (1e-12 if precision == "fp64" else 1e-3)
    # the block scatters back into the scattered-field region
    assert np.abs(o.probe(op[0])[:, 1:]).max() > 1e-3


def test_gpu_tfsf_with_reference_y_pml_vs_oracle():
    """The STAGED engine with the reference's y-slab PML: the TFSF planes
    lie between the slabs."""
    spec = tfsf_box(1, 0, True, n_axis=90, nsteps=150)
    spec["grid"]["periodic"] = [1, 0, 1]
    spec["pml"] = {"mode": "ref_y", "npml": [0, 10, 0], "order": 3.0, "R0": math.exp(-16.0),
                   "eps_low": 1.0, "eps_high": 1.0}
    g, o, _, _ = _both(spec)
    check("tfsf ref-y pml fields", _field_err(g, o), 1e-12)
