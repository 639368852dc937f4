"""The reference's own shipped 3D program against this build.

tests/golden/sim3_run.npz holds the outputs of
FDTDsimulations/3D/air-AuAntennasSiO2-Si/program.f90 itself, compiled where
it lies and run under mpirun -np 6 (tests/golden/gen_golden_sim3.py): its
three flux spectra, its incident, reflected and transmitted time traces.  So
the driver-level lines of that program -- the plane source on Dx
(:396-406), the DFT gates n < tfinStop and n > tfinStop with t = (n-1)*dt
(:441-540), probe rows at t = n*dt (:443-451), the flux sum (:613-685), the
periodic x and z wrap of its halo exchange and the 6-rank decomposition --
are pinned by a run of the reference, not by a restatement of it.
"""
import numpy as np
import pytest

from conftest import GOLDEN, check, relerr

FIX = f"{GOLDEN}/sim3_run.npz"


def _fix():
    return np.load(FIX)


def test_fixture_is_the_shipped_window():
    """The run's tfinStop and step count are those the reference printed
    (sim3/out.out), and its spectra carry signal."""
    import json
    import os

    from fdtd_amd import workloads

    d = _fix()
    log = json.load(open(os.path.join(GOLDEN, "sim3_out_values.json")))
    spec = workloads.sim3_shipped()
    assert int(d["tfin"]) == log["tfinStop"] == spec["tfinStop"]
    assert json.loads(str(d["timing"]))["steps"] == spec["nsteps"] == 4 * log["tfinStop"]
    assert np.abs(d["Syin"]).max() > 0 and np.abs(d["Syt"]).max() > 0


def test_oracle_follows_the_program_for_800_steps():
    """The C oracle on the same case: its incident-plane trace equals the
    program's first 800 rows to 1e-12 (FP64, both without contraction)."""
    import oracle
    from fdtd_amd import workloads

    d = _fix()
    spec = workloads.sim3_shipped(nsteps=800)
    sim, ids, pids = oracle.make_simulation(spec)
    sim.step(800)
    tr = sim.probe(pids[0])
    sim.close()
    want = d["infld_head"]
    assert tr.shape == want.shape
    assert np.array_equal(tr[:, 0], want[:, 0]) or relerr(tr[:, 0], want[:, 0]) < 1e-15
    assert np.abs(want[:, 1]).max() > 0.1  # the pulse has reached the plane
    check("oracle vs sim3 program, incident trace 800 steps", relerr(tr, want), 1e-12)


@pytest.mark.gpu
def test_gpu_reproduces_the_program_run():
    """All 13,696 steps on the GPU in FP64: the three spectra and the three
    traces equal the program's to 1e-12."""
    import fdtd_amd
    from fdtd_amd import workloads

    d = _fix()
    spec = workloads.sim3_shipped()
    sim, ids, pids = fdtd_amd.make_simulation(spec, precision="fp64")
    sim.step(spec["nsteps"])
    lam = 2 * np.pi * fdtd_amd.C0 / np.array(spec["dft"][0]["omega"]) / 1e-9
    assert relerr(lam, d["lam"]) < 1e-14
    for name, pid in (("Syin", ids[0]), ("Syr", ids[1]), ("Syt", ids[2])):
        check(f"sim3 program {name}", relerr(sim.flux(pid), d[name]), 1e-12)
    inc = sim.probe(pids[0])
    tra = sim.probe(pids[1])
    sim.close()
    tfin = int(d["tfin"])
    for name, trace in (("infld", inc), ("rfld", inc), ("tfld", tra)):
        rows = d[name]
        n = rows[:, 0].astype(np.int64)
        if name == "infld":
            assert n.max() < tfin
        if name == "rfld":
            assert n.min() == tfin + 1
        check(f"sim3 program {name}", relerr(trace[n - 1], rows[:, 1:]), 1e-12)
