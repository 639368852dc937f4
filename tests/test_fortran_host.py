"""The Fortran90 host (fortran/host3d.f90) drives the library through the
iso_c_binding interface (fortran/fdtd_c.f90): its flux and probe files equal
the same workload driven from Python through ctypes, and its derived
parameters equal the values the reference printed (sim3/out.out)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG, relerr

HOST = os.path.join(PKG, "fortran", "host3d")


@pytest.mark.gpu
def test_fortran_host_matches_the_python_driven_library(tmp_path):
    import fdtd_amd
    from fdtd_amd import workloads

    nsteps = 1500
    res = subprocess.run([HOST, str(nsteps), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    log = json.load(open(os.path.join(GOLDEN, "sim3_out_values.json")))
    assert f"{log['tfinStop']}" in res.stdout  # the host derived the reference's window
    spec = workloads.sim3_shipped(nsteps=nsteps)
    assert spec["tfinStop"] == log["tfinStop"]
    sim, ids, pids = fdtd_amd.make_simulation(spec)
    sim.step(nsteps)
    for name, pid in (("r0Syin.dat", ids[0]), ("r0Syr.dat", ids[1]), ("r0Syt.dat", ids[2])):
        f = np.loadtxt(os.path.join(tmp_path, name))
        assert f.shape == (100, 2)
        S = sim.flux(pid)
        lam = 2 * np.pi * fdtd_amd.C0 / np.array(spec["dft"][0]["omega"]) / 1e-9
        assert np.allclose(f[:, 0], lam, rtol=1e-14)
        if np.abs(S).max() > 0:
            assert relerr(f[:, 1], S) < 1e-14
        else:
            assert not f[:, 1].any()
    tr = np.loadtxt(os.path.join(tmp_path, "r0tfld.dat"))
    assert relerr(tr, sim.probe(pids[1])) < 1e-14
    inc = np.loadtxt(os.path.join(tmp_path, "r0infld.dat"))
    assert relerr(inc, sim.probe(pids[0])[: len(inc)]) < 1e-14
    assert np.abs(inc[:, 1]).max() > 0.1  # the pulse has passed the incident plane


HOST2D = os.path.join(PKG, "fortran", "host2d")


@pytest.mark.gpu
def test_fortran_2d_host_writes_the_reference_files(tmp_path):
    """fortran/host2d.f90 runs the reference's shipped 2D case through the
    2D C ABI and writes the files the reference's 3-rank run wrote; each
    equals the shipped file (golden fixture) to the parity bar."""
    for d in ("transforms", "timesLiney", "timesGnuplotmap"):
        os.makedirs(os.path.join(tmp_path, d))
    res = subprocess.run([HOST2D, "13696", str(tmp_path), "1"], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "3424" in res.stdout and "13696" in res.stdout
    g = np.load(os.path.join(GOLDEN, "ref2d_shipped.npz"))
    T = os.path.join(tmp_path, "transforms")
    for name in ("r0Syin", "r0Syr", "r2Syt"):
        f = np.loadtxt(os.path.join(T, name + ".dat"))
        assert np.max(np.abs(f[:, 0] - g[name][:, 0])) <= 1e-9
        assert relerr(f[:, 1], g[name][:, 1]) <= 1e-12, name
    rta = np.loadtxt(os.path.join(T, "RTA_2Dfdtd.data"))
    assert np.max(np.abs(rta[:, 1:] - g["rta"][:, 1:])) <= 1e-10
    for name, key in (("r0infld", "infld"), ("r0rfld", "rfld"), ("r2tfld", "tfld")):
        f = np.loadtxt(os.path.join(T, name + ".dat"))
        assert f.shape == g[key].shape
        for c in range(3):
            assert relerr(f[:, c], g[key][:, c]) <= 1e-11, name
    for fr in (5, 10, 20):
        line = np.concatenate([np.loadtxt(os.path.join(tmp_path, "timesLiney", f"r{r}E_time{fr}.dat"))
                               for r in range(3)])
        for c in range(3):
            assert relerr(line[:, c], g[f"line{fr}"][:, c]) <= 1e-11
    m = np.loadtxt(os.path.join(tmp_path, "timesGnuplotmap", "r2E_time20.dat"))
    assert np.array_equal(m[:, :2].astype(np.int32), g["map_ij"])
    assert relerr(m[:, 2], g["map_hz"]) <= 1e-11
