"""The Fortran90 host (fortran/host3d.f90) drives the library through the
iso_c_binding interface (fortran/fdtd_c.f90): its flux and probe files equal
the same workload driven from Python through ctypes, and its derived
parameters equal the values the reference printed (sim3/out.out)."""
import io
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG, relerr

HOST = os.path.join(PKG, "fortran", "host3d")


def _ftn(text):
    """Fortran Ew.d output drops the exponent letter when the exponent has
    three digits ("0.12-150"): put it back before parsing."""
    return re.sub(r"(?<=[0-9.])([+-][0-9]{3})(?![0-9])", r"E\1", text)


def _load(path):
    with open(path) as f:
        return np.loadtxt(io.StringIO(_ftn(f.read())))


def _vtk_values(path):
    with open(path) as f:
        head = [next(f).rstrip("\n") for _ in range(10)]
        body = np.array(_ftn(f.read()).split(), dtype=np.float64)
    return head, body


@pytest.mark.gpu
def test_fortran_host_matches_the_python_driven_library(tmp_path):
    """Flux, per-x reflected products, traces and one frame of each kind
    (line, map, VTK) written by the Fortran host equal the same workload
    driven from Python, in the reference's file names and formats
    (sim3/program.f90:362-376,543-604,613-685; misc/constants_mod.f90:45-61)."""
    import fdtd_amd
    from fdtd_amd import workloads

    nsteps = 1500
    for d in ("transforms", "timesLiney", "timesGnuplotmap", "timesVTK"):
        os.makedirs(os.path.join(tmp_path, d))
    res = subprocess.run([HOST, str(nsteps), str(tmp_path), "7", str(nsteps)], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    log = json.load(open(os.path.join(GOLDEN, "sim3_out_values.json")))
    assert f"{log['tfinStop']}" in res.stdout  # the host derived the reference's window
    spec = workloads.sim3_shipped(nsteps=nsteps)
    assert spec["tfinStop"] == log["tfinStop"]
    g = spec["grid"]
    nx, ny, nz, ds = g["nx"], g["ny"], g["nz"], g["ds"]
    n_y = ny // 6
    ic, kc = nx // 2, nz // 2
    ptty = spec["probes"][1][1]
    rt = (ptty - 1) // n_y
    sim, ids, pids = fdtd_amd.make_simulation(spec)
    sim.step(nsteps)
    T = os.path.join(tmp_path, "transforms")
    lam = 2 * np.pi * fdtd_amd.C0 / np.array(spec["dft"][0]["omega"]) / 1e-9
    for name, pid in (("r0Syin.dat", ids[0]), ("r0Syr.dat", ids[1]), (f"r{rt}Syt.dat", ids[2])):
        f = np.loadtxt(os.path.join(T, name))
        assert f.shape == (100, 2)
        S = sim.flux(pid)
        assert np.allclose(f[:, 0], lam, rtol=1e-14)
        if np.abs(S).max() > 0:
            assert relerr(f[:, 1], S) < 1e-14
        else:
            assert not f[:, 1].any()
    tr = _load(os.path.join(T, f"r{rt}tfld.dat"))
    assert relerr(tr, sim.probe(pids[1])) < 1e-14
    inc = _load(os.path.join(T, "r0infld.dat"))
    assert relerr(inc, sim.probe(pids[0])[: len(inc)]) < 1e-14
    assert np.abs(inc[:, 1]).max() > 0.1  # the pulse has passed the incident plane

    # per-x products of the reflected plane on z = kc, one block per frequency
    tf = _load(os.path.join(T, "r0tf_r.dat")).reshape(100, nx, 6)
    dre = {c: sim.dft(ids[1], c)[0][:, kc - 1, :] for c in (0, 2, 3, 5)}
    dim = {c: sim.dft(ids[1], c)[1][:, kc - 1, :] for c in (0, 2, 3, 5)}
    assert np.allclose(tf[:, :, 0], lam[:, None], atol=1e-5)
    assert np.allclose(tf[:, :, 1], (np.arange(1, nx + 1) * ds / 1e-9)[None, :], atol=1e-5)
    want = [dre[3] * dre[2], dim[3] * dim[2], dre[5] * dre[0], dim[5] * dim[0]]
    for c, w in enumerate(want):
        assert relerr(tf[:, :, 2 + c], w.T) < 1e-9

    F = np.stack([sim.field(c) for c in range(6)], axis=-1)  # (nx, ny, nz, 6)
    for r in range(6):
        rows = slice(r * n_y, (r + 1) * n_y)
        jg = np.arange(r * n_y + 1, (r + 1) * n_y + 1)
        line = np.loadtxt(os.path.join(tmp_path, "timesLiney", f"r{r}E_time1.dat"))
        assert line.shape == (n_y, 7)
        assert np.allclose(line[:, 0], jg * ds / 1e-9, atol=6e-6)
        assert np.allclose(line[:, 1:], F[ic - 1, rows, kc - 1, :], atol=6e-6, rtol=0)
        m = _load(os.path.join(tmp_path, "timesGnuplotmap", f"r{r}Ekc_time1.dat"))
        assert m.shape == (nx * n_y, 8)
        assert np.allclose(m[:, 0], np.tile(np.arange(1, nx + 1), n_y) * ds / 1e-9, rtol=1e-15)
        assert np.allclose(m[:, 1], np.repeat(jg, nx) * ds / 1e-9, rtol=1e-15)
        want = F[:, rows, kc - 1, :].transpose(1, 0, 2).reshape(-1, 6)
        assert relerr(m[:, 2:], want) < 1e-15
    assert np.abs(F[:, :, kc - 1, 0]).max() > 0.1
    for r in (0, 5):
        head, body = _vtk_values(os.path.join(tmp_path, "timesVTK", f"r{r}Ex_time1.vtk"))
        assert head[0] == "# vtk DataFile Version 2.0" and head[1] == "Ex" and head[3] == "DATASET STRUCTURED_POINTS"
        assert head[4].split() == ["DIMENSIONS", str(nx), str(n_y), str(nz)]
        assert head[7].split() == ["POINT_DATA", str(nx * n_y * nz)]
        want = F[:, r * n_y:(r + 1) * n_y, :, 5].reshape(-1, order="F")
        assert body.shape == want.shape
        assert relerr(body, want) < 1e-15
    assert not os.path.exists(os.path.join(tmp_path, "timesVTK", "r1Ex_time1.vtk"))


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
