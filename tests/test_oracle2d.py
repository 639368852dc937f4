"""The 2D oracle against the reference's shipped 2D run (golden vectors from
FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate/: transforms/,
timesLiney/, timesGnuplotmap/), plus closed-form pins of its parts."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT


def golden(name):
    return np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))


@pytest.fixture(scope="module")
def shipped_run():
    import oracle
    from fdtd_amd import sim2d

    spec = sim2d.ref2d_shipped()
    sim, ids, pids = oracle.make_simulation2d(spec)
    sim.commit()
    fe = spec["frame_every"]
    frames = {}
    n = 0
    # the shipped output comes from a build whose complex division is C's
    with oracle.complex_division("shipped"):
        for f in range(1, 21):
            sim.step(fe)
            n += fe
            if f in (5, 10, 20):
                frames[f] = (sim.field(sim2d.EX2), sim.field(sim2d.HZ2))
        sim.step(spec["nsteps"] - n)
    out = {"spec": spec, "flux": [sim.flux(i) for i in ids], "probes": [sim.probe(p) for p in pids],
           "frames": frames}
    sim.close()
    return out


def test_shipped_case_parameters_match_the_reference_log():
    """sim2/out.out: the values the reference printed for its run."""
    from fdtd_amd import sim2d

    s = sim2d.ref2d_shipped()
    assert (s["grid"]["nx"], s["grid"]["ny"]) == (125, 1266)
    assert s["tfinStop"] == 3424 and s["nsteps"] == 13696 and s["frame_every"] == 684
    assert abs(s["pml"]["sigmam"] - 2.3989393988497120e17) < 1e3
    assert s["materials"][1]["region"] == {"kind": "box", "lo": [50, 1008], "hi": [74, 1012]}
    assert s["materials"][3]["region"]["lo"][1] == 1084 and s["materials"][3]["region"]["hi"][1] == 1256
    assert s["probes"] == [[62, 14], [62, 1170]]


def test_spectra_match_the_shipped_run(shipped_run):
    g = golden("ref2d_shipped")
    for S, name in zip(shipped_run["flux"], ("r0Syin", "r0Syr", "r2Syt")):
        ref = g[name][:, 1]
        assert np.max(np.abs(S - ref)) <= 1e-13 * np.max(np.abs(ref)), name
    from fdtd_amd import sim2d
    lam = 2 * math.pi * sim2d.C0 / np.array(shipped_run["spec"]["dft"][0]["omega"]) / 1e-9
    assert np.max(np.abs(lam - g["r0Syin"][:, 0])) <= 1e-10


def test_traces_match_the_shipped_run(shipped_run):
    g = golden("ref2d_shipped")
    tf = shipped_run["spec"]["tfinStop"]
    p0, p1 = shipped_run["probes"]
    for got, ref in ((p0[:tf - 1], g["infld"]), (p0[tf:], g["rfld"]), (p1, g["tfld"])):
        assert got.shape == ref.shape
        # the reference printed 19-20 significant digits (F30.19 for rfld/tfld)
        for c in range(3):
            assert np.max(np.abs(got[:, c] - ref[:, c])) <= 1e-12 * np.max(np.abs(ref[:, c]))


def test_line_frames_and_map_match_the_shipped_run(shipped_run):
    g = golden("ref2d_shipped")
    ic = shipped_run["spec"]["ic"]
    for f, (ex, hz) in shipped_run["frames"].items():
        ref = g[f"line{f}"]
        assert np.max(np.abs(ex[ic - 1, :] - ref[:, 1])) <= 1e-12 * np.max(np.abs(ref[:, 1]))
        assert np.max(np.abs(hz[ic - 1, :] - ref[:, 2])) <= 1e-12 * np.max(np.abs(ref[:, 2]))
    ex, hz = shipped_run["frames"][int(g["map_frame"])]
    ij = g["map_ij"]
    got = hz[ij[:, 0] - 1, ij[:, 1] - 1]
    assert np.max(np.abs(got - g["map_hz"])) <= 1e-12 * np.max(np.abs(g["map_hz"]))


def _vacuum_box(nx, ny, ds=2e-9, ku=0.5, periodic=True):
    import oracle
    from fdtd_amd import sim2d

    # cells no material covers keep E (the reference has no default medium):
    # vacuum is a dielectric of eps 1 over the whole domain
    spec = {"grid": {"nx": nx, "ny": ny, "ds": ds, "courant": ku, "periodic_x": periodic},
            "materials": [{"type": "dielectric", "region": sim2d.box2((1, 1), (nx, ny)), "eps": 1.0}]}
    return oracle.make_simulation2d(spec)[0]


def test_periodic_standing_wave_follows_the_discrete_dispersion():
    """No PML, periodic x, zero ghosts in y (Hz(j=0) = 0, Ex(j=ny+1) = 0):
    Hz = cos(kx i) sin(ky j) with kx = 2 pi p/nx, ky = pi (2m+1)/(2 ny+1) is
    a mode of the discrete curl-curl, so its amplitude a(n) obeys
    a(n+1) + a(n-1) = 2 cos(w) a(n), sin^2(w/2) = Ku^2 (sin^2(kx/2) + sin^2(ky/2))."""
    from fdtd_amd import sim2d

    nx, ny, ku = 32, 24, 0.5
    sim = _vacuum_box(nx, ny, ku=ku)
    kx = 2 * math.pi * 2 / nx
    ky = math.pi * 3 / (2 * ny + 1)
    i = np.arange(1, nx + 1)[:, None]
    j = np.arange(1, ny + 1)[None, :]
    hz0 = np.cos(kx * i) * np.sin(ky * j)
    sim.set_field(sim2d.HZ2, hz0)
    w = 2 * math.asin(ku * math.sqrt(math.sin(kx / 2) ** 2 + math.sin(ky / 2) ** 2))
    amps = []
    for _ in range(60):
        sim.step(1)
        amps.append(float(np.sum(sim.field(sim2d.HZ2) * hz0) / np.sum(hz0 * hz0)))
    a = np.array(amps)
    # a(n+1) + a(n-1) = 2 cos(w) a(n) for the discrete mode
    resid = a[2:] + a[:-2] - 2 * math.cos(w) * a[1:-1]
    assert np.max(np.abs(resid)) < 1e-12
    sim.close()


def test_pml_absorbs_a_normally_incident_pulse():
    """With the reference's layout (x grading over npmlx cells inside the y
    slabs) a normally incident pulse comes back from each slab below 1e-3 of
    its amplitude.  The split fields never wrap in x (BERpmly_HzExEy_mod.f90
    updates i = 1..N_x with zero ghosts), so the x grading is what absorbs
    at the slab's x ends: without it, or in a domain a few cells wide, the
    slabs reflect strongly -- the reference's behaviour, kept."""
    import oracle
    from fdtd_amd import sim2d

    nx, ny, P = 64, 1600, 10
    f0 = sim2d.C0 / 600e-9
    st = 1.0 / f0
    spec = {"grid": {"nx": nx, "ny": ny, "ds": 4e-9, "courant": 0.5, "periodic_x": True},
            "pml": {"npml": [P, P], "order": 3.0, "R0": math.exp(-16.0), "eps_low": 1.0, "eps_high": 1.0},
            "materials": [{"type": "dielectric", "region": sim2d.box2((1, P), (nx, ny - P)), "eps": 1.0}],
            "sources": [{"j": 400, "t0": 4 * st, "st": st, "w0": 2 * math.pi * f0}],
            "probes": [[nx // 2, 200]]}
    sim, _, pids = oracle.make_simulation2d(spec)
    sim.step(8000)
    a = np.abs(sim.probe(pids[0])[:, 1])
    direct = a[1000:2500].max()   # source -> probe
    low = a[3000:4500].max()      # probe -> low slab -> probe
    high = a[5500:7500].max()     # source -> high slab -> probe
    assert direct > 0.9
    assert low < 1e-3 * direct and high < 1e-3 * direct
    sim.close()
