"""The scope's workloads on the GPU against the FP64 CPU oracle, at sizes
the oracle runs in seconds:

  C1  the 96^3 gold sphere in full (10-cell PML on all faces, 2000 steps,
      16 wavelengths on one z plane, 2 probes);
  C2  the 512^3 gold dimer's layout at 96^3 (the 12-cell PML, the seed's
      gap, r = 12 nm spheres, 64 wavelengths on the three orthogonal centre
      planes, the 8 probes), 1200 steps;
  C3  the graphene sheet over 16 jittered gold disks at 96 x 96 x 64
      (Kubo graphene with the Raman term, disks r = 6 nm, h = 8 nm), 600
      steps;
  C5  served jobs (sphere, rod and dimer from the served stream's draws at
      64-80 cells a side) through the job pool.

Tolerances, normwise (max |a-b| / max |b|): FP64 fields, spectra and probe
traces 1e-12; FP32 spectra 1e-5 against the FP64 oracle (FP32 fields, FP64
dispersive state and DFT accumulation), FP32 probe traces 1e-3.

The FP32 two-kernel step writes the monitor-plane samples from the H kernel
(up to four planes; further planes use the separate sample kernel):
FDTD_DFT_FUSE=0 turns that off, and the two runs must agree bit for bit."""
import functools

import numpy as np
import pytest

from conftest import check, relerr

pytestmark = pytest.mark.gpu


def c1_spec():
    from fdtd_amd import workloads

    return workloads.c1_gold_sphere()


def c2_spec(extra_planes=False):
    from fdtd_amd import workloads

    spec = workloads.c2_gold_dimer(n=96, radius=12.0, nsteps=1200)
    if extra_planes:  # five planes: four sampled by the H kernel, one by the sample kernel
        om = spec["dft"][0]["omega"]
        spec["dft"] += [{"axis": 0, "index": 70, "omega": om[::4], "comps": [1, 2, 4, 5]},
                        {"axis": 2, "index": 20, "omega": om[::2], "comps": [0, 1, 3, 4]}]
    return spec


def c3_spec():
    from fdtd_amd import workloads

    spec = workloads.c3_graphene_disks(nsteps=600, nx=96, ny=96, nz=64, nfreq=16, disk_radius=6.0,
                                       disk_height=8.0)
    return spec


SPECS = {"c1": c1_spec, "c2": c2_spec, "c2_5planes": lambda: c2_spec(True), "c3": c3_spec}


@functools.lru_cache(maxsize=None)
def oracle_run(name):
    """Fields, spectra and probe traces of the FP64 oracle (run once per
    workload per session)."""
    import oracle

    spec = SPECS[name]()
    sim, ids, pids = oracle.make_simulation(spec)
    sim.step(spec["nsteps"])
    out = {"fields": [sim.field(q) for q in range(6)], "flux": [sim.flux(i) for i in ids],
           "probes": [sim.probe(p) for p in pids]}
    sim.close()
    return out


def gpu_run(spec, precision, env=None, monkeypatch=None, dft=False):
    import fdtd_amd

    for k, v in (env or {}).items():
        monkeypatch.setenv(k, v)
    sim, ids, pids = fdtd_amd.make_simulation(spec, precision=precision)
    sim.step(spec["nsteps"])
    out = {"fields": [sim.field(q) for q in range(6)], "flux": [sim.flux(i) for i in ids],
           "probes": [sim.probe(p) for p in pids]}
    if dft:
        comps = [d["comps"] for d in spec["dft"]]
        out["dft"] = [[np.stack(sim.dft(i, c)) for c in cs] for i, cs in zip(ids, comps)]
    sim.close()
    return out


def _compare(name, g, o, precision):
    if precision == "fp64":
        for lab, comps in (("E", (0, 1, 2)), ("H", (3, 4, 5))):
            check(f"{name} fp64 {lab}", relerr(np.stack([g["fields"][q] for q in comps]),
                                              np.stack([o["fields"][q] for q in comps])), 1e-12)
    stol, ptol = (1e-12, 1e-12) if precision == "fp64" else (1e-5, 1e-3)
    for q, (a, b) in enumerate(zip(g["flux"], o["flux"])):
        check(f"{name} {precision} spectrum {q}", relerr(a, b), stol)
    for q, (a, b) in enumerate(zip(g["probes"], o["probes"])):
        check(f"{name} {precision} probe {q}", relerr(a, b), ptol)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["c1", "c2", "c3"])
def test_workload_vs_oracle(name, precision):
    spec = SPECS[name]()
    _compare(name, gpu_run(spec, precision), oracle_run(name), precision)


def test_c2_planes_are_all_exercised():
    """The reduced dimer keeps the three orthogonal centre planes (one per
    axis, 64 wavelengths), the 8 probes and the gold cells; every plane sees
    the pulse."""
    spec = SPECS["c2"]()
    assert sorted(d["axis"] for d in spec["dft"]) == [0, 1, 2]
    assert all(len(d["omega"]) == 64 for d in spec["dft"]) and len(spec["probes"]) == 8
    o = oracle_run("c2")
    assert all(np.abs(f).max() > 0 for f in o["flux"])


@pytest.mark.parametrize("onepass", ["0", "1"])
@pytest.mark.parametrize("name", ["c2", "c2_5planes"])
def test_fp32_fused_monitor_samples_equal_the_sample_kernel(name, onepass, monkeypatch):
    """FP32: the monitor samples written by the two-kernel step's H kernel
    and by the separate sample kernel give bit-identical accumulators and
    spectra (three planes: all fused; five: four fused, one not), and the
    spectra hold 1e-5 against the FP64 oracle, in both step forms (the
    one-pass step samples with the sample kernel either way)."""
    spec = SPECS[name]()
    a = gpu_run(spec, "fp32", env={"FDTD_ONEPASS": onepass}, monkeypatch=monkeypatch, dft=True)
    b = gpu_run(spec, "fp32", env={"FDTD_ONEPASS": onepass, "FDTD_DFT_FUSE": "0"}, monkeypatch=monkeypatch,
                dft=True)
    for pa, pb in zip(a["dft"], b["dft"]):
        for x, y in zip(pa, pb):
            assert np.array_equal(x, y)
    for x, y in zip(a["flux"], b["flux"]):
        assert np.array_equal(x, y)
    o = oracle_run(name)
    for q, (x, y) in enumerate(zip(a["flux"], o["flux"])):
        check(f"{name} fp32 spectrum {q}", relerr(x, y), 1e-5)


def served_jobs():
    """Sphere, rod and dimer jobs with the served stream's draws at 64, 72
    and 80 cells a side, 1500 steps."""
    from fdtd_amd import serve

    return [(0.0, serve.job_spec(np.random.default_rng(11 + q), side=64 + 8 * q, geom=g, nsteps=1500))
            for q, g in enumerate(("sphere", "rod", "dimer"))]


def test_served_jobs_vs_oracle():
    """Each served FP32 job's spectrum, stepped in the pool beside the
    others, against the FP64 oracle (1e-5); and the same in the mix as
    alone, bit for bit."""
    import oracle
    from fdtd_amd import serve

    jobs = served_jobs()
    mixed = serve.served_run(jobs, max_active=3, chunk_steps=64, realtime=False)
    alone = serve.sequential_run(jobs)
    for q, (_, spec) in enumerate(jobs):
        sim, ids, _ = oracle.make_simulation(spec)
        sim.step(spec["nsteps"])
        ref = sim.flux(ids[0])
        sim.close()
        check(f"served {spec['name']} fp32 spectrum", relerr(mixed["flux"][q], ref), 1e-5)
        assert np.array_equal(mixed["flux"][q], alone["flux"][q])
