"""2D TEz stepper on the GPU, through the C ABI (include/fdtd2d.h): the
reference's shipped 2D run (golden files), the oracle on sweep instances
(gold disk + Kubo graphene line), and batches against the same simulations
stepped alone."""
import os

import numpy as np
import pytest

from conftest import check, ROOT, relerr

pytestmark = pytest.mark.gpu


def golden(name):
    return np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))


def _shipped(precision):
    from fdtd_amd import sim2d

    spec = sim2d.ref2d_shipped()
    sim, ids, pids = sim2d.make_simulation2d(spec, precision=precision)
    fe = spec["frame_every"]
    frames, n = {}, 0
    for f in range(1, 21):
        sim.step(fe)
        n += fe
        if f in (5, 10, 20):
            frames[f] = (sim.field(sim2d.EX2), sim.field(sim2d.HZ2))
    sim.step(spec["nsteps"] - n)
    out = spec, [sim.flux(i) for i in ids], [sim.probe(p) for p in pids], frames
    sim.close()
    return out


def _max_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_shipped_run_fp64_matches_the_reference_files():
    """Spectra within 1e-12, traces and line frames within 1e-11 of the
    reference's own 3-rank output (relative to each quantity's maximum)."""
    g = golden("ref2d_shipped")
    spec, flux, probes, frames = _shipped("fp64")
    for S, name in zip(flux, ("r0Syin", "r0Syr", "r2Syt")):
        assert _max_rel(S, g[name][:, 1]) <= 1e-12, name
    tf = spec["tfinStop"]
    for got, ref in ((probes[0][:tf - 1], g["infld"]), (probes[0][tf:], g["rfld"]), (probes[1], g["tfld"])):
        for c in range(3):
            assert _max_rel(got[:, c], ref[:, c]) <= 1e-11
    ic = spec["ic"]
    for f, (ex, hz) in frames.items():
        assert _max_rel(ex[ic - 1, :], g[f"line{f}"][:, 1]) <= 1e-11
        assert _max_rel(hz[ic - 1, :], g[f"line{f}"][:, 2]) <= 1e-11
    ij = g["map_ij"]
    assert _max_rel(frames[20][1][ij[:, 0] - 1, ij[:, 1] - 1], g["map_hz"]) <= 1e-11


def test_shipped_run_fp32_spectra_error():
    """FP32 fields, FP64 DFT: spectra within 2e-4 of the reference's."""
    g = golden("ref2d_shipped")
    _, flux, _, _ = _shipped("fp32")
    errs = [_max_rel(S, g[name][:, 1]) for S, name in zip(flux, ("r0Syin", "r0Syr", "r2Syt"))]
    print("fp32 spectra max rel err (in, r, t):", errs)
    assert max(errs) <= 2e-4


def _sweep(side, nsteps, seed=0):
    from fdtd_amd import sim2d

    return sim2d.sweep_instance(np.random.default_rng(seed), side=side, nsteps=nsteps)


def test_sweep_instance_fp64_matches_the_oracle():
    import oracle
    from fdtd_amd import sim2d

    spec = _sweep(160, 2000, seed=3)
    gs, gi, gp = sim2d.make_simulation2d(spec, precision="fp64")
    os_, oi, op = oracle.make_simulation2d(spec)
    gs.step(spec["nsteps"])
    os_.step(spec["nsteps"])
    for a, b in zip(gi, oi):
        assert _max_rel(gs.flux(a), os_.flux(b)) <= 1e-12
    for comp in (sim2d.EX2, sim2d.EY2, sim2d.HZ2, sim2d.DX2, sim2d.DY2):
        check("test_gpu2d:87", relerr(gs.field(comp), os_.field(comp)), 1e-12)
    check("test_gpu2d:88", _max_rel(gs.probe(gp[0])[:, 1:], os_.probe(op[0])[:, 1:]), 1e-12)
    gs.close()
    os_.close()


def test_non_periodic_cp4_graphene_fp64_matches_the_oracle():
    """No x wrap, the reference's CP4 graphene with the 2D Raman term and
    a gold box, stepped in uneven calls."""
    import oracle
    from fdtd_amd import graphene_ref_dict, sim2d

    spec = _sweep(96, 900, seed=5)
    spec["grid"]["periodic_x"] = False
    gr = graphene_ref_dict()
    gr.update(ref_init_order=0, wNL=2.0e14, dnuNL=1.0e13, xi3=1e-3)
    spec["materials"][2]["params"] = gr
    spec["materials"].append({"type": "drude2cp", "region": sim2d.box2((10, 30), (40, 36))})
    gs, gi, _ = sim2d.make_simulation2d(spec, precision="fp64")
    os_, oi, _ = oracle.make_simulation2d(spec)
    for n in (1, 31, 33, 100, 735):
        gs.step(n)
        os_.step(n)
    assert gs.steps_done == os_.steps_done == 900
    for a, b in zip(gi, oi):
        assert _max_rel(gs.flux(a), os_.flux(b)) <= 1e-12
    for comp in (sim2d.EX2, sim2d.EY2, sim2d.HZ2):
        check("test_gpu2d:114", relerr(gs.field(comp), os_.field(comp)), 1e-12)
    gs.close()
    os_.close()


def test_batch_equals_each_simulation_alone():
    """Three sizes, three step counts, one batch: every result is bitwise
    the result of the same simulation stepped alone (FP32)."""
    from fdtd_amd import sim2d

    specs = [_sweep(96, 700, 11), _sweep(130, 1000, 12), _sweep(200, 1300, 13)]
    alone = []
    for sp in specs:
        s, ids, pids = sim2d.make_simulation2d(sp, precision="fp32")
        s.step(sp["nsteps"])
        alone.append(([s.flux(i) for i in ids], s.field(sim2d.HZ2), s.probe(pids[0])))
        s.close()
    built = [sim2d.make_simulation2d(sp, precision="fp32") for sp in specs]
    for s, _, _ in built:
        s.commit()
    batch = sim2d.Batch2D([s for s, _, _ in built])
    batch.run([sp["nsteps"] for sp in specs])
    batch.sync()
    for (s, ids, pids), sp, (fl, hz, pr) in zip(built, specs, alone):
        assert s.steps_done == sp["nsteps"]
        for a, b in zip([s.flux(i) for i in ids], fl):
            assert np.array_equal(a, b)
        assert np.array_equal(s.field(sim2d.HZ2), hz)
        assert np.array_equal(s.probe(pids[0]), pr)
    # a second run continues every member from where it stopped
    batch.run([10, 0, 33])
    batch.sync()
    assert [s.steps_done for s, _, _ in built] == [sp["nsteps"] + d for sp, d in zip(specs, (10, 0, 33))]
    batch.close()
    for s, _, _ in built:
        s.close()


def test_direct_launches_equal_the_graph(monkeypatch):
    """FDTD2D_GRAPH=0 enqueues each 32-step block's launches directly instead
    of replaying the captured graph: bitwise the same results (FP32)."""
    from fdtd_amd import sim2d

    specs = [_sweep(96, 300, 21), _sweep(130, 250, 22)]
    out = []
    for direct in (False, True):
        if direct:
            monkeypatch.setenv("FDTD2D_GRAPH", "0")
        built = [sim2d.make_simulation2d(sp, precision="fp32") for sp in specs]
        for s, _, _ in built:
            s.commit()
        batch = sim2d.Batch2D([s for s, _, _ in built])
        batch.run([sp["nsteps"] for sp in specs])
        batch.sync()
        out.append([([s.flux(i) for i in ids], s.field(sim2d.HZ2), s.probe(pids[0])) for s, ids, pids in built])
        batch.close()
        for s, _, _ in built:
            s.close()
    for (fa, ha, pa), (fb, hb, pb) in zip(*out):
        assert all(np.array_equal(a, b) for a, b in zip(fa, fb))
        assert np.array_equal(ha, hb) and np.array_equal(pa, pb)
