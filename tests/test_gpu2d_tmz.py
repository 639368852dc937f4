"""2D TMz (Ez, Hx, Hy) on the GPU through the C ABI (include/fdtd2d.h)
against the oracle's TMz restatement (itself equal, bit for bit, to the 3D
oracle at nz = 1), and TMz members in a batch beside TEz ones."""
import numpy as np
import pytest

from conftest import check, relerr

pytestmark = pytest.mark.gpu


def _max_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("periodic_x,graphene", [(False, "cp4"), (True, "kubo"), (True, "cp4")])
def test_tmz_fp64_matches_the_oracle(periodic_x, graphene):
    """FP64: spectra within 1e-12, fields and traces within 1e-12 (relative
    to each quantity's maximum; both are built without multiply-add
    contraction), stepped in uneven calls across the 32-step graph blocks."""
    import oracle
    from fdtd_amd import sim2d

    spec = sim2d.tmz_case(periodic_x=periodic_x, graphene=graphene, nsteps=600)
    gs, gi, gp = sim2d.make_simulation2d(spec, precision="fp64")
    os_, oi, op = oracle.make_simulation2d(spec)
    for n in (1, 31, 33, 100, 435):
        gs.step(n)
        os_.step(n)
    assert gs.steps_done == os_.steps_done == 600
    for a, b in zip(gi, oi):
        assert _max_rel(gs.flux(a), os_.flux(b)) <= 1e-12
        for comp in (sim2d.EZ2, sim2d.HX2):
            for x, y in zip(gs.dft(a, comp), os_.dft(b, comp)):
                check("test_gpu2d_tmz:35", _max_rel(x, y), 1e-12)
    for comp in (sim2d.EZ2, sim2d.HX2, sim2d.HY2, sim2d.DZ2):
        check("test_gpu2d_tmz:37", relerr(gs.field(comp), os_.field(comp)), 1e-12)
    for a, b in zip(gp, op):
        check("test_gpu2d_tmz:39", _max_rel(gs.probe(a)[:, 1:], os_.probe(b)[:, 1:]), 1e-12)
    gs.close()
    os_.close()


def test_tmz_fp32_spectra_error():
    """FP32 fields with FP64 DFT accumulation against the FP64 oracle."""
    import oracle
    from fdtd_amd import sim2d

    spec = sim2d.tmz_case(periodic_x=False, graphene="cp4", nsteps=600)
    gs, gi, _ = sim2d.make_simulation2d(spec, precision="fp32")
    os_, oi, _ = oracle.make_simulation2d(spec)
    gs.step(spec["nsteps"])
    os_.step(spec["nsteps"])
    errs = [_max_rel(gs.flux(a), os_.flux(b)) for a, b in zip(gi, oi)]
    print("TMz fp32 spectra max rel err:", errs)
    assert max(errs) <= 1e-4
    gs.close()
    os_.close()


def test_mixed_polarisation_batch_equals_each_alone():
    """TEz and TMz members of one FP32 batch, different sizes and step
    counts: every result is bitwise the result of the member alone."""
    from fdtd_amd import sim2d

    specs = [sim2d.sweep_instance(np.random.default_rng(21), side=96, nsteps=500),
             sim2d.tmz_case(periodic_x=False, graphene="cp4", nsteps=700),
             sim2d.tmz_case(periodic_x=True, graphene="kubo", nx=72, ny=60, nsteps=333)]
    comps = [sim2d.HZ2, sim2d.EZ2, sim2d.HY2]
    alone = []
    for sp, c in zip(specs, comps):
        s, ids, pids = sim2d.make_simulation2d(sp, precision="fp32")
        s.step(sp["nsteps"])
        alone.append(([s.flux(i) for i in ids], s.field(c), s.probe(pids[0])))
        s.close()
    built = [sim2d.make_simulation2d(sp, precision="fp32") for sp in specs]
    for s, _, _ in built:
        s.commit()
    batch = sim2d.Batch2D([s for s, _, _ in built])
    batch.run([sp["nsteps"] for sp in specs])
    batch.sync()
    for (s, ids, pids), sp, c, (fl, f, pr) in zip(built, specs, comps, alone):
        assert s.steps_done == sp["nsteps"]
        for a, b in zip([s.flux(i) for i in ids], fl):
            assert np.array_equal(a, b)
        assert np.array_equal(s.field(c), f)
        assert np.array_equal(s.probe(pids[0]), pr)
    batch.close()
    for s, _, _ in built:
        s.close()


def test_tmz_argument_errors():
    from fdtd_amd import abi, sim2d

    s = sim2d.Simulation2D(40, 48, 2e-9, 0.6, periodic_x=True, polarization="tmz")
    with pytest.raises(abi.FdtdError):
        s.set_pml((4, 8))  # x PML on a periodic axis
    with pytest.raises(abi.FdtdError):
        s.add_dft_line(10, np.array([1e15]), sim2d.DFT_HZ)
    s.set_pml((0, 8))
    s.add_drude2cp(sim2d.disk((20, 6), 4.0))
    with pytest.raises(abi.FdtdError):
        s.commit()  # dispersive cells inside the PML
    s.close()


def test_tmz_sweep_instance_fp64_matches_the_oracle():
    """A TMz member of the batched sweep (gold disk, Kubo graphene line,
    periodic x, y PML only)."""
    import oracle
    from fdtd_amd import sim2d

    spec = sim2d.sweep_instance(np.random.default_rng(4), side=128, nsteps=1500, polarization="tmz")
    gs, gi, gp = sim2d.make_simulation2d(spec, precision="fp64")
    os_, oi, op = oracle.make_simulation2d(spec)
    gs.step(spec["nsteps"])
    os_.step(spec["nsteps"])
    for a, b in zip(gi, oi):
        assert _max_rel(gs.flux(a), os_.flux(b)) <= 1e-12
    for comp in (sim2d.EZ2, sim2d.HX2, sim2d.HY2):
        check("test_gpu2d_tmz:122", relerr(gs.field(comp), os_.field(comp)), 1e-12)
    check("test_gpu2d_tmz:123", _max_rel(gs.probe(gp[0])[:, 1:], os_.probe(op[0])[:, 1:]), 1e-12)
    gs.close()
    os_.close()
