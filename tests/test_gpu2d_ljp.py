"""2D Lorentz-Jp cells on the GPU (fdtd2d_add_lorentz_jp, TEz and TMz)
against the oracle's restatement (itself pinned to the 3D oracle at
nz = 1), FP32 spectra, and a Lorentz-Jp member of a batch."""
import numpy as np
import pytest

from conftest import check, relerr
from test_oracle2d_ljp import ljp_case

pytestmark = pytest.mark.gpu


def _max_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("polarization,preset", [("tez", "gold"), ("tez", "si"), ("tez", "film_gold"),
                                                 ("tez", "graphene3_1nm"), ("tmz", "si"), ("tmz", "film_gold"),
                                                 ("tmz", "graphene3_1nm")])
def test_lorentz_jp_fp64_matches_the_oracle(polarization, preset):
    import oracle
    from fdtd_amd import sim2d

    spec = ljp_case(polarization, preset)
    gs, gi, gp = sim2d.make_simulation2d(spec, precision="fp64")
    os_, oi, op = oracle.make_simulation2d(spec)
    for n in (1, 31, 100, 268):
        gs.step(n)
        os_.step(n)
    comps = (sim2d.EZ2, sim2d.HX2, sim2d.HY2) if polarization == "tmz" else (sim2d.EX2, sim2d.EY2, sim2d.HZ2)
    for c in comps:
        check(f"test_gpu2d_ljp:32 {c}", relerr(gs.field(c), os_.field(c)), 1e-12)
    for a, b in zip(gi, oi):
        assert _max_rel(gs.flux(a), os_.flux(b)) <= 1e-12
    for a, b in zip(gp, op):
        check("test_gpu2d_ljp:36", _max_rel(gs.probe(a)[:, 1:], os_.probe(b)[:, 1:]), 1e-12)
    gs.close()
    os_.close()


def test_lorentz_jp_fp32_spectra():
    import oracle
    from fdtd_amd import sim2d

    spec = ljp_case("tez", "film_gold")
    gs, gi, _ = sim2d.make_simulation2d(spec, precision="fp32")
    os_, oi, _ = oracle.make_simulation2d(spec)
    gs.step(spec["nsteps"])
    os_.step(spec["nsteps"])
    errs = [_max_rel(gs.flux(a), os_.flux(b)) for a, b in zip(gi, oi)]
    print("2D Lorentz-Jp fp32 spectra max rel err:", errs)
    assert max(errs) <= 1e-4
    gs.close()
    os_.close()


def test_lorentz_jp_member_of_a_batch_equals_alone():
    from fdtd_amd import sim2d

    specs = [ljp_case("tmz", "graphene3_1nm", nsteps=300), ljp_case("tez", "si", nsteps=250),
             sim2d.sweep_instance(np.random.default_rng(5), side=96, nsteps=200)]
    alone = []
    for sp in specs:
        s, ids, _ = sim2d.make_simulation2d(sp, precision="fp32")
        s.step(sp["nsteps"])
        alone.append([s.flux(i) for i in ids])
        s.close()
    built = [sim2d.make_simulation2d(sp, precision="fp32") for sp in specs]
    for s, _, _ in built:
        s.commit()
    b = sim2d.Batch2D([s for s, _, _ in built])
    b.run([sp["nsteps"] for sp in specs])
    b.sync()
    for (s, ids, _), fl in zip(built, alone):
        for i, f in zip(ids, fl):
            assert np.array_equal(s.flux(i), f)
    b.close()
    for s, _, _ in built:
        s.close()


def test_lorentz_jp_in_the_pml_is_refused():
    from fdtd_amd import abi, sim2d

    s = sim2d.Simulation2D(30, 90, 2e-9, 0.5, polarization="tmz")
    s.set_pml((0, 8))
    s.add_lorentz_jp(sim2d.box2((1, 3), (30, 4)), {"preset": "si"})
    with pytest.raises(abi.FdtdError):
        s.commit()
    s.close()
