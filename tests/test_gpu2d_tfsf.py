"""TFSF along y on the 2D GPU path (include/fdtd2d.h, both polarisations)
against the oracle's restatement (itself pinned to the 3D oracle's TFSF at
nz = 1), the empty-grid scattered region, and a TFSF member in a batch."""
import numpy as np
import pytest

from conftest import check, relerr
from test_oracle2d_tfsf import tfsf_case

pytestmark = pytest.mark.gpu


def _max_rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("polarization,periodic_x", [("tez", True), ("tez", False), ("tmz", True), ("tmz", False)])
def test_tfsf_fp64_matches_the_oracle(polarization, periodic_x):
    """FP64: spectra 1e-12, fields and probes 1e-12, the incident line 1e-13
    (the pulse's exp and sin are the device's)."""
    import oracle
    from fdtd_amd import sim2d

    spec = tfsf_case(polarization, periodic_x=periodic_x, eps_bg=1.21 if not periodic_x else 1.0)
    gs, gi, gp = sim2d.make_simulation2d(spec, precision="fp64")
    os_, oi, op = oracle.make_simulation2d(spec)
    for n in (1, 40, 200, 259):
        gs.step(n)
        os_.step(n)
    comps = (sim2d.EZ2, sim2d.HX2, sim2d.HY2) if polarization == "tmz" else (sim2d.EX2, sim2d.EY2, sim2d.HZ2)
    for c in comps:
        check(f"test_gpu2d_tfsf:32 {c}", relerr(gs.field(c), os_.field(c)), 1e-12)
    for a, b in zip(gi, oi):
        assert _max_rel(gs.flux(a), os_.flux(b)) <= 1e-12
    for a, b in zip(gp, op):
        check("test_gpu2d_tfsf:36", _max_rel(gs.probe(a)[:, 1:], os_.probe(b)[:, 1:]), 1e-12)
    for x, y in zip(gs.tfsf_line(), os_.tfsf_line()):
        assert _max_rel(x, y) <= 1e-13
    gs.close()
    os_.close()


@pytest.mark.parametrize("polarization", ["tez", "tmz"])
def test_tfsf_empty_grid_fp64(polarization):
    from fdtd_amd import sim2d

    spec = tfsf_case(polarization, scatterer=False, nsteps=420)
    s, _, _ = sim2d.make_simulation2d(spec, precision="fp64")
    s.step(spec["nsteps"])
    f = s.field(sim2d.EZ2 if polarization == "tmz" else sim2d.EX2)
    inside = np.abs(f[:, 40:120]).max()
    outside = max(np.abs(f[:, 12:37]).max(), np.abs(f[:, 123:148]).max())
    assert inside > 1e-2
    assert outside <= 1e-10 * inside
    s.close()


def test_tfsf_fp32_spectra():
    import oracle
    from fdtd_amd import sim2d

    spec = tfsf_case("tez")
    gs, gi, _ = sim2d.make_simulation2d(spec, precision="fp32")
    os_, oi, _ = oracle.make_simulation2d(spec)
    gs.step(spec["nsteps"])
    os_.step(spec["nsteps"])
    errs = [_max_rel(gs.flux(a), os_.flux(b)) for a, b in zip(gi, oi)]
    print("2D TFSF fp32 spectra max rel err:", errs)
    assert max(errs) <= 1e-4
    gs.close()
    os_.close()


def test_tfsf_member_of_a_batch_equals_alone():
    from fdtd_amd import sim2d

    specs = [tfsf_case("tmz", periodic_x=False), sim2d.sweep_instance(np.random.default_rng(3), side=96, nsteps=300),
             tfsf_case("tez", nsteps=420)]
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
