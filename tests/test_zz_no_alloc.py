"""fdtd_get_flux_noalloc (the flux through the library's fixed device array,
for callers that cannot allocate) against fdtd_get_flux, through the C ABI:
bit for bit, on two calls in a row with the second on a plane of fewer
frequencies (the array keeps the first call's longer contents), and the
refusals: more frequencies than the array holds, an uncommitted handle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sphere(nfreq, nsteps=120):
    from fdtd_amd import workloads

    spec = workloads.c1_gold_sphere(nsteps=nsteps, n=32)
    spec["pml"]["npml"] = [6, 6, 6]
    spec["materials"][1]["region"]["radius"] = 5.0
    spec["sources"][0]["index"] = 8
    om = list(workloads.wavelength_frequencies(400e-9, 900e-9, nfreq))
    spec["dft"] = [{"axis": 2, "index": 24, "omega": om, "comps": [0, 1, 3, 4]},
                   {"axis": 2, "index": 20, "omega": om[::8], "comps": [0, 1, 3, 4]}]
    return spec


def test_flux_without_allocation_equals_fdtd_get_flux():
    import fdtd_amd

    spec = _sphere(64)
    sim, ids, _ = fdtd_amd.make_simulation(spec, precision="fp64")
    sim.step(spec["nsteps"])
    big, small = ids  # 64 and 8 frequencies
    a = sim.flux_noalloc(big)
    b = sim.flux_noalloc(small)  # second call, fewer frequencies
    assert np.abs(a).max() > 0 and np.abs(b).max() > 0
    assert np.array_equal(a, sim.flux(big))
    assert np.array_equal(b, sim.flux(small))
    # a second handle on the same GPU shares the array
    spec32 = _sphere(16, nsteps=80)
    s32, ids32, _ = fdtd_amd.make_simulation(spec32, precision="fp32")
    s32.step(spec32["nsteps"])
    for i in ids32:
        assert np.array_equal(s32.flux_noalloc(i), s32.flux(i))
    assert np.array_equal(sim.flux_noalloc(big), a)
    s32.close()
    sim.close()


def test_flux_without_allocation_refuses_what_it_cannot_do():
    import fdtd_amd
    from fdtd_amd import box
    from fdtd_amd.abi import FdtdError

    sim = fdtd_amd.Simulation(8, 8, 8, 2e-9, 0.5, precision="fp64")
    sim.add_dielectric(box((1, 1, 1), (8, 8, 8)), 1.0)
    many = sim.add_dft_plane(2, 4, np.linspace(1e15, 3e15, 1025), comps=[0, 1, 3, 4])
    few = sim.add_dft_plane(2, 5, np.linspace(1e15, 3e15, 4), comps=[0, 1, 3, 4])
    with pytest.raises(FdtdError):
        sim.flux_noalloc(few)  # not committed yet: committing allocates
    sim.step(3)
    with pytest.raises(FdtdError):
        sim.flux_noalloc(many)  # 1025 frequencies, more than the array's 1024
    assert np.abs(sim.flux(many)).max() >= 0  # the allocating entry point takes it
    assert np.array_equal(sim.flux_noalloc(few), sim.flux(few))
    sim.close()
