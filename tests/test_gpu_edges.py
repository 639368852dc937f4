"""Edge cases of the 3D path through the C ABI: ragged grids (sizes that
are no multiple of the 64-lane rows or of the tiles), unequal and absent PML
per axis, a periodic axis beside PML slabs, probes in the corner cells;
stepping in pieces against stepping at once; step(0); the argument checks
of the set-up calls.  Tolerances as in test_gpu_parity.py (FP64 fields,
spectra and traces 1e-12)."""
import numpy as np
import pytest

from conftest import check, relerr

pytestmark = pytest.mark.gpu


def _ragged(n, P, periodic, nsteps=200):
    from fdtd_amd import box, gold_params_dict, sphere
    from fdtd_amd.workloads import pulse_for_band, wavelength_frequencies

    nx, ny, nz = n
    c = ((nx + 1) / 2.0, (ny + 1) / 2.0, (nz + 1) / 2.0)
    pulse = pulse_for_band(400e-9, 900e-9, 600e-9)
    pulse["t0"] = 2.0 * pulse["st"]  # the peak leaves the source plane by step ~115 (dt = 16.7 as)
    return {
        "grid": {"nx": nx, "ny": ny, "nz": nz, "ds": 10e-9, "courant": 0.5, "periodic": list(periodic)},
        "pml": {"mode": "all_faces", "npml": list(P), "order": 3.0, "R0": 1e-6},
        "materials": [{"type": "dielectric", "region": box((1, 1, 1), n), "eps": 1.0},
                      {"type": "dielectric", "region": box((2, 2, P[2] + 4), (nx // 2, ny - 1, nz - P[2] - 3)),
                       "eps": 2.1},
                      {"type": "drude2cp", "region": sphere(c, min(nx, ny, nz) / 5.0), "params": gold_params_dict()}],
        "sources": [dict(axis=2, index=P[2] + 2, comp=0, amplitude=1.0, **pulse)],
        "dft": [{"axis": 2, "index": nz - P[2] - 2, "omega": list(wavelength_frequencies(400e-9, 900e-9, 5)),
                 "comps": [0, 1, 3, 4]}],
        "probes": [[1, 1, 1], [nx, ny, nz], [nx, 1, nz // 2], [int(c[0]), int(c[1]), int(c[2])]],
        "nsteps": nsteps,
    }


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("n,P,periodic", [((37, 29, 23), (4, 3, 5), (0, 0, 0)),
                                          ((65, 17, 19), (2, 4, 3), (0, 0, 0)),
                                          ((130, 12, 21), (5, 0, 3), (0, 1, 0)),
                                          ((9, 70, 14), (0, 6, 2), (1, 0, 0))])
def test_ragged_grids_vs_oracle(n, P, periodic, precision):
    """FP64 at the parity bar; FP32 fields 2e-4, spectra 1e-5, traces 1e-3
    (FP32 fields, FP64 dispersive state and DFT accumulators)."""
    import fdtd_amd
    import oracle

    spec = _ragged(n, P, periodic)
    g, gids, gp = fdtd_amd.make_simulation(spec, precision=precision)
    o, oids, op = oracle.make_simulation(spec)
    g.step(spec["nsteps"])
    o.step(spec["nsteps"])
    ftol, stol, ptol = (1e-12, 1e-12, 1e-12) if precision == "fp64" else (2e-4, 1e-5, 1e-3)
    for q in range(6):
        check(f"ragged {n} {precision} field {q}", relerr(g.field(q), o.field(q)), ftol)
    check(f"ragged {n} {precision} flux", relerr(g.flux(gids[0]), o.flux(oids[0])), stol)
    for a, b in zip(gp, op):
        ref = o.probe(b)
        if np.abs(ref).max() > 0:
            assert relerr(g.probe(a), ref) < ptol, (n, a)
        else:
            assert np.abs(g.probe(a)).max() == 0
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stepping_in_pieces_equals_stepping_at_once(precision):
    """Fields bitwise; the DFT folds its blocks of samples at different
    points, so spectra agree to rounding."""
    import fdtd_amd

    spec = _ragged((37, 29, 23), (4, 3, 5), (0, 0, 0), nsteps=100)
    a, aids, ap = fdtd_amd.make_simulation(spec, precision=precision)
    b, bids, bp = fdtd_amd.make_simulation(spec, precision=precision)
    a.step(100)
    for k in (0, 1, 30, 0, 2, 33, 34):
        b.step(k)
    for q in range(6):
        assert np.array_equal(a.field(q), b.field(q)), q
    assert relerr(a.flux(aids[0]), b.flux(bids[0])) < 1e-13
    for x, y in zip(ap, bp):
        assert np.array_equal(a.probe(x), b.probe(y))
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stepping_in_whole_dft_blocks_equals_stepping_at_once_bitwise(precision):
    """The DFT folds its blocks at steps that are multiples of the block
    (64) and where a call ends: calls that end on such steps (64, 128, then
    the rest) fold at the same steps as one call, so the spectra are equal
    bit for bit as well as the fields and the probes."""
    import fdtd_amd

    spec = _ragged((37, 29, 23), (4, 3, 5), (0, 0, 0), nsteps=200)
    a, aids, ap = fdtd_amd.make_simulation(spec, precision=precision)
    b, bids, bp = fdtd_amd.make_simulation(spec, precision=precision)
    a.step(200)
    for k in (64, 64, 72):
        b.step(k)
    for q in range(6):
        assert np.array_equal(a.field(q), b.field(q)), q
    assert np.array_equal(a.flux(aids[0]), b.flux(bids[0]))
    assert np.abs(a.flux(aids[0])).max() > 0
    for x, y in zip(ap, bp):
        assert np.array_equal(a.probe(x), b.probe(y))
    a.close()
    b.close()


def test_step_zero_changes_nothing():
    import fdtd_amd

    spec = _ragged((37, 29, 23), (4, 3, 5), (0, 0, 0), nsteps=10)
    s, _, _ = fdtd_amd.make_simulation(spec, precision="fp64")
    s.step(10)
    before = [s.field(q).copy() for q in range(6)]
    s.step(0)
    for q in range(6):
        assert np.array_equal(before[q], s.field(q))
    s.close()


def test_set_up_arguments_are_checked():
    import fdtd_amd
    from fdtd_amd import box, cylinder
    from fdtd_amd.abi import FdtdError

    def sim():
        return fdtd_amd.Simulation(20, 20, 20, 2e-9, 0.5, precision="fp64")

    s = sim()
    s.set_pml("all_faces", (11, 2, 2))  # thicker than half the grid: refused at commit
    with pytest.raises(FdtdError):
        s.commit()
    s.close()
    s = sim()
    # boxes reaching past the grid are clipped to it, as the reference's
    # material set-up clips them; a region of no known shape is refused
    s.add_dielectric(box((0, 1, 1), (25, 5, 5)), 2.0)
    with pytest.raises(FdtdError):
        s.add_dielectric(cylinder((5, 5, 5), 2.0, 3.0, 3), 2.0)  # no axis 3
    with pytest.raises(FdtdError):
        s.add_plane_source(2, 21, 0, 1e-15, 1e-15, 1e15)  # plane outside the grid
    with pytest.raises(FdtdError):
        s.add_plane_source(2, 5, 0, 1e-15, 0.0, 1e15)  # zero pulse width
    with pytest.raises(FdtdError):
        s.add_probe(0, 5, 5)
    with pytest.raises(FdtdError):
        s.add_dft_plane(3, 5, np.array([1e15]), comps=[0])
    with pytest.raises(FdtdError):
        s.add_dft_plane(2, 5, np.array([1e15]), comps=[])
    with pytest.raises(FdtdError):
        s.step(-1)
    s.add_dielectric(box((1, 1, 1), (20, 20, 20)), 1.0)
    s.step(1)
    with pytest.raises(FdtdError):
        s.add_probe(5, 5, 5)  # set-up after the first step
    s.close()
    p = fdtd_amd.Simulation(20, 20, 20, 2e-9, 0.5, precision="fp64", periodic=(1, 0, 0))
    p.set_pml("all_faces", (2, 2, 2))  # PML on a periodic axis
    with pytest.raises(FdtdError):
        p.commit()
    p.close()
