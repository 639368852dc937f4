"""The one-pass step (k_march: E and H of a step in one kernel, each
wavefront marching a 64 x R tile through a chunk of planes over ping-pong
copies) against the two-kernel FUSED step (k_fused_e + k_fused_h,
FDTD_ONEPASS=0) on the same inputs: the per-cell arithmetic is the same
term for term and neither build contracts multiply-adds, so fields,
spectra and probe traces are bitwise equal in FP64 and in FP32.  Cases
cover partial tiles and chunks (tile height and chunk depth varied, grids
off the 64-lane rows), periodic wrap of the halo column, row and plane,
ALL_FACES PML on some or all faces, dispersive cells (written by k_ade
before the pass), plane sources through PML cells, and TFSF terms."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _periodic_box():
    from fdtd_amd import box, sphere
    from fdtd_amd.workloads import pulse_for_band

    n = (130, 31, 9)
    return {"grid": {"nx": n[0], "ny": n[1], "nz": n[2], "ds": 2e-9, "courant": 0.5, "periodic": [1, 1, 1]},
            "materials": [{"type": "dielectric", "region": box((1, 1, 1), n), "eps": 1.0},
                          {"type": "dielectric", "region": sphere((40, 15, 5), 6.0), "eps": 2.25}],
            "sources": [dict(axis=0, index=7, comp=1, amplitude=1.0, **pulse_for_band(400e-9, 900e-9, 600e-9))],
            "probes": [[1, 1, 1], [130, 31, 9], [60, 16, 5]],
            "nsteps": 300}


def _cases():
    from test_gpu_parity import _small_sphere
    from test_gpu_tfsf import _tfsf_through_pml
    from fdtd_amd import workloads

    graphene = workloads.c3_graphene_disks(nsteps=120, nx=48, ny=48, nz=40, nfreq=8)
    graphene["pml"]["npml"] = [6, 6, 6]
    for m in graphene["materials"]:
        if m["type"] == "drude2cp":
            m["region"]["radius"] = 3.0
            m["region"]["half_length"] = 2.0
    graphene["sources"][0]["index"] = 8
    graphene["dft"] = [{"axis": 2, "index": 32, "omega": graphene["dft"][0]["omega"], "comps": [0, 1, 3, 4]}]
    return {"gold_sphere": _small_sphere(nsteps=200), "periodic_box": _periodic_box(),
            "graphene_disks": graphene, "tfsf_pml": _tfsf_through_pml(nsteps=200), "x_edges": _x_edges()}


def _x_edges():
    """96 cells along x with a 12-cell x PML: one column of plain tiles (x
    33-64 in FP32) between general x edges 32 cells wide, a gold sphere and
    an x-normal source plane inside the low edge, an x-normal DFT plane and
    probes in the high one."""
    from fdtd_amd import workloads
    from fdtd_amd import sphere

    spec = workloads.c1_gold_sphere(nsteps=200, n=96)
    spec["grid"] = dict(spec["grid"], ny=40, nz=36)
    spec["pml"]["npml"] = [12, 6, 6]
    spec["materials"][0]["region"] = dict(spec["materials"][0]["region"], hi=[96, 40, 36])
    spec["materials"][1]["region"] = sphere((22.5, 20.5, 18.5), 5.0)
    spec["sources"] = [dict(spec["sources"][0], axis=0, index=20, comp=2)]
    spec["dft"] = [dict(spec["dft"][0], axis=0, index=80, comps=[1, 2, 4, 5])]
    spec["probes"] = [[22, 20, 18], [80, 20, 18], [5, 20, 18]]
    return spec


def _run(spec, precision, env, monkeypatch):
    import fdtd_amd

    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sim, ids, pids = fdtd_amd.make_simulation(spec, precision=precision, engine="fused")
    sim.step(spec["nsteps"])
    return sim, ids, pids


def _same(a, ia, pa, b, ib, pb):
    for c in range(6):
        assert np.array_equal(a.field(c), b.field(c)), c
    for x, y in zip(ia, ib):
        assert np.array_equal(a.flux(x), b.flux(y))
    for x, y in zip(pa, pb):
        assert np.array_equal(a.probe(x), b.probe(y))
    assert max(np.abs(a.field(c)).max() for c in range(3)) > 1e-4  # something propagated


@pytest.mark.parametrize("case", ["gold_sphere", "periodic_box", "graphene_disks", "tfsf_pml"])
@pytest.mark.parametrize("by,kc", [(2, 64), (4, 7), (8, 5), (2, 3), (4, 16), (8, 64)])
def test_onepass_equals_two_kernel_step_fp64(case, by, kc, monkeypatch):
    spec = _cases()[case]
    a, ia, pa = _run(spec, "fp64", {"FDTD_ONEPASS": "0"}, monkeypatch)
    b, ib, pb = _run(spec, "fp64", {"FDTD_ONEPASS": "1", "FDTD_ONEPASS_BY": str(by), "FDTD_ONEPASS_KC": str(kc)},
                     monkeypatch)
    _same(a, ia, pa, b, ib, pb)


@pytest.mark.parametrize("case", ["gold_sphere", "periodic_box", "graphene_disks", "tfsf_pml", "x_edges"])
@pytest.mark.parametrize("by,kc", [(2, 64), (4, 9)])
def test_onepass_equals_two_kernel_step_fp32(case, by, kc, monkeypatch):
    spec = _cases()[case]
    a, ia, pa = _run(spec, "fp32", {"FDTD_ONEPASS": "0"}, monkeypatch)
    b, ib, pb = _run(spec, "fp32", {"FDTD_ONEPASS": "1", "FDTD_ONEPASS_BY": str(by), "FDTD_ONEPASS_KC": str(kc)},
                     monkeypatch)
    _same(a, ia, pa, b, ib, pb)


def _x_face_cases():
    """Ragged grids for the x-face tiles: a tile holding both x slabs (37
    cells), the second tile's lanes reaching into the high slab (100), the
    first tile left general because the cell after it lies in the high slab
    (70, periodic z without z slabs)."""
    from test_gpu_edges import _ragged

    r37 = _ragged((37, 29, 23), (4, 3, 5), (0, 0, 0), nsteps=150)
    del r37["materials"][1]  # no dielectric block across the whole tile
    return {"ragged_37": r37,
            "ragged_100": _ragged((100, 24, 30), (12, 4, 4), (0, 0, 0), nsteps=150),
            "ragged_70_zper": _ragged((70, 20, 16), (12, 4, 0), (0, 0, 1), nsteps=150)}


@pytest.mark.parametrize("case", ["gold_sphere", "x_edges", "graphene_disks", "ragged_37", "ragged_100",
                                  "ragged_70_zper"])
@pytest.mark.parametrize("bx", ["2", "4"])
def test_x_face_tiles_equal_general_x_edges_fp32(case, bx, monkeypatch):
    """FP32 one-pass tiles whose lanes reach into an x PML slab (x-face
    tiles, their own launch) update those lanes with the split-field rule
    of the general kernels: the run equals the one with the x edges left to
    the general kernels (FDTD_XFACE=0) bit for bit, and the x-face launch
    did run (its timer counted launches, and its algorithmic bytes include
    x-slab cells)."""
    from fdtd_amd import abi

    own = case in ("gold_sphere", "x_edges", "graphene_disks")
    spec = _cases()[case] if own else _x_face_cases()[case]
    # 4-plane chunks on the small ragged grids, so that some chunks miss the
    # source plane, the sphere and the dielectric block
    env = {"FDTD_ONEPASS": "1", "FDTD_ONEPASS_BY": "4", "FDTD_ONEPASS_KC": "16" if own else "4",
           "FDTD_XFACE_BY": bx}
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    import fdtd_amd

    b, ib, pb = fdtd_amd.make_simulation(spec, precision="fp32", engine="fused")
    b.set_profiling(True)
    b.step(spec["nsteps"])
    assert b.kernel_stats(abi.K_XF)[1] > 0
    assert b.algorithmic_bytes(abi.K_XF) > 0
    monkeypatch.setenv("FDTD_XFACE", "0")
    a, ia, pa = fdtd_amd.make_simulation(spec, precision="fp32", engine="fused")
    a.set_profiling(True)
    a.step(spec["nsteps"])
    assert a.kernel_stats(abi.K_XF)[1] == 0
    _same(a, ia, pa, b, ib, pb)


def test_autotune_leaves_the_state_untouched(monkeypatch):
    """fdtd_autotune times the step forms with dry steps on the zero fields:
    the run that follows equals an untuned run bit for bit (every step form
    does the same arithmetic), the step count stays 0, and tuning after a
    step or after set_field is refused."""
    import fdtd_amd
    from fdtd_amd.abi import FdtdError

    spec = _cases()["gold_sphere"]
    monkeypatch.setenv("FDTD_ONEPASS", "1")
    a, ia, _ = fdtd_amd.make_simulation(spec, engine="fused")
    by, kc = a.autotune(2)
    assert (by == 0 and kc in (1, 2, 4, 8)) or (by in (2, 4, 8) and kc in (16, 32, 64, 128))
    assert a.steps_done == 0
    for c in range(6):
        assert not np.abs(a.field(c)).any()
    a.step(spec["nsteps"])
    b, ib, _ = fdtd_amd.make_simulation(spec, engine="fused")
    b.step(spec["nsteps"])
    for c in range(6):
        assert np.array_equal(a.field(c), b.field(c))
    assert np.array_equal(a.flux(ia[0]), b.flux(ib[0]))
    with pytest.raises(FdtdError):
        a.autotune(1)
    c_, _, _ = fdtd_amd.make_simulation(spec, engine="fused")
    c_.commit()
    c_.set_field(0, np.zeros(c_.shape))
    with pytest.raises(FdtdError):
        c_.autotune(1)


@pytest.mark.parametrize("case", ["gold_sphere", "periodic_box", "graphene_disks", "tfsf_pml"])
@pytest.mark.parametrize("zm", [2, 4, 8])
def test_zmarch_equals_one_plane_tiles_bitwise(case, zm, monkeypatch):
    """The two-kernel step with tiles of zm planes (each thread carrying the
    neighbour plane in registers) loads the same values and does the same
    arithmetic as one-plane tiles: FP64 fields, fluxes and probes are
    bitwise equal, partial chunks (nz not a multiple of zm) and mixed rows
    between uniform ones included."""
    spec = _cases()[case]
    a, ia, pa = _run(spec, "fp64", {"FDTD_ONEPASS": "0", "FDTD_ZMARCH": "1"}, monkeypatch)
    b, ib, pb = _run(spec, "fp64", {"FDTD_ONEPASS": "0", "FDTD_ZMARCH": str(zm)}, monkeypatch)
    for c in range(6):
        assert np.array_equal(a.field(c), b.field(c)), (case, c)
    for x, y in zip(ia, ib):
        assert np.array_equal(a.flux(x), b.flux(y))
    for x, y in zip(pa, pb):
        assert np.array_equal(a.probe(x), b.probe(y))
    assert max(np.abs(a.field(c)).max() for c in range(3)) > 1e-4


def test_zmarch_fp32_bitwise():
    import fdtd_amd

    spec = _cases()["gold_sphere"]
    out = []
    for zm in ("1", "4"):
        import os
        os.environ["FDTD_ZMARCH"] = zm
        try:
            sim, ids, _ = fdtd_amd.make_simulation(spec, precision="fp32", engine="fused")
        finally:
            del os.environ["FDTD_ZMARCH"]
        sim.step(spec["nsteps"])
        out.append([sim.field(c) for c in range(6)] + [sim.flux(i) for i in ids])
    for x, y in zip(*out):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_full_size_dimer_forms_agree_and_resume(precision, monkeypatch):
    """The benchmark's own 512^3 gold dimer (12-cell PML, 64 wavelengths on
    three planes, 8 probes) at full size, where the oracle cannot follow:
    the one-pass step with the tile shape the autotuner keeps equals the
    two-kernel step bit for bit after 96 steps (fields, spectra, probes),
    and a run saved at step 40 and resumed in the other form lands on the
    same bits."""
    import fdtd_amd
    from fdtd_amd import workloads

    spec = workloads.c2_gold_dimer(nsteps=96)
    a, ia, pa = _run(spec, precision, {"FDTD_ONEPASS": "0"}, monkeypatch)
    monkeypatch.setenv("FDTD_ONEPASS", "1")
    monkeypatch.setenv("FDTD_ONEPASS_BY", "4")
    monkeypatch.setenv("FDTD_ONEPASS_KC", "16")
    b, ib, pb = fdtd_amd.make_simulation(spec, precision=precision, engine="fused")
    b.step(40)
    ck = b.checkpoint()
    b.step(56)
    _same(a, ia, pa, b, ib, pb)
    b.close()
    monkeypatch.setenv("FDTD_ONEPASS", "0")
    c, ic, pc = fdtd_amd.make_simulation(spec, precision=precision, engine="fused")
    c.restore(ck)
    c.step(56)
    _same(a, ia, pa, c, ic, pc)
    a.close()
    c.close()
