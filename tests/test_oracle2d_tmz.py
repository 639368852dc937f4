"""The 2D TMz restatement of the oracle (fdtd2d_oracle.c) against the 3D
oracle run with nz = 1 and periodic z: the 3D FUSED stepping with no z
variation carries exactly Ez, Hx, Hy, so every TMz field, DFT accumulator,
flux and probe value must agree bit for bit.  The 3D oracle is itself pinned
against the reference's compiled 3D modules and closed forms; a sign, index
or ordering slip in the TMz restatement shows up here as a mismatch."""
import numpy as np
import pytest


def as_3d(spec):
    """The z-invariant 3D equivalent of a TMz spec (nz = 1, periodic z)."""
    from fdtd_amd import EZ, EX, HX, HZ, box, cylinder

    g = spec["grid"]
    P = spec["pml"]["npml"]
    mats = []
    for m in spec["materials"]:
        r = m["region"]
        if r["kind"] == "box":
            reg = box((r["lo"][0], r["lo"][1], 1), (r["hi"][0], r["hi"][1], 1))
        else:
            reg = cylinder((r["center"][0], r["center"][1], 1.0), r["radius"], 0.5, 2)
        mats.append(dict(m, region=reg))
    return {
        "grid": {"nx": g["nx"], "ny": g["ny"], "nz": 1, "ds": g["ds"], "courant": g["courant"],
                 "periodic": (1 if g["periodic_x"] else 0, 0, 1)},
        "pml": {"mode": "all_faces", "npml": [P[0], P[1], 0], "order": spec["pml"]["order"],
                "R0": spec["pml"]["R0"]},
        "materials": mats,
        "sources": [dict(axis=1, index=s["j"], comp=EZ, t0=s["t0"], st=s["st"], w0=s["w0"])
                    for s in spec["sources"]],
        "dft": [{"axis": 1, "index": d["j"], "omega": d["omega"], "comps": (EX, EZ, HX, HZ),
                 "n_begin": d.get("n_begin", 1), "n_end": d.get("n_end", 0)} for d in spec["dft"]],
        "probes": [[i, j, 1] for i, j in spec["probes"]],
    }


@pytest.mark.parametrize("periodic_x,graphene", [(False, "cp4"), (True, "kubo"), (True, "cp4")])
def test_tmz_restatement_equals_the_3d_oracle_at_nz1(periodic_x, graphene):
    import oracle
    from fdtd_amd import EZ, EX, EY, HX, HY, HZ, DZ
    from fdtd_amd import sim2d

    spec = sim2d.tmz_case(periodic_x=periodic_x, graphene=graphene, nsteps=260)
    s2, ids2, pids2 = oracle.make_simulation2d(spec)
    s3, ids3, pids3 = oracle.make_simulation(as_3d(spec))
    s2.step(spec["nsteps"])
    s3.step(spec["nsteps"])
    ez = s2.field(sim2d.EZ2)
    assert np.max(np.abs(ez)) > 1e-3  # the pulse is on the grid
    for c2, c3 in ((sim2d.EZ2, EZ), (sim2d.HX2, HX), (sim2d.HY2, HY), (sim2d.DZ2, DZ)):
        np.testing.assert_array_equal(s2.field(c2), s3.field(c3)[:, :, 0])
    for c3 in (EX, EY, HZ):  # the TEz set stays zero in 3D
        assert not np.any(s3.field(c3))
    for a, b in zip(ids2, ids3):
        for c2, c3 in ((sim2d.EZ2, EZ), (sim2d.HX2, HX)):
            r2, i2 = s2.dft(a, c2)
            r3, i3 = s3.dft(b, c3)
            np.testing.assert_array_equal(r2, r3[:, 0, :])
            np.testing.assert_array_equal(i2, i3[:, 0, :])
        np.testing.assert_array_equal(s2.flux(a), s3.flux(b))
    for a, b in zip(pids2, pids3):
        p2, p3 = s2.probe(a), s3.probe(b)
        np.testing.assert_array_equal(p2, p3[:, [0, 3, 4]])
    s2.close()
    s3.close()


def test_tmz_rejects_what_it_does_not_model():
    import oracle
    from fdtd_amd import abi
    from fdtd_amd import sim2d

    spec = sim2d.tmz_case(periodic_x=True)
    with pytest.raises(abi.FdtdError):  # x PML on a periodic axis
        sim = sim2d.Simulation2D(40, 48, 2e-9, 0.6, periodic_x=True, binding=oracle.binding2d(),
                                 polarization="tmz")
        sim.set_pml((4, 8))
    sim = sim2d.Simulation2D(40, 48, 2e-9, 0.6, periodic_x=True, binding=oracle.binding2d(), polarization="tmz")
    sim.set_pml((0, 8))
    sim.add_drude2cp(sim2d.disk((20, 6), 4.0))  # reaches into the y PML
    with pytest.raises(abi.FdtdError):
        sim.commit()
    sim.close()
    sim = sim2d.Simulation2D(40, 48, 2e-9, 0.6, binding=oracle.binding2d(), polarization="tmz")
    with pytest.raises(abi.FdtdError):  # a TEz component on a TMz line
        sim.add_dft_line(10, np.array([1e15]), sim2d.DFT_EX)
    del spec
