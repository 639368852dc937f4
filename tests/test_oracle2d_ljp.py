"""The 2D Lorentz-Jp restatement (fdtd2d_oracle.c; the mod2 LorentzJp
modules, whose pole tables and update are those of the 3D family) against
the 3D oracle at nz = 1, whose Lorentz-Jp update is pinned against the
reference's compiled 3D modules (tests/test_oracle_golden.py)."""
import math

import numpy as np
import pytest


def ljp_case(polarization, preset, nsteps=400):
    from fdtd_amd import sim2d
    from fdtd_amd.workloads import pulse_for_band

    nx, ny = 30, 90
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    om = 2.0 * math.pi * 299792458.0 / np.linspace(400e-9, 900e-9, 8)
    params = {"preset": preset}
    if preset.startswith("graphene"):
        params.update(wNL=2.976e14, dnuNL=3.0e12, xi3=0.05)
    return {"grid": {"nx": nx, "ny": ny, "ds": 2e-9, "courant": 0.5, "periodic_x": True,
                     "polarization": polarization},
            "materials": [{"type": "dielectric", "region": sim2d.box2((1, 1), (nx, ny)), "eps": 1.0},
                          {"type": "lorentz_jp", "region": sim2d.disk((15.5, 50), 6.0), "params": params},
                          {"type": "lorentz_jp", "region": sim2d.box2((1, 20), (nx, 21)), "params": params}],
            "sources": [dict(j=14, **pul)],
            "dft": [{"j": 12, "omega": om}, {"j": 75, "omega": om}],
            "probes": [[15, 50], [15, 20]],
            "nsteps": nsteps}


def as_3d(spec):
    from fdtd_amd import EX, EY, EZ, HX, HY, HZ, box, cylinder

    g = spec["grid"]
    tmz = g["polarization"] == "tmz"
    mats = []
    for m in spec["materials"]:
        r = m["region"]
        if r["kind"] == "box":
            reg = box((r["lo"][0], r["lo"][1], 1), (r["hi"][0], r["hi"][1], 1))
        else:
            reg = cylinder((r["center"][0], r["center"][1], 1.0), r["radius"], 0.5, 2)
        mats.append(dict(m, region=reg))
    return {"grid": {"nx": g["nx"], "ny": g["ny"], "nz": 1, "ds": g["ds"], "courant": g["courant"],
                     "periodic": (1, 0, 1)},
            "materials": mats,
            "sources": [dict(axis=1, index=s["j"], comp=EZ if tmz else EX, t0=s["t0"], st=s["st"], w0=s["w0"])
                        for s in spec["sources"]],
            "dft": [{"axis": 1, "index": d["j"], "omega": d["omega"], "comps": (EX, EZ, HX, HZ)}
                    for d in spec["dft"]],
            "probes": [[i, j, 1] for i, j in spec["probes"]]}


@pytest.mark.parametrize("preset", ["si", "film_gold", "graphene3_1nm"])
def test_tmz_lorentz_jp_equals_the_3d_oracle_bitwise(preset):
    import oracle
    from fdtd_amd import EZ, HX, HY
    from fdtd_amd import sim2d

    spec = ljp_case("tmz", preset)
    s2, i2, _ = oracle.make_simulation2d(spec)
    s3, i3, _ = oracle.make_simulation(as_3d(spec))
    s2.step(spec["nsteps"])
    s3.step(spec["nsteps"])
    assert np.abs(s2.field(sim2d.EZ2)).max() > 1e-3
    for c2, c3 in ((sim2d.EZ2, EZ), (sim2d.HX2, HX), (sim2d.HY2, HY)):
        np.testing.assert_array_equal(s2.field(c2), s3.field(c3)[:, :, 0])
    for a, b in zip(i2, i3):
        np.testing.assert_array_equal(s2.flux(a), s3.flux(b))
    s2.close()
    s3.close()


@pytest.mark.parametrize("preset", ["gold", "si", "film_gold", "graphene3_1nm"])
def test_tez_lorentz_jp_equals_the_3d_oracle(preset):
    """TEz: the two restatements sum the curl of E for Hz in another order,
    so the bar is rounding (1e-12 normwise)."""
    import oracle
    from fdtd_amd import EX, EY, HZ
    from fdtd_amd import sim2d

    spec = ljp_case("tez", preset)
    s2, i2, _ = oracle.make_simulation2d(spec)
    s3, i3, _ = oracle.make_simulation(as_3d(spec))
    s2.step(spec["nsteps"])
    s3.step(spec["nsteps"])
    assert np.abs(s2.field(sim2d.EX2)).max() > 1e-3
    for c2, c3 in ((sim2d.EX2, EX), (sim2d.EY2, EY), (sim2d.HZ2, HZ)):
        b = s3.field(c3)[:, :, 0]
        assert np.abs(s2.field(c2) - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-30), (preset, c2)
    for a, b in zip(i2, i3):
        f3 = s3.flux(b)
        assert np.abs(s2.flux(a) - f3).max() <= 1e-12 * np.abs(f3).max()
    s2.close()
    s3.close()
