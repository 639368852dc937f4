"""Generate the golden vectors from the reference's own Fortran modules.

Runs oracle/_ref/ref3d (built by ``make -C oracle ref`` from the reference
module sources, in this development container only) on the reference-envelope
workloads of fdtd_amd.workloads and stores small .npz fixtures next to this
script: the workload spec (JSON), the DFT frequency grid, final fields (or
slices of them for the 96^3 case), DFT accumulators (or samples), flux
spectra and probe traces.  The reference never travels; only these vectors do.

Usage: python tests/golden/gen_golden.py [case ...]
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "fortran-mpi-fdtd_amd"))
from fdtd_amd import workloads  # noqa: E402

REF_EXE = os.path.join(REPO, "oracle", "_ref", "ref3d")
MTYPE = {"dielectric": 1, "drude2cp": 2, "graphene": 3}
# the Lorentz-Jp modules: mtype 4 + FDTD_LJP_* (gold, si, titanium, film gold,
# graphene, graphene3 1nm, non-dispersive graphene); only their own tables
LJP_MTYPE = {"gold": 4, "si": 5, "titanium": 6, "film_gold": 7, "graphene": 8, "graphene3_1nm": 9,
             "graphene_nd": 10}


def namelist(spec):
    g, p = spec["grid"], spec["pml"]
    mats = spec["materials"]
    lines = ["&grid",
             f" nx={g['nx']}, ny={g['ny']}, nz={g['nz']}, ds={g['ds']!r}, Ku={g['courant']!r},",
             f" nsteps={spec['nsteps']}, pbcx={'.true.' if g['periodic'][0] else '.false.'},"
             f" pbcz={'.true.' if g['periodic'][2] else '.false.'}", "/",
             "&pml",
             f" npmlx={p['npml'][0]}, npmly={p['npml'][1]}, npmlz={p['npml'][2]}, order={p['order']!r},"
             f" R0={p['R0']!r}, eps_low={p['eps_low']!r}, eps_high={p['eps_high']!r}", "/",
             "&mats", f" nmat={len(mats)},"]
    for q, m in enumerate(mats, start=1):
        r = m["region"]
        assert r["kind"] == "box"
        b = [r["lo"][0], r["hi"][0], r["lo"][1], r["hi"][1], r["lo"][2], r["hi"][2]]
        if m["type"] == "lorentz_jp":
            lp = m["params"]
            assert set(lp) <= {"preset", "wNL", "dnuNL", "xi3"}, "the modules take only their own tables"
            mt = LJP_MTYPE[lp["preset"]]
            if mt >= 8:
                lines.append(f" gwNL({q})={lp['wNL']!r}, gdnuNL({q})={lp['dnuNL']!r}, gxi3({q})={lp['xi3']!r},")
        else:
            mt = MTYPE[m["type"]]
        lines.append(f" mtype({q})={mt}, mbox(1:6,{q})={','.join(str(v) for v in b)},")
        if m["type"] == "dielectric":
            lines.append(f" meps({q})={m['eps']!r},")
        if m["type"] == "graphene":
            gp = m["params"]
            assert gp["model"] == "cp4" and gp["ref_init_order"] == 1 and gp["normal_axis"] == 1
            lines.append(f" gwNL({q})={gp['wNL']!r}, gdnuNL({q})={gp['dnuNL']!r}, gxi3({q})={gp['xi3']!r},")
    lines.append("/")
    s = spec["sources"][0]
    assert s["axis"] == 1
    lines += ["&src", f" s_on=.true., s_index={s['index']}, s_comp={s['comp'] + 1}, s_t0={s['t0']!r},"
              f" s_st={s['st']!r}, s_w0={s['w0']!r}, s_amp={s['amplitude']!r}", "/"]
    d = spec["dft"]
    assert all(x["axis"] == 1 for x in d)
    lines += ["&dft", f" ndft={len(d)}, nfreq={len(d[0]['omega'])},"
              f" d_index(1:{len(d)})={','.join(str(x['index']) for x in d)},"
              f" d_nb(1:{len(d)})={','.join(str(x['n_begin']) for x in d)},"
              f" d_ne(1:{len(d)})={','.join(str(x['n_end']) for x in d)},"
              f" lam_init={spec['lam'][0]!r}, lam_end={spec['lam'][1]!r}", "/"]
    pr = spec["probes"]
    lines += ["&probes", f" nprobe={len(pr)},"]
    for q, (i, j, k) in enumerate(pr, start=1):
        lines.append(f" p_ijk(1:3,{q})={i},{j},{k},")
    lines.append("/")
    return "\n".join(lines).replace("e-", "d-").replace("e+", "d+") + "\n"


def run_reference(spec):
    g = spec["grid"]
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    nd = len(spec["dft"])
    nf = len(spec["dft"][0]["omega"])
    npr = len(spec["probes"])
    with tempfile.TemporaryDirectory() as tmp:
        cfg = os.path.join(tmp, "cfg.nml")
        with open(cfg, "w") as f:
            f.write(namelist(spec))
        res = subprocess.run([REF_EXE, cfg, tmp], capture_output=True, text=True, timeout=3600)
        if res.returncode != 0:
            raise RuntimeError(res.stdout + res.stderr)
        fl = np.fromfile(os.path.join(tmp, "fields.bin"), dtype="<f8")
        dft = np.fromfile(os.path.join(tmp, "dft.bin"), dtype="<f8")
        prb = np.fromfile(os.path.join(tmp, "probes.bin"), dtype="<f8")
    ncell = nx * ny * nz
    fields = fl.reshape(9, ncell)
    arg = dft[:nf]
    off = nf
    acc = {}
    for p in range(nd):
        for c, name in enumerate(("Ex", "Ez", "Hx", "Hz")):
            n1 = nx * nz * nf
            acc[(p, name, "re")] = dft[off:off + n1]
            acc[(p, name, "im")] = dft[off + n1:off + 2 * n1]
            off += 2 * n1
    sy = dft[off:off + nf * nd].reshape(nd, nf)
    prb = prb.reshape(npr, spec["nsteps"], 7)
    return fields, arg, acc, sy, prb, res.stdout


def save(spec, full_fields=True):
    fields, arg, acc, sy, prb, log = run_reference(spec)
    g = spec["grid"]
    nx, ny, nz = g["nx"], g["ny"], g["nz"]
    out = {"spec": np.frombuffer(json.dumps(spec).encode(), dtype=np.uint8), "arg": arg, "flux": sy, "probes": prb}
    names = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dx", "Dy", "Dz")
    for q, nm in enumerate(names):
        f3 = fields[q].reshape(nz, ny, nx)  # [k][j][i]
        if full_fields:
            out["f_" + nm] = f3
        else:
            out["fy_" + nm] = f3[:, ny // 2, :]          # y mid-plane [k][i]
            out["fz_" + nm] = f3[nz // 2, :, :]          # z mid-plane [j][i]
    nf = len(arg)
    for (p, nm, part), v in acc.items():
        a = v.reshape(nf, nz, nx)  # [m][k][i]
        out[f"dft{p}_{nm}_{part}"] = a if full_fields else a[:, ::8, ::8]
    path = os.path.join(HERE, spec["name"] + ".npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes;", log.strip().splitlines()[-1])


CASES = {
    "ref_box_small": (workloads.ref_box_small, True),
    "ref_graphene_small": (workloads.ref_graphene_small, True),
    "ref_c1_twin": (workloads.ref_c1_twin, False),
    "ref_ljp_metals": (workloads.ref_ljp_metals, True),
    "ref_ljp_graphene": (workloads.ref_ljp_graphene, True),
    "ref_ljp_titanium": (lambda: workloads.ref_ljp_metals(nsteps=300, titanium=True), True),
}

if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for nm in names:
        fn, full = CASES[nm]
        save(fn(), full_fields=full)
