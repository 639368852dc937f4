"""Loader of the CPU oracle (fdtd_oracle.c).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker.  It binds liboracle.so (built by
``make -C oracle``) with the same ctypes prototypes as the HIP library, so a
``fdtd_amd.Simulation`` can run on either.
"""
import ctypes as C
import os
import sys

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(ORACLE_DIR)
# OFDTD_ORACLE_LIB: another build of the same sources (the sanitizer build)
LIB = os.environ.get("OFDTD_ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")
REF_EXE = os.path.join(ORACLE_DIR, "_ref", "ref3d")
sys.path.insert(0, os.path.join(REPO_DIR, "fortran-mpi-fdtd_amd"))

from fdtd_amd import abi  # noqa: E402

_b = None


def binding():
    global _b
    if _b is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB} (run make -C oracle)")
        dll = C.CDLL(LIB)
        _b = abi.Binding(dll, "ofdtd_", extra=False)
        dll.ofdtd_drude2cp_eps.restype = None
        dll.ofdtd_drude2cp_eps.argtypes = [C.POINTER(abi.Drude2CP), C.c_double, abi._PD, abi._PD]
        dll.ofdtd_constants.restype = None
        dll.ofdtd_constants.argtypes = [abi._PD]
    return _b


class complex_division:
    """Context manager selecting how the oracle takes the real part of the
    dispersive update's complex quotient: "modules" ((a*c + b*d) / (c*c +
    d*d), as the reference's modules compiled here evaluate it; the default)
    or "shipped" (C's complex division, as the build behind the reference's
    shipped 2D output evaluates it)."""

    def __init__(self, form):
        self.mode = {"modules": 0, "shipped": 1}[form]

    def __enter__(self):
        dll = binding().dll
        self.prev = C.c_int.in_dll(dll, "ofdtd_cdiv_mode").value
        dll.ofdtd_set_cdiv_mode(C.c_int(self.mode))
        return self

    def __exit__(self, *exc):
        binding().dll.ofdtd_set_cdiv_mode(C.c_int(self.prev))


_b2 = None


def binding2d():
    """The 2D restatement (fdtd2d_oracle.c) under the ofdtd2d_ prefix."""
    global _b2
    if _b2 is None:
        from fdtd_amd import abi2d
        dll = binding().dll
        err = dll.ofdtd2d_last_error
        err.restype = C.c_char_p
        err.argtypes = []
        _b2 = abi2d.Binding2D(dll, "ofdtd2d_", err, batches=False)
        _b2.preset_binding = binding()  # ofdtd_lorentz_jp_preset
    return _b2


def make_simulation2d(spec, precision="fp64"):
    from fdtd_amd.sim2d import make_simulation2d as ms
    return ms(spec, precision=precision, binding=binding2d())


def drude2cp_eps(params, w):
    """Closed-form Drude+2CP permittivity (misc/my_matfunctions_mod.f90:8-31)."""
    from fdtd_amd import _drude2cp
    b = binding()
    re, im = C.c_double(), C.c_double()
    p = _drude2cp(params)
    b.dll.ofdtd_drude2cp_eps(C.byref(p), float(w), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def constants():
    out = (C.c_double * 4)()
    binding().dll.ofdtd_constants(out)
    return tuple(out)


def make_simulation(spec, precision="fp64"):
    from fdtd_amd import make_simulation as ms
    return ms(spec, precision=precision, binding=binding())


# ---- oracle/_ref/ref3d (the reference's own 3D modules) ------------------------
MTYPE = {"dielectric": 1, "drude2cp": 2, "graphene": 3}
# the Lorentz-Jp modules: mtype 4 + FDTD_LJP_* (gold, si, titanium, film gold,
# graphene, graphene3 1nm, non-dispersive graphene); only their own tables
LJP_MTYPE = {"gold": 4, "si": 5, "titanium": 6, "film_gold": 7, "graphene": 8, "graphene3_1nm": 9,
             "graphene_nd": 10}


def ref3d_namelist(spec):
    """The namelist config of oracle/_ref/ref3d (ref_driver.f90) for a
    reference-envelope spec: REF_Y PML, boxes, one y-plane source, y-plane DFTs."""
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
        elif m["type"] == "lorentz_d":
            assert m["params"] == {"preset": "si"}, "mod3/si_mod.f90 has its own table only"
            mt = 12
        elif m["type"] == "dielectric" and m.get("form") == "E":
            mt = 11  # mod3/simpledielectricLorentzJp_mod.f90
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
