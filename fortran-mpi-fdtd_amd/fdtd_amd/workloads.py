"""Workload descriptions (JSON-style specs consumed by ``apply_spec``).

Reference-envelope cases follow the shape of the shipped 3D driver
(sim3/program.f90:44-156,230-287): y-slab PML, periodic x/z, boxes applied
in update order, a soft Dx plane source on a y plane, time-gated DFT planes.
The 96^3 / 512^3 / 384x384x256 cases are the gold-sphere, gold-dimer and
graphene-over-disks workloads; random draws use one
``numpy.random.default_rng(seed)`` in the order listed.
"""
import math

import numpy as np

from . import C0, NANO, PI, box, cylinder, gold_params_dict, reference_frequencies, sphere, wavelength_frequencies
from .abi import EX, EY, EZ, HX, HY, HZ

COURANT_3D = 0.99 / math.sqrt(3.0)


def pulse_for_band(lam_lo, lam_hi, lam_c):
    """Gaussian-modulated sine centred at lam_c whose 1/e spectral half-width
    reaches the nearer band edge: st = 1/dw, t0 = 4 st."""
    wc = 2.0 * PI * C0 / lam_c
    dw = min(2.0 * PI * C0 / lam_lo - wc, wc - 2.0 * PI * C0 / lam_hi)
    st = 1.0 / dw
    return {"t0": 4.0 * st, "st": st, "w0": wc}


def _vacuum(n):
    return {"type": "dielectric", "region": box((1, 1, 1), n), "eps": 1.0}


# ---------------------------------------------------------------------------
# reference-envelope cases (REF_Y PML, STAGED engine, Fortran-checkable)
def ref_box_small(nsteps=1200):
    """A scaled-down sim3 layout: air / gold antenna / SiO2 / Si along y,
    y-PML 10 with x/z grading depth 3, periodic x/z."""
    nx, ny, nz = 12, 72, 8
    ds = 3.999 * NANO
    P = 10
    air_hi, au_lo = 50, 46
    lam0 = 633.0 * NANO
    w0 = 2.0 * PI * C0 / lam0
    st = 0.5e-15
    tfin = 420
    sposy = P + 2
    ptiny = sposy + 2
    ptty = 59
    arg = reference_frequencies(500e-9, 800e-9, 20)
    return {
        "name": "ref_box_small",
        "grid": {"nx": nx, "ny": ny, "nz": nz, "ds": ds, "courant": 0.5, "periodic": [1, 0, 1]},
        "pml": {"mode": "ref_y", "npml": [3, P, 3], "order": 3.0, "R0": math.exp(-16.0),
                "eps_low": 1.0, "eps_high": 3.88 ** 2},
        "materials": [
            {"type": "dielectric", "region": box((1, P + 1, 1), (nx, air_hi, nz)), "eps": 1.0},
            {"type": "drude2cp", "region": box((4, au_lo, 1), (9, air_hi, nz)), "params": gold_params_dict()},
            {"type": "dielectric", "region": box((1, air_hi + 1, 1), (nx, 56, nz)), "eps": 1.46 ** 2},
            {"type": "dielectric", "region": box((1, 57, 1), (nx, ny - P, nz)), "eps": 3.88 ** 2},
        ],
        "sources": [{"axis": 1, "index": sposy, "comp": EX, "t0": 4.0 * st, "st": st, "w0": w0, "amplitude": 1.0}],
        "dft": [
            {"axis": 1, "index": ptiny, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": tfin - 1},
            {"axis": 1, "index": ptiny, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": tfin + 1, "n_end": 0},
            {"axis": 1, "index": ptty, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": 0},
        ],
        "probes": [[6, ptiny, 4], [6, ptty, 4], [6, 48, 4]],
        "nsteps": nsteps,
        "lam": [500e-9, 800e-9],
    }


def ref_si_small(nsteps=1200):
    """ref_box_small with its Si substrate, and a block in the air above the
    gold, as the D-form Lorentz silicon of mod3/si_mod.f90 (three poles, the
    first undamped) instead of the constant eps 3.88^2."""
    spec = ref_box_small(nsteps=nsteps)
    spec["name"] = "ref_si_small"
    nx, ny, nz = spec["grid"]["nx"], spec["grid"]["ny"], spec["grid"]["nz"]
    P = spec["pml"]["npml"][1]
    mats = spec["materials"]
    mats[3] = {"type": "lorentz_d", "region": box((1, 57, 1), (nx, ny - P, nz)), "params": {"preset": "si"}}
    mats.insert(1, {"type": "lorentz_d", "region": box((4, 30, 1), (9, 34, nz)), "params": {"preset": "si"}})
    return spec


def ref_graphene_small(nsteps=900):
    """A graphene sheet (one cell thick, normal y) with the Raman term over
    a gold strip, y-PML, periodic x/z; the reference's CP4 parameters."""
    nx, ny, nz = 10, 60, 6
    ds = 2.0 * NANO
    P = 10
    lam0 = 600.0 * NANO
    st = 0.4e-15
    gr = {"model": "cp4", "normal_axis": 1, "ref_init_order": 1, "thickness": 3.4e-10, "eps_inf": 1.0,
          "gamma": 9e15, "omega_d": 5.2796e15, "A": [13.1068, 1.8417, 0.4191, 0.7737],
          "phi": [-1.8391, 0.3505, 3.0565, -0.1199], "Omega": [6.1264e14, 6.6843e15, 5.6289e15, 1.118e16],
          "Gamma": [1e14, 7.905e14, 6.7526e14, 1.0005e14],
          "wNL": 2.976e14, "dnuNL": 3.0e12, "xi3": 0.05}
    arg = reference_frequencies(450e-9, 900e-9, 12)
    return {
        "name": "ref_graphene_small",
        "grid": {"nx": nx, "ny": ny, "nz": nz, "ds": ds, "courant": 0.5, "periodic": [1, 0, 1]},
        "pml": {"mode": "ref_y", "npml": [2, P, 2], "order": 3.0, "R0": math.exp(-16.0),
                "eps_low": 1.0, "eps_high": 2.1316},
        "materials": [
            {"type": "dielectric", "region": box((1, P + 1, 1), (nx, 40, nz)), "eps": 1.0},
            {"type": "graphene", "region": box((1, 30, 1), (nx, 30, nz)), "params": gr},
            {"type": "drude2cp", "region": box((3, 36, 1), (8, 40, nz)), "params": gold_params_dict()},
            {"type": "dielectric", "region": box((1, 41, 1), (nx, ny - P, nz)), "eps": 2.1316},
        ],
        "sources": [{"axis": 1, "index": P + 2, "comp": EX, "t0": 4.0 * st, "st": st,
                     "w0": 2.0 * PI * C0 / lam0, "amplitude": 1.0}],
        "dft": [
            {"axis": 1, "index": 16, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": 0},
            {"axis": 1, "index": 45, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": 0},
        ],
        "probes": [[5, 29, 3], [5, 30, 3], [5, 45, 3]],
        "nsteps": nsteps,
        "lam": [450e-9, 900e-9],
    }


def ref_ljp_metals(nsteps=1000, titanium=False, eform=False):
    """Lorentz-Jp metals (gold antenna, Si and film-gold slabs) between air
    and SiO2 along y, y-PML 10 with x/z grading depth 2, periodic x/z.
    Every cell has exactly one writer (no reference driver calls
    update_EJp, so the case keeps E-form and D-form regions apart).
    titanium=True puts the titanium table (mod3/titaniumJp_mod.f90, negative
    damping on pole 1: the recursion grows without bound) in row 50.
    eform=True makes every dielectric but the rows around the source (which
    adds to D) the E-form one written for these simulations
    (mod3/simpledielectricLorentzJp_mod.f90, "form": "E"): the same material
    to the library, E += Ku*ga*curl H in the reference."""
    nx, ny, nz = 8, 72, 6
    ds = 3.999 * NANO
    P = 10
    lam0 = 633.0 * NANO
    st = 0.5e-15
    arg = reference_frequencies(500e-9, 800e-9, 16)
    ljp = lambda name: {"preset": name}  # noqa: E731
    mats = [
        {"type": "dielectric", "region": box((1, P + 1, 1), (nx, 40, nz)), "eps": 1.0},
        {"type": "dielectric", "region": box((1, 41, 1), (1, 43, nz)), "eps": 1.0},
        {"type": "lorentz_jp", "region": box((2, 41, 1), (nx - 1, 43, nz)), "params": ljp("gold")},
        {"type": "dielectric", "region": box((nx, 41, 1), (nx, 43, nz)), "eps": 1.0},
        {"type": "dielectric", "region": box((1, 44, 1), (nx, 45, nz)), "eps": 1.0},
        {"type": "lorentz_jp", "region": box((1, 46, 1), (nx, 48, nz)), "params": ljp("si")},
        {"type": "dielectric", "region": box((1, 49, 1), (nx, 49, nz)), "eps": 1.0},
        {"type": "lorentz_jp", "region": box((1, 50, 1), (nx, 50, nz)), "params": ljp("titanium")} if titanium
        else {"type": "dielectric", "region": box((1, 50, 1), (nx, 50, nz)), "eps": 1.0},
        {"type": "lorentz_jp", "region": box((1, 51, 1), (nx, 52, nz)), "params": ljp("film_gold")},
        {"type": "dielectric", "region": box((1, 53, 1), (nx, ny - P, nz)), "eps": 1.46 ** 2},
    ]
    if eform:
        # the source adds to D (sim3/program.f90:396-406), which E-form cells
        # never read: the rows around it stay D-form
        mats[0] = {"type": "dielectric", "region": box((1, P + 4, 1), (nx, 40, nz)), "eps": 1.0}
        for m in mats:
            if m["type"] == "dielectric":
                m["form"] = "E"
        mats.insert(0, {"type": "dielectric", "region": box((1, P + 1, 1), (nx, P + 3, nz)), "eps": 1.0})
    return {
        "name": "ref_ljp_titanium" if titanium else ("ref_ljp_eform" if eform else "ref_ljp_metals"),
        "grid": {"nx": nx, "ny": ny, "nz": nz, "ds": ds, "courant": 0.5, "periodic": [1, 0, 1]},
        "pml": {"mode": "ref_y", "npml": [2, P, 2], "order": 3.0, "R0": math.exp(-16.0),
                "eps_low": 1.0, "eps_high": 1.46 ** 2},
        "materials": mats,
        "sources": [{"axis": 1, "index": P + 2, "comp": EX, "t0": 4.0 * st, "st": st,
                     "w0": 2.0 * PI * C0 / lam0, "amplitude": 1.0}],
        "dft": [
            {"axis": 1, "index": 16, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": 0},
            {"axis": 1, "index": 56, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": 0},
        ],
        "probes": [[4, 42, 3], [4, 47, 3], [4, 51, 3], [4, 56, 3]],
        "nsteps": nsteps,
        "lam": [500e-9, 800e-9],
    }


def ref_ljp_graphene(nsteps=900):
    """The three Lorentz-Jp graphene sheets (one cell thick, normal y) with
    their third-order oscillator, in air over SiO2, y-PML, periodic x/z; the
    4-pole sheet drives Ez with JpT_x as mod3/grapheneLorentzJp_mod.f90:68
    does, which couples the Ex-polarised pulse into Ez."""
    nx, ny, nz = 6, 60, 6
    ds = 2.0 * NANO
    P = 10
    lam0 = 600.0 * NANO
    st = 0.4e-15
    arg = reference_frequencies(450e-9, 900e-9, 12)
    nl = {"wNL": 2.976e14, "dnuNL": 3.0e12, "xi3": 2e-7}
    mats = [
        {"type": "dielectric", "region": box((1, P + 1, 1), (nx, 29, nz)), "eps": 1.0},
        {"type": "lorentz_jp", "region": box((1, 30, 1), (nx, 30, nz)), "params": {"preset": "graphene", **nl}},
        {"type": "dielectric", "region": box((1, 31, 1), (nx, 32, nz)), "eps": 1.0},
        {"type": "lorentz_jp", "region": box((1, 33, 1), (nx, 33, nz)),
         "params": {"preset": "graphene3_1nm", **nl}},
        {"type": "dielectric", "region": box((1, 34, 1), (nx, 35, nz)), "eps": 1.0},
        {"type": "lorentz_jp", "region": box((1, 36, 1), (nx, 36, nz)), "params": {"preset": "graphene_nd", **nl}},
        {"type": "dielectric", "region": box((1, 37, 1), (nx, 40, nz)), "eps": 1.0},
        {"type": "dielectric", "region": box((1, 41, 1), (nx, ny - P, nz)), "eps": 2.1316},
    ]
    return {
        "name": "ref_ljp_graphene",
        "grid": {"nx": nx, "ny": ny, "nz": nz, "ds": ds, "courant": 0.5, "periodic": [1, 0, 1]},
        "pml": {"mode": "ref_y", "npml": [2, P, 2], "order": 3.0, "R0": math.exp(-16.0),
                "eps_low": 1.0, "eps_high": 2.1316},
        "materials": mats,
        "sources": [{"axis": 1, "index": P + 2, "comp": EX, "t0": 4.0 * st, "st": st,
                     "w0": 2.0 * PI * C0 / lam0, "amplitude": 1.0}],
        "dft": [
            {"axis": 1, "index": 16, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": 0},
            {"axis": 1, "index": 45, "omega": list(arg), "comps": [EX, EZ, HX, HZ], "n_begin": 1, "n_end": 0},
        ],
        "probes": [[3, 30, 3], [3, 33, 3], [3, 36, 3], [3, 45, 3]],
        "nsteps": nsteps,
        "lam": [450e-9, 900e-9],
    }


def sim1_tfsf_shipped(nsteps=None):
    """The reference's 1D TFSF run (sim1/program.f90:18-216) as a 3D grid of
    1 x npos x 1 cells, periodic in x and z: ds = 1 nm, Ku = 0.99, air /
    SiO2 / Si along y, a one-sided TFSF boundary at TFSF_bd = 25 fed by an
    incident line in air with its own PML (:134, the line's PML is the 1D
    module's) and the pulse at spos = 20 (beta = 2), the reflection probe at
    monitor_ptr = 30 and the incident-line probe at monitor_pti = 20.
    Parameters computed as the program computes them, integer truncations
    included.  The main grid's y PML (10 cells, air below, Si above,
    :133) is the ALL_FACES split PML over media that fill it: the 3D
    reference slab PML keeps its split ghosts zero, which in a one-cell-wide
    periodic grid couples the transverse components, and Ku = 0.99 is above
    the 3D limit, so that form does not reduce to the 1D run."""
    lam0 = 633.0 * NANO
    f0 = C0 / lam0
    ds = 1.0 * NANO
    Ku = 0.99
    dt = Ku * ds / C0
    npos = math.floor((20.0 * lam0 + 344.0 * NANO) / ds)
    P = 10
    jc = int(0.5 * npos)
    st = 2.0 * (1.0 / f0)
    t0 = 4.0 * st
    spos = P + math.floor(10.0 * NANO / ds)
    bd = spos + math.floor(5.0 * NANO / ds)
    pti, ptr = bd - math.floor(5.0 * NANO / ds), bd + math.floor(5.0 * NANO / ds)
    tfin = math.floor((ds * float(jc)) / (C0 / 1.0) / dt)
    Nt = 6 * tfin + 3 * math.floor(t0 / dt)
    air = (P + 1, P + 1 + math.floor(20.0 * lam0 / ds) - 1)
    sio2 = (air[1] + 1, air[1] + 1 + math.floor(284.0 * NANO / ds) - 1)
    si = (sio2[1] + 1, npos - P)
    return {
        "name": "sim1_tfsf_shipped",
        "grid": {"nx": 1, "ny": npos, "nz": 1, "ds": ds, "courant": Ku, "periodic": [1, 0, 1]},
        "pml": {"mode": "all_faces", "npml": [0, P, 0], "order": 3.0, "R0": math.exp(-16.0)},
        "materials": [
            {"type": "dielectric", "region": box((1, 1, 1), (1, air[1], 1)), "eps": 1.0},
            {"type": "dielectric", "region": box((1, sio2[0], 1), (1, sio2[1], 1)), "eps": 1.46 ** 2},
            {"type": "dielectric", "region": box((1, si[0], 1), (1, npos, 1)), "eps": 3.772 ** 2},
        ],
        "tfsf": [{"axis": 1, "comp": EX, "lo": bd, "hi": 0, "src_index": spos, "npml": P, "eps": 1.0,
                  "order": 3.0, "R0": math.exp(-16.0), "t0": t0, "st": st, "w0": 2.0 * PI * f0, "amplitude": 1.0}],
        "probes": [[1, ptr, 1]],
        "tfsf_probes": [[0, pti]],
        "freq": list(reference_frequencies(200e-9, 1000e-9, 100)),
        "tfinStop": tfin,
        "nsteps": Nt if nsteps is None else nsteps,
    }


def running_spectrum(trace, omega, dt, ds, e_col, h_col):
    """sim1/program.f90:270-326 from a probe trace: re/im += cos/sin(w t)*F*dt
    with t = (n-1)*dt, then S = (-(reH*reE) - (imH*imE))*ds*z0."""
    z0 = math.sqrt((PI * 4e-7) / 8.85418781762039e-12)  # misc/constants_mod.f90
    omega = np.asarray(omega)
    n = trace.shape[0]
    t = np.arange(n, dtype=np.float64) * dt
    re_e = np.zeros(len(omega)); im_e = np.zeros(len(omega))
    re_h = np.zeros(len(omega)); im_h = np.zeros(len(omega))
    for q in range(n):
        c, s_ = np.cos(omega * t[q]), np.sin(omega * t[q])
        e, h = trace[q, e_col], trace[q, h_col]
        re_e = re_e + c * e * dt; im_e = im_e + s_ * e * dt
        re_h = re_h + c * h * dt; im_h = im_h + s_ * h * dt
    return (-(re_h * re_e) - (im_h * im_e)) * ds * z0


def ref_c1_twin(nsteps=2000, n=96):
    """Reference-envelope twin of the 96^3 gold-sphere case: y-PML only,
    periodic x/z, a gold cube of the sphere's volume, Dx source on a y plane,
    a transmission DFT plane (16 frequencies, reference spacing).  `n`
    resizes the cube grid (same cell size, PML and particle)."""
    ds = 2.0 * NANO
    P = 10
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    arg = reference_frequencies(400e-9, 900e-9, 16)
    side = int(round((4.0 / 3.0 * PI * 10 ** 3) ** (1.0 / 3.0)))  # 16 cells
    lo = (n - side) // 2 + 1
    return {
        "name": "ref_c1_twin" if n == 96 else f"ref_c1_twin_{n}",
        "grid": {"nx": n, "ny": n, "nz": n, "ds": ds, "courant": COURANT_3D, "periodic": [1, 0, 1]},
        "pml": {"mode": "ref_y", "npml": [P, P, P], "order": 3.0, "R0": 1e-6, "eps_low": 1.0, "eps_high": 1.0},
        "materials": [
            {"type": "dielectric", "region": box((1, P + 1, 1), (n, n - P, n)), "eps": 1.0},
            {"type": "drude2cp", "region": box((lo, lo, lo), (lo + side - 1,) * 3), "params": gold_params_dict()},
        ],
        "sources": [dict(axis=1, index=P + 2, comp=EX, amplitude=1.0, **pul)],
        "dft": [{"axis": 1, "index": n - P - 2, "omega": list(arg), "comps": [EX, EZ, HX, HZ],
                 "n_begin": 1, "n_end": 0}],
        "probes": [[n // 2, n - P - 2, n // 2], [n // 2, lo - 2, n // 2]],
        "nsteps": nsteps,
        "lam": [400e-9, 900e-9],
    }


def sim3_shipped(nsteps=None):
    """The shipped 3D case of sim3/program.f90:44-156,230-287 on one domain:
    125 x 1266 x 25 cells of 3.999 nm, Ku = 0.5, y-PML 10 (x/z grading 10),
    air / Au antenna / SiO2 / Si, 633 nm source, 100 frequencies over
    500-800 nm on the incident, reflected and transmitted planes.  The same
    parameters as fortran/host3d.f90."""
    ds = 3.999 * NANO
    nx, ny, nz = math.floor(500e-9 / ds), math.floor(5064e-9 / ds), math.floor(25 * ds / ds)
    ic, kc = nx // 2, nz // 2
    P = 10
    lam0 = 633e-9
    f0 = C0 / lam0
    st = 1.0 / f0
    air = (P + 1, int(0.8 * ny))
    au_x = (int(ic - 0.5 * math.floor(100e-9 / ds) + 1), int(ic + 0.5 * math.floor(100e-9 / ds)))
    au_y = (air[1] - math.floor(20e-9 / ds) + 1, air[1])
    sio2 = (air[1] + 1, air[1] + 1 + math.floor(284e-9 / ds) - 1)
    si = (sio2[1] + 1, ny - P)
    ptty = int(0.5 * (si[1] + si[0]))
    sposy = P + 2
    ptiny = sposy + 2
    dt = 0.5 * ds / C0
    tfin = math.floor(((ds * (air[1] - sposy)) + 1.0 * 4.5 * lam0 / 1.0) / (C0 / 1.0) / dt)
    arg = list(reference_frequencies(500e-9, 800e-9, 100))
    comps = [EX, EZ, HX, HZ]
    return {
        "name": "sim3_shipped",
        "tfinStop": tfin,
        "grid": {"nx": nx, "ny": ny, "nz": nz, "ds": ds, "courant": 0.5, "periodic": [1, 0, 1]},
        "pml": {"mode": "ref_y", "npml": [P, P, P], "order": 3.0, "R0": math.exp(-16.0),
                "eps_low": 1.0, "eps_high": 3.88 ** 2},
        "materials": [
            {"type": "dielectric", "region": box((1, air[0], 1), (nx, air[1], nz)), "eps": 1.0},
            {"type": "drude2cp", "region": box((au_x[0], au_y[0], 1), (au_x[1], au_y[1], nz)),
             "params": gold_params_dict()},
            {"type": "dielectric", "region": box((1, sio2[0], 1), (nx, sio2[1], nz)), "eps": 1.46 ** 2},
            {"type": "dielectric", "region": box((1, si[0], 1), (nx, si[1], nz)), "eps": 3.88 ** 2},
        ],
        "sources": [{"axis": 1, "index": sposy, "comp": EX, "t0": 4.0 * st, "st": st, "w0": 2.0 * PI * f0,
                     "amplitude": 1.0}],
        "dft": [{"axis": 1, "index": ptiny, "omega": arg, "comps": comps, "n_begin": 1, "n_end": tfin - 1},
                {"axis": 1, "index": ptiny, "omega": arg, "comps": comps, "n_begin": tfin + 1, "n_end": 0},
                {"axis": 1, "index": ptty, "omega": arg, "comps": comps, "n_begin": 1, "n_end": 0}],
        "probes": [[ic, ptiny, kc], [ic, ptty, kc]],
        "nsteps": 4 * tfin if nsteps is None else nsteps,
    }


# ---------------------------------------------------------------------------
# all-faces PML workloads (FUSED engine)
def c1_gold_sphere(nsteps=2000, n=96):
    """96^3, ds = 2 nm, 10-cell PML on all faces (m = 3, R = 1e-6), vacuum
    with a centred r = 20 nm gold sphere, x-polarised Gaussian plane source
    propagating +z at z = 12, DFT on one z plane at 16 wavelengths spaced
    evenly over 400-900 nm."""
    ds = 2.0 * NANO
    P = 10
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    om = wavelength_frequencies(400e-9, 900e-9, 16)
    c = (n + 1) / 2.0
    return {
        "name": f"gold_sphere_{n}",
        "grid": {"nx": n, "ny": n, "nz": n, "ds": ds, "courant": COURANT_3D, "periodic": [0, 0, 0]},
        "pml": {"mode": "all_faces", "npml": [P, P, P], "order": 3.0, "R0": 1e-6},
        "materials": [_vacuum((n, n, n)),
                      {"type": "drude2cp", "region": sphere((c, c, c), 20e-9 / ds), "params": gold_params_dict()}],
        "sources": [dict(axis=2, index=P + 2, comp=EX, amplitude=1.0, **pul)],
        "dft": [{"axis": 2, "index": n - P - 2, "omega": list(om), "comps": [EX, EY, HX, HY],
                 "n_begin": 1, "n_end": 0}],
        "probes": [[int(c), int(c), int(c + 12)], [int(c), int(c), P + 4]],
        "nsteps": nsteps,
    }


def c2_gold_dimer(seed=0, nsteps=4000, n=512, nfreq=64, radius=40.0):
    """512^3, ds = 1 nm, 12-cell PML, a gold sphere dimer (r = 40 nm, gap
    ~ U(2, 10) nm from the seed) along x, x-polarised Gaussian plane source,
    DFT at 64 wavelengths (400-900 nm) on 3 orthogonal planes through the
    dimer centre, 8 point probes.  `n` and `radius` (nm) shrink the case
    (same PML, planes, probes and draws) to sizes the CPU oracle runs."""
    rng = np.random.default_rng(seed)
    gap_nm = float(rng.uniform(2.0, 10.0))
    ds = 1.0 * NANO
    P = 12
    r = float(radius)
    c = (n + 1) / 2.0
    off = r + gap_nm / 2.0
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    om = list(wavelength_frequencies(400e-9, 900e-9, nfreq))
    mid = int(c)
    probes = [[mid, mid, mid], [int(c - off), mid, mid], [int(c + off), mid, mid], [mid, mid, P + 4],
              [mid, mid, n - P - 4], [mid, P + 4, mid], [P + 4, mid, mid], [mid, mid, int(c + r + 10)]]
    return {
        "name": f"gold_dimer_{n}",
        "gap_nm": gap_nm,
        "grid": {"nx": n, "ny": n, "nz": n, "ds": ds, "courant": COURANT_3D, "periodic": [0, 0, 0]},
        "pml": {"mode": "all_faces", "npml": [P, P, P], "order": 3.0, "R0": 1e-6},
        "materials": [_vacuum((n, n, n)),
                      {"type": "drude2cp", "region": sphere((c - off, c, c), r), "params": gold_params_dict()},
                      {"type": "drude2cp", "region": sphere((c + off, c, c), r), "params": gold_params_dict()}],
        "sources": [dict(axis=2, index=P + 2, comp=EX, amplitude=1.0, **pul)],
        "dft": [{"axis": 0, "index": mid, "omega": om, "comps": [EY, EZ, HY, HZ]},
                {"axis": 1, "index": mid, "omega": om, "comps": [EX, EZ, HX, HZ]},
                {"axis": 2, "index": mid, "omega": om, "comps": [EX, EY, HX, HY]}],
        "probes": probes,
        "nsteps": nsteps,
    }


def c3_graphene_disks(seed=0, nsteps=6000, nx=384, ny=384, nz=256, nfreq=128, disk_radius=30.0,
                      disk_height=20.0):
    """384x384x256, ds = 1 nm, 12-cell PML; a one-cell graphene sheet at
    z = 128 (Kubo intraband Drude, E_F = 0.4 eV, tau = 0.1 ps, T = 300 K,
    plus the Raman term); beneath it a 4x4 array of gold disks (r = 30 nm,
    h = 20 nm) with centres jittered by U(-3, 3) nm from the seed; DFT at
    128 wavelengths.  Grid and disk sizes shrink the case (same draws)."""
    rng = np.random.default_rng(seed)
    ds = 1.0 * NANO
    P = 12
    zs = nz // 2
    pitch_x = (nx - 2 * P) / 4.0
    pitch_y = (ny - 2 * P) / 4.0
    mats = [_vacuum((nx, ny, nz))]
    for a in range(4):
        for b in range(4):
            jx, jy = rng.uniform(-3.0, 3.0, size=2)
            cx = P + pitch_x * (a + 0.5) + 0.5 + jx
            cy = P + pitch_y * (b + 0.5) + 0.5 + jy
            mats.append({"type": "drude2cp",
                         "region": cylinder((cx, cy, zs - disk_height / 2.0 - 0.5), disk_radius, disk_height / 2.0, 2),
                         "params": gold_params_dict()})
    gr = {"model": "kubo", "normal_axis": 2, "eps_inf": 1.0, "fermi_ev": 0.4, "tau": 1e-13,
          "temperature": 300.0, "wNL": 2.976e14, "dnuNL": 3.0e12, "xi3": 0.05}
    mats.append({"type": "graphene", "region": box((P + 1, P + 1, zs), (nx - P, ny - P, zs)), "params": gr})
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    om = list(wavelength_frequencies(400e-9, 900e-9, nfreq))
    return {
        "name": f"graphene_disks_{nx}x{ny}x{nz}",
        "grid": {"nx": nx, "ny": ny, "nz": nz, "ds": ds, "courant": COURANT_3D, "periodic": [0, 0, 0]},
        "pml": {"mode": "all_faces", "npml": [P, P, P], "order": 3.0, "R0": 1e-6},
        "materials": mats,
        "sources": [dict(axis=2, index=P + 2, comp=EX, amplitude=1.0, **pul)],
        "dft": [{"axis": 2, "index": P + 6, "omega": om, "comps": [EX, EY, HX, HY]},
                {"axis": 2, "index": nz - P - 4, "omega": om, "comps": [EX, EY, HX, HY]}],
        "probes": [[nx // 2, ny // 2, zs], [nx // 2, ny // 2, zs + 20]],
        "nsteps": nsteps,
    }


def scaled(spec_fn, **kw):
    return spec_fn(**kw)
