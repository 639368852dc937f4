"""Host-side mirror of the reference's 2D TEz driver over include/fdtd2d.h.

The reference's 2D driver (sim2/program.f90, sim2 =
FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate/) sets up a y-slab
split-field PML, materials applied in order, a soft line source on Dx, DFT
lines of Ex and Hz, point traces, and steps a time loop.  ``Simulation2D``
keeps that vocabulary; ``Batch2D`` steps many simulations together.
"""
import ctypes as C
import math

import numpy as np

from . import C0, NANO, PI, _drude2cp, _graphene, gold_params_dict
from . import abi
from . import abi2d
from .abi2d import DFT_EX, DFT_EZ, DFT_HX, DFT_HZ, DX2, DY2, DZ2, EX2, EY2, EZ2, HX2, HY2, HZ2, TEZ, TMZ  # noqa: F401

_binding2d = None


def library2d():
    """The HIP library's 2D entry points (fdtd2d_*).  Raises if not built."""
    global _binding2d
    if _binding2d is None:
        from . import library
        b = library()
        _binding2d = abi2d.Binding2D(b.dll, "fdtd2d_", b.last_error)
        _binding2d.preset_binding = b  # fdtd_lorentz_jp_preset
    return _binding2d


def box2(lo, hi):
    """1-based inclusive box (xinit, yinit) .. (xend, yend)."""
    return {"kind": "box", "lo": [int(lo[0]), int(lo[1])], "hi": [int(hi[0]), int(hi[1])]}


def disk(center, radius):
    """Cells whose centres lie within `radius` (cells) of `center`."""
    return {"kind": "disk", "center": [float(center[0]), float(center[1])], "radius": float(radius)}


def _region2d(d):
    r = abi2d.Region2D()
    r.kind = abi2d.REGION_BOX if d["kind"] == "box" else abi2d.REGION_DISK
    for a in range(2):
        r.lo[a] = int(d.get("lo", [0, 0])[a])
        r.hi[a] = int(d.get("hi", [0, 0])[a])
        r.center[a] = float(d.get("center", [0.0, 0.0])[a])
    r.radius = float(d.get("radius", 0.0))
    return r


class Simulation2D:
    """One 2D domain, TEz (the reference's: Hz, Ex, Ey) or TMz (Ez, Hx, Hy),
    on the HIP library by default or on the oracle's binding."""

    def __init__(self, nx, ny, ds, courant, precision="fp64", periodic_x=True, binding=None, polarization="tez"):
        self.b = binding or library2d()
        g = abi2d.Grid2D()
        g.nx, g.ny = int(nx), int(ny)
        g.precision = abi.FP32 if precision in ("fp32", abi.FP32) else abi.FP64
        g.periodic_x = int(bool(periodic_x))
        g.ds, g.courant = float(ds), float(courant)
        g.polarization = TMZ if polarization in ("tmz", TMZ) else TEZ
        self.tmz = g.polarization == TMZ
        self.grid = g
        self.shape = (int(nx), int(ny))
        self.h = C.c_void_p()
        self.b.check(self.b.create(C.byref(self.h), C.byref(g)))
        self._dfts = []

    def close(self):
        if self.h:
            self.b.destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pml(self, npml, order=3.0, R0=math.exp(-16.0), sigmam=0.0, eps_low=1.0, eps_high=1.0):
        p = abi2d.Pml2D()
        p.npml[0], p.npml[1] = int(npml[0]), int(npml[1])
        p.order, p.R0, p.sigmam, p.eps_low, p.eps_high = order, R0, sigmam, eps_low, eps_high
        self.b.check(self.b.set_pml(self.h, C.byref(p)))

    def add_dielectric(self, region, eps):
        self.b.check(self.b.add_dielectric(self.h, C.byref(_region2d(region)), float(eps)))

    def add_drude2cp(self, region, params=None):
        p = _drude2cp(params or gold_params_dict())
        self.b.check(self.b.add_drude2cp(self.h, C.byref(_region2d(region)), C.byref(p)))

    def add_graphene(self, region, params):
        p = _graphene(params)
        self.b.check(self.b.add_graphene(self.h, C.byref(_region2d(region)), C.byref(p)))

    def add_lorentz_jp(self, region, params):
        """A Lorentz-Jp material: params as fdtd_amd.lorentz_jp takes them
        (a preset name plus wNL, dnuNL, xi3 for the graphene kinds)."""
        from . import lorentz_jp
        p = lorentz_jp(params, getattr(self.b, "preset_binding", None))
        self.b.check(self.b.add_lorentz_jp(self.h, C.byref(_region2d(region)), C.byref(p)))

    def add_line_source(self, j, t0, st, w0, amplitude=1.0):
        p = abi.Pulse(t0, st, w0, amplitude)
        self.b.check(self.b.add_line_source(self.h, int(j), C.byref(p)))

    def add_dft_line(self, j, omega, mask=None, n_begin=1, n_end=0):
        """mask None: both components of the polarisation (Ex, Hz or Ez, Hx)."""
        if mask is None:
            mask = (DFT_EZ | DFT_HX) if self.tmz else (DFT_EX | DFT_HZ)
        om = np.ascontiguousarray(omega, dtype=np.float64)
        i = C.c_int()
        self.b.check(self.b.add_dft_line(self.h, int(j), len(om), om.ctypes.data_as(abi._PD), int(mask),
                                         int(n_begin), int(n_end), C.byref(i)))
        self._dfts.append(len(om))
        return i.value

    def add_probe(self, i, j):
        r = C.c_int()
        self.b.check(self.b.add_probe(self.h, int(i), int(j), C.byref(r)))
        return r.value

    def add_tfsf(self, lo, src_index, t0, st, w0, amplitude=1.0, hi=0, npml=10, eps=1.0, order=3.0,
                 R0=math.exp(-16.0)):
        """Plane wave along y by total-field / scattered-field (fdtd2d_add_tfsf)."""
        t = abi2d.Tfsf2D(int(lo), int(hi), int(src_index), int(npml), float(eps), float(order), float(R0),
                         abi.Pulse(t0, st, w0, amplitude))
        i = C.c_int()
        self.b.check(self.b.add_tfsf(self.h, C.byref(t), C.byref(i)))
        return i.value

    def tfsf_line(self, tid=0):
        e = np.empty(self.shape[1] + 2)
        h = np.empty(self.shape[1] + 2)
        self.b.check(self.b.get_tfsf_line(self.h, int(tid), e.ctypes.data_as(abi._PD), h.ctypes.data_as(abi._PD)))
        return e, h

    def commit(self):
        self.b.check(self.b.commit(self.h))

    def step(self, n=1):
        self.b.check(self.b.step(self.h, int(n)))

    def sync(self):
        self.b.check(self.b.sync(self.h))

    @property
    def steps_done(self):
        n = C.c_int()
        self.b.check(self.b.get_step(self.h, C.byref(n)))
        return n.value

    @property
    def dt(self):
        d = C.c_double()
        self.b.check(self.b.get_dt(self.h, C.byref(d)))
        return d.value

    def field(self, comp):
        """Interior field as an (nx, ny) Fortran-ordered float64 array."""
        out = np.empty(self.shape, dtype=np.float64, order="F")
        self.b.check(self.b.get_field(self.h, int(comp), out.ctypes.data_as(abi._PD)))
        return out

    def set_field(self, comp, values):
        v = np.asfortranarray(values, dtype=np.float64)
        assert v.shape == self.shape
        self.b.check(self.b.set_field(self.h, int(comp), v.ctypes.data_as(abi._PD)))

    def dft(self, line, comp):
        nf = self._dfts[line]
        re = np.empty((self.shape[0], nf), dtype=np.float64, order="F")
        im = np.empty((self.shape[0], nf), dtype=np.float64, order="F")
        self.b.check(self.b.get_dft(self.h, int(line), int(comp), re.ctypes.data_as(abi._PD),
                                    im.ctypes.data_as(abi._PD)))
        return re, im

    def flux(self, line):
        S = np.empty(self._dfts[line], dtype=np.float64)
        self.b.check(self.b.get_flux(self.h, int(line), S.ctypes.data_as(abi._PD)))
        return S

    def probe(self, pid):
        n = C.c_int()
        self.b.check(self.b.get_probe_count(self.h, int(pid), C.byref(n)))
        out = np.empty((n.value, 3), dtype=np.float64)
        if n.value:
            self.b.check(self.b.get_probe(self.h, int(pid), out.ctypes.data_as(abi._PD)))
        return out


class Batch2D:
    """Many committed Simulation2D handles stepped together (HIP only)."""

    def __init__(self, sims, binding=None):
        self.b = binding or library2d()
        self.sims = list(sims)
        arr = (C.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        self.h = C.c_void_p()
        self.b.check(self.b.batch_create(C.byref(self.h), arr, len(self.sims)))

    def run(self, nsteps):
        n = np.ascontiguousarray(nsteps, dtype=np.int32)
        assert len(n) == len(self.sims)
        self.b.check(self.b.batch_run(self.h, n.ctypes.data_as(C.POINTER(C.c_int32))))

    def sync(self):
        self.b.check(self.b.batch_sync(self.h))

    def close(self):
        if self.h:
            self.b.batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def apply_spec2d(sim, spec):
    p = spec.get("pml")
    if p:
        sim.set_pml(p["npml"], p.get("order", 3.0), p.get("R0", math.exp(-16.0)), p.get("sigmam", 0.0),
                    p.get("eps_low", 1.0), p.get("eps_high", 1.0))
    for m in spec.get("materials", []):
        if m["type"] == "dielectric":
            sim.add_dielectric(m["region"], m["eps"])
        elif m["type"] == "drude2cp":
            sim.add_drude2cp(m["region"], m.get("params") or gold_params_dict())
        elif m["type"] == "graphene":
            sim.add_graphene(m["region"], m["params"])
        elif m["type"] == "lorentz_jp":
            sim.add_lorentz_jp(m["region"], m["params"])
        else:
            raise ValueError(m["type"])
    for s in spec.get("sources", []):
        sim.add_line_source(s["j"], s["t0"], s["st"], s["w0"], s.get("amplitude", 1.0))
    for t in spec.get("tfsf", []):
        sim.add_tfsf(t["lo"], t["src_index"], t["t0"], t["st"], t["w0"], t.get("amplitude", 1.0), t.get("hi", 0),
                     t.get("npml", 10), t.get("eps", 1.0), t.get("order", 3.0), t.get("R0", math.exp(-16.0)))
    ids = [sim.add_dft_line(d["j"], np.array(d["omega"]), d.get("mask"), d.get("n_begin", 1),
                            d.get("n_end", 0)) for d in spec.get("dft", [])]
    pids = [sim.add_probe(*pr) for pr in spec.get("probes", [])]
    return ids, pids


def make_simulation2d(spec, precision="fp64", binding=None):
    g = spec["grid"]
    sim = Simulation2D(g["nx"], g["ny"], g["ds"], g["courant"], precision=precision,
                       periodic_x=g.get("periodic_x", True), binding=binding,
                       polarization=g.get("polarization", "tez"))
    ids, pids = apply_spec2d(sim, spec)
    return sim, ids, pids


# ---- workloads -------------------------------------------------------------------
def sweep_instance(rng, side=None, nsteps=None, ds=1e-9, npml=16, nfreq=32, polarization="tez"):
    """One simulation of the batched 2D sweep, drawn from `rng` in this order:
    grid side (uniform 256-2048 cells), gold cylinder radius (uniform
    10-80 nm), graphene Fermi level (uniform 0.1-0.6 eV), step count (uniform
    3000-8000), polarisation (TEz or TMz, even odds).  Layout along y:
    16-cell PML, line source, reflection line, the gold disk at mid-height
    with a Kubo graphene line (tau 0.1 ps, 300 K) two cells above it,
    transmission line, PML; x periodic (TEz: the reference's x grading
    inside the slabs; TMz: y slabs only).  `side`/`nsteps` override the
    drawn values and `polarization` ("tez", "tmz" or "drawn") the drawn
    polarisation (the draws still happen, so the stream stays aligned)."""
    from .workloads import pulse_for_band

    s_draw = int(rng.integers(256, 2049))
    r_nm = float(rng.uniform(10.0, 80.0))
    ef = float(rng.uniform(0.1, 0.6))
    n_draw = int(rng.integers(3000, 8001))
    tm_draw = bool(rng.random() < 0.5)
    tm = tm_draw if polarization == "drawn" else polarization == "tmz"
    n = side or s_draw
    steps = nsteps or n_draw
    P = npml
    c = (n + 1) / 2.0
    r = min(r_nm * 1e-9 / ds, n / 2.0 - P - 8)
    g_row = int(round(c + r)) + 2
    gr = {"model": "kubo", "normal_axis": 1, "thickness": 3.4e-10, "eps_inf": 1.0, "fermi_ev": ef,
          "tau": 1e-13, "temperature": 300.0, "wNL": 0.0, "dnuNL": 0.0, "xi3": 0.0}
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    om = 2.0 * PI * C0 / np.linspace(400e-9, 900e-9, nfreq)
    return {
        "name": f"sweep2d_{'tmz' if tm else 'tez'}_{n}",
        "grid": {"nx": n, "ny": n, "ds": ds, "courant": 0.5, "periodic_x": True,
                 "polarization": "tmz" if tm else "tez"},
        "pml": {"npml": [0 if tm else P, P], "order": 3.0, "R0": math.exp(-16.0), "eps_low": 1.0, "eps_high": 1.0},
        "materials": [{"type": "dielectric", "region": box2((1, P), (n, n - P)), "eps": 1.0},
                      {"type": "drude2cp", "region": disk((c, c), r)},
                      {"type": "graphene", "region": box2((1, g_row), (n, g_row)), "params": gr}],
        "sources": [dict(j=P + 2, **pul)],
        "dft": [{"j": P + 4, "omega": om}, {"j": n - P - 4, "omega": om}],
        "probes": [[int(c), n - P - 4]],
        "nsteps": steps,
        "radius_nm": r * ds / 1e-9,
        "fermi_ev": ef,
    }


def tmz_case(periodic_x=False, nx=40, ny=48, nsteps=300, graphene="cp4", precision_hint=None):
    """A small TMz (Ez, Hx, Hy) case: PML on the y faces (and the x faces
    when x is not periodic), dielectric layers below (over the source row)
    and above, a gold disk, a graphene line
    above it (CP4 with the Raman term, or Kubo), an Ez line source, DFT lines
    of Ez and Hx below and above the particle, two probes."""
    from .workloads import pulse_for_band
    from . import graphene_ref_dict

    P = (0 if periodic_x else 6, 8)
    ds = 2e-9
    cx, cy = (nx + 1) / 2.0, ny / 2.0
    if graphene == "kubo":
        gr = {"model": "kubo", "normal_axis": 1, "thickness": 3.4e-10, "eps_inf": 1.0, "fermi_ev": 0.4,
              "tau": 1e-13, "temperature": 300.0, "wNL": 2.976e14, "dnuNL": 3.0e12, "xi3": 0.05}
    else:
        gr = dict(graphene_ref_dict())
        gr.update(ref_init_order=0, wNL=2.976e14, dnuNL=3.0e12, xi3=0.05)
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    om = 2.0 * PI * C0 / np.linspace(400e-9, 900e-9, 12)
    g_row = int(round(cy + 6)) + 2
    return {
        "name": f"tmz_{'per' if periodic_x else 'pml'}_{graphene}",
        "grid": {"nx": nx, "ny": ny, "ds": ds, "courant": 0.6, "periodic_x": periodic_x, "polarization": "tmz"},
        "pml": {"npml": list(P), "order": 3.0, "R0": math.exp(-16.0)},
        "materials": [{"type": "dielectric", "region": box2((1, 1), (nx, ny)), "eps": 1.0},
                      {"type": "dielectric", "region": box2((1, 1), (nx, P[1] + 4)), "eps": 1.44},
                      {"type": "dielectric", "region": box2((1, ny - P[1] - 9), (nx, ny)), "eps": 2.25},
                      {"type": "drude2cp", "region": disk((cx, cy), 6.0)},
                      {"type": "graphene", "region": box2((P[0] + 1, g_row), (nx - P[0], g_row)), "params": gr}],
        "sources": [dict(j=P[1] + 3, **pul)],
        "dft": [{"j": P[1] + 5, "omega": om}, {"j": ny - P[1] - 3, "omega": om, "n_begin": 40}],
        "probes": [[int(cx), P[1] + 5], [3, g_row]],
        "nsteps": nsteps,
    }


def sweep_stream(seed=0, count=256, side=None, nsteps=None, polarization="drawn"):
    """`count` sweep instances from one seed (TEz and TMz drawn by default)."""
    rng = np.random.default_rng(seed)
    return [sweep_instance(rng, side=side, nsteps=nsteps, polarization=polarization) for _ in range(count)]


def ref2d_frequencies(lam_init, lam_end, nfreq):
    """sim2/program.f90:121-126,294-298: arg(m) = 2*pi*freq with freq
    advanced by repeated addition of df from c0/lam_end."""
    finit = C0 / lam_end
    fend = C0 / lam_init
    df = (fend - finit) / nfreq
    arg = []
    freq = finit
    for _ in range(nfreq):
        arg.append(2.0 * PI * freq)
        freq = freq + df
    return np.array(arg)


def ref2d_shipped(nsteps=None):
    """The reference's shipped 2D case (sim2/program.f90:40-300): 125 x 1266
    cells of 3.999 nm, Ku = 0.5, 10-cell y PML (air below, Si above), a gold
    strip on air over SiO2 over Si, Dx line source at y = 12, DFT lines at
    y = 14 (incident n < tfinStop, reflected n > tfinStop) and y = 1170
    (transmitted), traces at x = 62 on both lines.  Parameters are computed
    as the reference computes them (Fortran integer truncation included)."""
    lambda0 = 633.0 * NANO
    f0 = C0 / lambda0
    ku = 0.5
    ds = 3.999 * NANO
    dt = ku * ds / C0
    nx = math.floor(500.0 * NANO / ds)
    ny = math.floor(5064.0 * NANO / ds)
    ic = nx // 2
    npml = 10
    order = 3.0
    R0 = math.exp(-16.0)
    sigmam = (order + 1.0) * C0 * math.log(1.0 / R0) / (2.0 * npml * ds)
    periodo = 1.0 / f0
    w0 = 2.0 * PI * f0
    st = 1.0 * periodo
    t0 = 4.0 * st
    sposy = npml + 2
    air_y1 = int(0.8 * ny)
    gl = math.floor(100.0 * NANO / ds)
    au_x0 = int(ic - 0.5 * gl + 1)
    au_x1 = int(ic + 0.5 * gl)
    au_y0 = air_y1 - math.floor(20.0 * NANO / ds) + 1
    sio2_y0 = air_y1 + 1
    sio2_y1 = sio2_y0 + math.floor(284.0 * NANO / ds) - 1
    si_y0 = sio2_y1 + 1
    si_y1 = ny - npml
    ptty = int(0.5 * (si_y0 + si_y1))
    dist = ds * float(air_y1 - sposy) + 1.0 * 4.5 * lambda0 / 1.0
    tfin = math.floor((dist / C0) / dt)
    nt = 4 * tfin
    arg = ref2d_frequencies(500e-9, 800e-9, 200)
    return {
        "name": "ref2d_shipped",
        "grid": {"nx": nx, "ny": ny, "ds": ds, "courant": ku, "periodic_x": True},
        "pml": {"npml": [npml, npml], "order": order, "sigmam": sigmam, "eps_low": 1.0, "eps_high": 3.88 ** 2},
        "materials": [
            {"type": "dielectric", "region": box2((1, npml), (nx, air_y1)), "eps": 1.0},
            {"type": "drude2cp", "region": box2((au_x0, au_y0), (au_x1, air_y1))},
            {"type": "dielectric", "region": box2((1, sio2_y0), (nx, sio2_y1)), "eps": 1.46 ** 2},
            {"type": "dielectric", "region": box2((1, si_y0), (nx, si_y1)), "eps": 3.88 ** 2},
        ],
        "sources": [{"j": sposy, "t0": t0, "st": st, "w0": w0}],
        "dft": [{"j": sposy + 2, "omega": arg, "n_begin": 1, "n_end": tfin - 1},
                {"j": sposy + 2, "omega": arg, "n_begin": tfin + 1, "n_end": 0},
                {"j": ptty, "omega": arg, "n_begin": 1, "n_end": 0}],
        "probes": [[ic, sposy + 2], [ic, ptty]],
        "nsteps": nt if nsteps is None else nsteps,
        "tfinStop": tfin,
        "frame_every": int(nt / 20.0),
        "ic": ic,
    }
