"""Host-side mirror of the reference's 3D module API over the gfx950 C ABI.

The reference drives its physics through Fortran type-bound procedures
(sim3/program.f90:205-293 set-up, :382-607 time loop).  This package keeps
that vocabulary -- grid, PML, materials applied in order, a plane source,
DFT monitors, probes, the time-loop call -- over ``include/fdtd.h``.

``Simulation`` binds the HIP library (``libfdtd_amd.so``); the same class
binds any library exporting the ABI under another prefix.  The library
must be built first (``__graft_entry__.build()``); there is no CPU fallback.
"""
import ctypes as C
import math
import os

import numpy as np

from . import abi
from .abi import (EX, EY, EZ, HX, HY, HZ, DX, DY, DZ, FP32, FP64, PML_ALL_FACES,  # noqa: F401
                  PML_NONE, PML_REF_Y, ENGINE_AUTO, ENGINE_FUSED, ENGINE_STAGED, FdtdError)

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(ROOT_DIR, "libfdtd_amd.so")

# misc/constants_mod.f90:3-24
PI = 4.0 * math.atan(1.0)
EP0 = 8.85418781762039e-12
MU0 = PI * 4e-7
C0 = 1.0 / math.sqrt(EP0 * MU0)
NANO = 1e-9

_binding = None


def library(path=None):
    """Load the HIP library (once).  Raises if it is not built."""
    global _binding
    if _binding is not None and path is None:
        return _binding
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library not built: {p} (run __graft_entry__.build())")
    b = abi.Binding(C.CDLL(p), "fdtd_")
    if path is None:
        _binding = b
    return b


# ---- regions ------------------------------------------------------------------
def box(lo, hi):
    """1-based inclusive global box (the reference's xinit..zend)."""
    return {"kind": "box", "lo": [int(v) for v in lo], "hi": [int(v) for v in hi]}


def sphere(center, radius, clip_lo=None, clip_hi=None):
    r = {"kind": "sphere", "center": [float(v) for v in center], "radius": float(radius)}
    if clip_lo is not None:
        r["lo"], r["hi"] = list(clip_lo), list(clip_hi)
    return r


def cylinder(center, radius, half_length, axis):
    return {"kind": "cylinder", "center": [float(v) for v in center], "radius": float(radius),
            "half_length": float(half_length), "axis": int(axis)}


def _region(d):
    r = abi.Region()
    kinds = {"box": abi.REGION_BOX, "sphere": abi.REGION_SPHERE, "cylinder": abi.REGION_CYLINDER}
    r.kind = kinds[d["kind"]]
    r.axis = int(d.get("axis", 0))
    for a in range(3):
        r.lo[a] = int(d.get("lo", [0, 0, 0])[a])
        r.hi[a] = int(d.get("hi", [0, 0, 0])[a])
        r.center[a] = float(d.get("center", [0.0, 0.0, 0.0])[a])
    r.radius = float(d.get("radius", 0.0))
    r.half_length = float(d.get("half_length", 0.0))
    return r


def gold_params(binding=None):
    p = abi.Drude2CP()
    (binding or library()).drude2cp_gold(C.byref(p))
    return p


def gold_params_dict():
    """mod3/gold_mod.f90:28-34 as a plain dict."""
    return {"eps_inf": 1.1431, "gamma": 1.0805e14, "omega_d": 1.3202e16,
            "A": [0.26698, 3.0834], "phi": [-1.2371, -1.0968],
            "Omega": [3.8711e15, 4.1684e15], "Gamma": [4.4642e14, 2.4555e15]}


def graphene_ref_dict():
    """mod3/graphene_mod.f90:33-41 (CP4) with the reference's init order."""
    return {"model": "cp4", "normal_axis": 1, "ref_init_order": 1, "thickness": 3.4e-10,
            "eps_inf": 1.0, "gamma": 9e15, "omega_d": 5.2796e15,
            "A": [13.1068, 1.8417, 0.4191, 0.7737], "phi": [-1.8391, 0.3505, 3.0565, -0.1199],
            "Omega": [6.1264e14, 6.6843e15, 5.6289e15, 1.118e16],
            "Gamma": [1e14, 7.905e14, 6.7526e14, 1.0005e14],
            "wNL": 0.0, "dnuNL": 0.0, "xi3": 0.0}


def lorentz_jp(d, binding=None):
    """fdtd_lorentz_jp from a dict: {"preset": name} (gold, si, titanium,
    film_gold, graphene, graphene3_1nm, graphene_nd: the Lorentz-Jp modules'
    tables) and/or explicit fields, which override the preset's."""
    b = binding or library()
    p = abi.LorentzJp()
    if "preset" in d:
        b.check(b.lorentz_jp_preset(abi.LJP_KINDS[d["preset"]], C.byref(p)))
    for name in ("npols", "jp_mask", "nonlinear", "ez_uses_jx"):
        if name in d:
            setattr(p, name, int(d[name]))
    for name in ("eps_inf", "wNL", "dnuNL", "xi3"):
        if name in d:
            setattr(p, name, float(d[name]))
    for name in ("xi", "w0", "dnu"):
        if name in d:
            arr = getattr(p, name)
            for l, v in enumerate(d[name]):
                arr[l] = float(v)
    return p


def _drude2cp(d):
    p = abi.Drude2CP()
    p.eps_inf, p.gamma, p.omega_d = d["eps_inf"], d["gamma"], d["omega_d"]
    for q in range(2):
        p.A[q], p.phi[q], p.Omega[q], p.Gamma[q] = d["A"][q], d["phi"][q], d["Omega"][q], d["Gamma"][q]
    return p


def _graphene(d):
    p = abi.Graphene()
    p.model = abi.GRAPHENE_CP4 if d.get("model", "cp4") == "cp4" else abi.GRAPHENE_KUBO
    p.normal_axis = int(d.get("normal_axis", 1))
    p.ref_init_order = int(d.get("ref_init_order", 0))
    p.thickness = float(d.get("thickness", 3.4e-10))
    p.eps_inf = float(d.get("eps_inf", 1.0))
    p.gamma = float(d.get("gamma", 0.0))
    p.omega_d = float(d.get("omega_d", 0.0))
    for q in range(4):
        p.A[q] = float(d.get("A", [0] * 4)[q])
        p.phi[q] = float(d.get("phi", [0] * 4)[q])
        p.Omega[q] = float(d.get("Omega", [0] * 4)[q])
        p.Gamma[q] = float(d.get("Gamma", [0] * 4)[q])
    p.fermi_ev = float(d.get("fermi_ev", 0.0))
    p.tau = float(d.get("tau", 0.0))
    p.temperature = float(d.get("temperature", 0.0))
    p.wNL = float(d.get("wNL", 0.0))
    p.dnuNL = float(d.get("dnuNL", 0.0))
    p.xi3 = float(d.get("xi3", 0.0))
    return p


class Simulation:
    """One 3D Yee domain on one GPU (or any backend exporting the ABI)."""

    def __init__(self, nx, ny, nz, ds, courant, precision="fp64", periodic=(0, 0, 0), engine="auto",
                 binding=None):
        self.b = binding or library()
        g = abi.Grid()
        g.nx, g.ny, g.nz = int(nx), int(ny), int(nz)
        g.precision = FP32 if precision in ("fp32", FP32) else FP64
        for a in range(3):
            g.periodic[a] = int(bool(periodic[a]))
        g.engine = {"auto": ENGINE_AUTO, "staged": ENGINE_STAGED, "fused": ENGINE_FUSED}.get(engine, engine)
        g.ds = float(ds)
        g.courant = float(courant)
        self.grid = g
        self.shape = (int(nx), int(ny), int(nz))
        self.h = C.c_void_p()
        self.b.check(self.b.create(C.byref(self.h), C.byref(g)))
        self._dfts = []
        self._probes = []

    # ---- lifetime
    def close(self):
        if self.h:
            self.b.destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- set-up (mirrors of the reference's initialize calls)
    def set_pml(self, mode, npml, order=3.0, R0=math.exp(-16.0), sigmam=0.0, eps_low=1.0, eps_high=1.0):
        p = abi.Pml()
        p.mode = {"none": PML_NONE, "ref_y": PML_REF_Y, "all_faces": PML_ALL_FACES}.get(mode, mode)
        for a in range(3):
            p.npml[a] = int(npml[a])
        p.order, p.R0, p.sigmam, p.eps_low, p.eps_high = order, R0, sigmam, eps_low, eps_high
        self.b.check(self.b.set_pml(self.h, C.byref(p)))

    def add_dielectric(self, region, eps):
        self.b.check(self.b.add_dielectric(self.h, C.byref(_region(region)), float(eps)))

    def add_drude2cp(self, region, params=None):
        p = _drude2cp(params or gold_params_dict())
        self.b.check(self.b.add_drude2cp(self.h, C.byref(_region(region)), C.byref(p)))

    def add_graphene(self, region, params):
        p = _graphene(params)
        self.b.check(self.b.add_graphene(self.h, C.byref(_region(region)), C.byref(p)))

    def add_lorentz_jp(self, region, params):
        p = lorentz_jp(params, self.b)
        self.b.check(self.b.add_lorentz_jp(self.h, C.byref(_region(region)), C.byref(p)))

    def add_lorentz_d(self, region, params):
        """fdtd_add_lorentz_d: the D-form Lorentz poles of mod3/si_mod.f90
        over a pole table ({"preset": "si"} is the module's own)."""
        p = lorentz_jp(params, self.b)
        self.b.check(self.b.add_lorentz_d(self.h, C.byref(_region(region)), C.byref(p)))

    def add_plane_source(self, axis, index, comp, t0, st, w0, amplitude=1.0):
        p = abi.Pulse(t0, st, w0, amplitude)
        self.b.check(self.b.add_plane_source(self.h, int(axis), int(index), int(comp), C.byref(p)))

    def add_dft_plane(self, axis, index, omega, comps=None, n_begin=1, n_end=0):
        om = np.ascontiguousarray(omega, dtype=np.float64)
        if comps is None:
            comps = {0: (EY, EZ, HY, HZ), 1: (EX, EZ, HX, HZ), 2: (EX, EY, HX, HY)}[axis]
        mask = 0
        for c in comps:
            mask |= 1 << c
        i = C.c_int()
        self.b.check(self.b.add_dft_plane(self.h, int(axis), int(index), len(om),
                                          om.ctypes.data_as(abi._PD), mask, int(n_begin), int(n_end), C.byref(i)))
        n = self.shape
        n1 = n[1] if axis == 0 else n[0]
        n2 = n[1] if axis == 2 else n[2]
        self._dfts.append((n1, n2, len(om)))
        return i.value

    def add_tfsf(self, axis, comp, lo, src_index, t0, st, w0, amplitude=1.0, hi=0, npml=10, eps=1.0, order=3.0,
                 R0=math.exp(-16.0)):
        """TFSF plane wave along +axis (fdtd_tfsf); returns its id."""
        t = abi.Tfsf()
        t.axis, t.comp, t.lo, t.hi, t.src_index, t.npml = int(axis), int(comp), int(lo), int(hi), int(src_index), int(npml)
        t.eps, t.order, t.R0 = float(eps), float(order), float(R0)
        t.pulse.t0, t.pulse.st, t.pulse.w0, t.pulse.amplitude = t0, st, w0, amplitude
        out = C.c_int()
        self.b.check(self.b.add_tfsf(self.h, C.byref(t), C.byref(out)))
        self._tfsf_n = getattr(self, "_tfsf_n", {})
        self._tfsf_n[out.value] = self.shape[int(axis)]
        return out.value

    def add_tfsf_probe(self, tfsf_id, index):
        out = C.c_int()
        self.b.check(self.b.add_tfsf_probe(self.h, int(tfsf_id), int(index), C.byref(out)))
        return out.value

    def tfsf_line(self, tfsf_id):
        n = self._tfsf_n[tfsf_id] + 2
        e = np.empty(n, dtype=np.float64)
        h = np.empty(n, dtype=np.float64)
        self.b.check(self.b.get_tfsf_line(self.h, int(tfsf_id), e.ctypes.data_as(abi._PD), h.ctypes.data_as(abi._PD)))
        return e, h

    def add_probe(self, i, j, k):
        r = C.c_int()
        self.b.check(self.b.add_probe(self.h, int(i), int(j), int(k), C.byref(r)))
        self._probes.append((i, j, k))
        return r.value

    def commit(self):
        self.b.check(self.b.commit(self.h))

    # ---- time loop
    def autotune(self, steps=5):
        """fdtd_autotune: time the FUSED step forms on this handle, keep the
        fastest; returns (by, kc), by = 0 for the two-kernel step."""
        by, kc = C.c_int(), C.c_int()
        self.b.check(self.b.autotune(self.h, int(steps), C.byref(by), C.byref(kc)))
        return by.value, kc.value

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

    # ---- results
    def field(self, comp):
        """Interior field as an (nx, ny, nz) Fortran-ordered float64 array."""
        out = np.empty(self.shape, dtype=np.float64, order="F")
        self.b.check(self.b.get_field(self.h, int(comp), out.ctypes.data_as(abi._PD)))
        return out

    def set_field(self, comp, values):
        v = np.asfortranarray(values, dtype=np.float64)
        assert v.shape == self.shape
        self.b.check(self.b.set_field(self.h, int(comp), v.ctypes.data_as(abi._PD)))

    def dft(self, plane, comp):
        n1, n2, nf = self._dfts[plane]
        re = np.empty((n1, n2, nf), dtype=np.float64, order="F")
        im = np.empty((n1, n2, nf), dtype=np.float64, order="F")
        self.b.check(self.b.get_dft(self.h, int(plane), int(comp), re.ctypes.data_as(abi._PD),
                                    im.ctypes.data_as(abi._PD)))
        return re, im

    def flux(self, plane):
        nf = self._dfts[plane][2]
        S = np.empty(nf, dtype=np.float64)
        self.b.check(self.b.get_flux(self.h, int(plane), S.ctypes.data_as(abi._PD)))
        return S

    def checkpoint(self):
        """fdtd_checkpoint_save: the handle's state after the steps done so
        far, as bytes (waits for queued work)."""
        n = C.c_size_t()
        self.b.check(self.b.checkpoint_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self.b.check(self.b.checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, data):
        """fdtd_checkpoint_load: continue from a checkpoint of a handle set
        up the same way (before this handle's first step)."""
        buf = C.create_string_buffer(bytes(data), len(data))
        self.b.check(self.b.checkpoint_load(self.h, buf, len(data)))

    def flux_noalloc(self, plane):
        """fdtd_get_flux_noalloc: the flux through the library's fixed device
        array, with no device allocation (the handle must be committed)."""
        nf = self._dfts[plane][2]
        S = np.empty(nf, dtype=np.float64)
        self.b.check(self.b.get_flux_noalloc(self.h, int(plane), S.ctypes.data_as(abi._PD)))
        return S

    def probe(self, pid):
        n = C.c_int()
        self.b.check(self.b.get_probe_count(self.h, int(pid), C.byref(n)))
        out = np.empty((n.value, 7), dtype=np.float64)
        if n.value:
            self.b.check(self.b.get_probe(self.h, int(pid), out.ctypes.data_as(abi._PD)))
        return out

    def energy(self):
        a, b = C.c_double(), C.c_double()
        self.b.check(self.b.energy(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- GPU introspection
    def memory_bytes(self):
        n = C.c_size_t()
        self.b.check(self.b.memory_bytes(self.h, C.byref(n)))
        return n.value

    def set_profiling(self, on):
        self.b.check(self.b.set_profiling(self.h, int(bool(on))))

    def kernel_stats(self, kernel):
        ms, n = C.c_double(), C.c_int64()
        self.b.check(self.b.kernel_stats(self.h, int(kernel), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def algorithmic_bytes(self, kernel):
        b = C.c_double()
        self.b.check(self.b.algorithmic_bytes(self.h, int(kernel), C.byref(b)))
        return b.value

    def cell_counts(self):
        a, b = C.c_int64(), C.c_int64()
        self.b.check(self.b.cell_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value


def reference_frequencies(lam_init, lam_end, nfreq):
    """sim3/program.f90:121-125,353-357: evenly spaced in frequency by
    repeated addition; the last point is fend - df."""
    finit = C0 / lam_end
    fend = C0 / lam_init
    df = (fend - finit) / nfreq
    arg = []
    freq = finit
    for _ in range(nfreq):
        arg.append(2.0 * PI * freq)
        freq = freq + df
    return np.array(arg)


def wavelength_frequencies(lam_init, lam_end, nfreq):
    """Angular frequencies of wavelengths spaced evenly over [lam_init, lam_end]."""
    lam = np.linspace(lam_init, lam_end, nfreq)
    return 2.0 * PI * C0 / lam


def apply_spec(sim, spec):
    """Apply a JSON-style workload description to a Simulation."""
    p = spec.get("pml")
    if p:
        sim.set_pml(p["mode"], p["npml"], p.get("order", 3.0), p.get("R0", math.exp(-16.0)),
                    p.get("sigmam", 0.0), p.get("eps_low", 1.0), p.get("eps_high", 1.0))
    for m in spec.get("materials", []):
        if m["type"] == "dielectric":
            sim.add_dielectric(m["region"], m["eps"])
        elif m["type"] == "drude2cp":
            sim.add_drude2cp(m["region"], m.get("params") or gold_params_dict())
        elif m["type"] == "graphene":
            sim.add_graphene(m["region"], m["params"])
        elif m["type"] == "lorentz_jp":
            sim.add_lorentz_jp(m["region"], m["params"])
        elif m["type"] == "lorentz_d":
            sim.add_lorentz_d(m["region"], m["params"])
        else:
            raise ValueError(m["type"])
    for s in spec.get("sources", []):
        sim.add_plane_source(s["axis"], s["index"], s["comp"], s["t0"], s["st"], s["w0"], s.get("amplitude", 1.0))
    ids = []
    for d in spec.get("dft", []):
        ids.append(sim.add_dft_plane(d["axis"], d["index"], np.array(d["omega"]), d.get("comps"),
                                     d.get("n_begin", 1), d.get("n_end", 0)))
    pids = [sim.add_probe(*pr) for pr in spec.get("probes", [])]
    tids = []
    for t in spec.get("tfsf", []):
        tids.append(sim.add_tfsf(t["axis"], t["comp"], t["lo"], t["src_index"], t["t0"], t["st"], t["w0"],
                                 t.get("amplitude", 1.0), t.get("hi", 0), t.get("npml", 10), t.get("eps", 1.0),
                                 t.get("order", 3.0), t.get("R0", math.exp(-16.0))))
    pids += [sim.add_tfsf_probe(tids[q], idx) for q, idx in spec.get("tfsf_probes", [])]
    return ids, pids


def make_simulation(spec, precision="fp64", engine="auto", binding=None):
    g = spec["grid"]
    sim = Simulation(g["nx"], g["ny"], g["nz"], g["ds"], g["courant"], precision=precision,
                     periodic=g.get("periodic", (0, 0, 0)), engine=engine, binding=binding)
    ids, pids = apply_spec(sim, spec)
    return sim, ids, pids
