"""ctypes mirror of include/fdtd.h (structs, enums, prototypes).

The same prototypes are bound for any shared library that exports the ABI
under a symbol prefix (``fdtd_`` for the HIP library).
"""
import ctypes as C

# status codes
FDTD_OK = 0
FDTD_ERR_ARG = -1
FDTD_ERR_STATE = -2
FDTD_ERR_NOMEM = -3
FDTD_ERR_HIP = -4
FDTD_ERR_UNSUPPORTED = -5
FDTD_ERR_NONFINITE = -6

FP64, FP32 = 0, 1
PML_NONE, PML_REF_Y, PML_ALL_FACES = 0, 1, 2
EX, EY, EZ, HX, HY, HZ, DX, DY, DZ = range(9)
COMP_NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dx", "Dy", "Dz")
REGION_BOX, REGION_SPHERE, REGION_CYLINDER = 0, 1, 2
GRAPHENE_CP4, GRAPHENE_KUBO = 0, 1
ENGINE_AUTO, ENGINE_STAGED, ENGINE_FUSED = 0, 1, 2
K_E, K_H, K_ADE, K_PML_E, K_PML_H, K_MAT, K_MON, K_GEN_E, K_GEN_H, K_XF, K_COUNT = range(11)
KERNEL_NAMES = ("e_update", "h_update", "ade", "pml_e", "pml_h", "material", "monitor", "gen_e", "gen_h", "x_face")


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("precision", C.c_int32),
        ("periodic", C.c_int32 * 3),
        ("engine", C.c_int32),
        ("ds", C.c_double),
        ("courant", C.c_double),
    ]


class Pml(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("npml", C.c_int32 * 3),
        ("order", C.c_double),
        ("R0", C.c_double),
        ("sigmam", C.c_double),
        ("eps_low", C.c_double),
        ("eps_high", C.c_double),
    ]


class Region(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("axis", C.c_int32),
        ("lo", C.c_int32 * 3),
        ("hi", C.c_int32 * 3),
        ("center", C.c_double * 3),
        ("radius", C.c_double),
        ("half_length", C.c_double),
    ]


class Drude2CP(C.Structure):
    _fields_ = [
        ("eps_inf", C.c_double),
        ("gamma", C.c_double),
        ("omega_d", C.c_double),
        ("A", C.c_double * 2),
        ("phi", C.c_double * 2),
        ("Omega", C.c_double * 2),
        ("Gamma", C.c_double * 2),
    ]


class Graphene(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("normal_axis", C.c_int32),
        ("ref_init_order", C.c_int32),
        ("reserved", C.c_int32),
        ("thickness", C.c_double),
        ("eps_inf", C.c_double),
        ("gamma", C.c_double),
        ("omega_d", C.c_double),
        ("A", C.c_double * 4),
        ("phi", C.c_double * 4),
        ("Omega", C.c_double * 4),
        ("Gamma", C.c_double * 4),
        ("fermi_ev", C.c_double),
        ("tau", C.c_double),
        ("temperature", C.c_double),
        ("wNL", C.c_double),
        ("dnuNL", C.c_double),
        ("xi3", C.c_double),
    ]


class LorentzJp(C.Structure):
    _fields_ = [
        ("npols", C.c_int32),
        ("jp_mask", C.c_int32),
        ("nonlinear", C.c_int32),
        ("ez_uses_jx", C.c_int32),
        ("eps_inf", C.c_double),
        ("xi", C.c_double * 4),
        ("w0", C.c_double * 4),
        ("dnu", C.c_double * 4),
        ("wNL", C.c_double),
        ("dnuNL", C.c_double),
        ("xi3", C.c_double),
    ]


LJP_GOLD, LJP_SI, LJP_TITANIUM, LJP_FILM_GOLD, LJP_GRAPHENE, LJP_GRAPHENE3_1NM, LJP_GRAPHENE_ND = range(7)
LJP_KINDS = {"gold": LJP_GOLD, "si": LJP_SI, "titanium": LJP_TITANIUM, "film_gold": LJP_FILM_GOLD,
             "graphene": LJP_GRAPHENE, "graphene3_1nm": LJP_GRAPHENE3_1NM, "graphene_nd": LJP_GRAPHENE_ND}


class Pulse(C.Structure):
    _fields_ = [("t0", C.c_double), ("st", C.c_double), ("w0", C.c_double),
                ("amplitude", C.c_double)]


class Tfsf(C.Structure):
    _fields_ = [
        ("axis", C.c_int32),
        ("comp", C.c_int32),
        ("lo", C.c_int32),
        ("hi", C.c_int32),
        ("src_index", C.c_int32),
        ("npml", C.c_int32),
        ("eps", C.c_double),
        ("order", C.c_double),
        ("R0", C.c_double),
        ("pulse", Pulse),
    ]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PD = C.POINTER(C.c_double)

PROTOTYPES = {
    "drude2cp_gold": (None, [C.POINTER(Drude2CP)]),
    "graphene_ref": (None, [C.POINTER(Graphene)]),
    "create": (_I, [C.POINTER(_P), C.POINTER(Grid)]),
    "destroy": (_I, [_P]),
    "last_error": (C.c_char_p, []),
    "set_pml": (_I, [_P, C.POINTER(Pml)]),
    "add_dielectric": (_I, [_P, C.POINTER(Region), _D]),
    "add_drude2cp": (_I, [_P, C.POINTER(Region), C.POINTER(Drude2CP)]),
    "add_graphene": (_I, [_P, C.POINTER(Region), C.POINTER(Graphene)]),
    "add_lorentz_jp": (_I, [_P, C.POINTER(Region), C.POINTER(LorentzJp)]),
    "add_lorentz_d": (_I, [_P, C.POINTER(Region), C.POINTER(LorentzJp)]),
    "lorentz_jp_preset": (_I, [C.c_int32, C.POINTER(LorentzJp)]),
    "add_plane_source": (_I, [_P, _I, _I, _I, C.POINTER(Pulse)]),
    "add_dft_plane": (_I, [_P, _I, _I, _I, _PD, _I, _I, _I, C.POINTER(_I)]),
    "add_probe": (_I, [_P, _I, _I, _I, C.POINTER(_I)]),
    "add_tfsf": (_I, [_P, C.POINTER(Tfsf), C.POINTER(_I)]),
    "add_tfsf_probe": (_I, [_P, _I, _I, C.POINTER(_I)]),
    "get_tfsf_line": (_I, [_P, _I, _PD, _PD]),
    "commit": (_I, [_P]),
    "step": (_I, [_P, _I]),
    "sync": (_I, [_P]),
    "get_step": (_I, [_P, C.POINTER(_I)]),
    "get_dt": (_I, [_P, _PD]),
    "get_field": (_I, [_P, _I, _PD]),
    "set_field": (_I, [_P, _I, _PD]),
    "get_dft": (_I, [_P, _I, _I, _PD, _PD]),
    "get_flux": (_I, [_P, _I, _PD]),
    "get_probe_count": (_I, [_P, _I, C.POINTER(_I)]),
    "get_probe": (_I, [_P, _I, _PD]),
    "energy": (_I, [_P, _PD, _PD]),
}

# entry points only the HIP library has
GPU_PROTOTYPES = {
    "get_flux_noalloc": (_I, [_P, _I, _PD]),
    "checkpoint_size": (_I, [_P, C.POINTER(C.c_size_t)]),
    "checkpoint_save": (_I, [_P, C.c_void_p, C.c_size_t]),
    "checkpoint_load": (_I, [_P, C.c_void_p, C.c_size_t]),
    "memory_bytes": (_I, [_P, C.POINTER(C.c_size_t)]),
    "set_profiling": (_I, [_P, _I]),
    "kernel_stats": (_I, [_P, _I, _PD, C.POINTER(C.c_int64)]),
    "algorithmic_bytes": (_I, [_P, _I, _PD]),
    "cell_counts": (_I, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "estimate_bytes": (_I, [_P, C.POINTER(C.c_size_t)]),
    "autotune": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I)]),
    "pool_create": (_I, [C.POINTER(_P), C.c_size_t, _I, _I]),
    "pool_destroy": (_I, [_P]),
    "pool_submit": (_I, [_P, _P, _I, C.POINTER(_I)]),
    "pool_progress": (_I, [_P, C.POINTER(_I), C.POINTER(_I)]),
    "pool_status": (_I, [_P, _I, C.POINTER(_I), _PD, _PD]),
    "pool_cancel": (_I, [_P, _I]),
    "pool_release": (_I, [_P, _I]),
    "pool_usage": (_I, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
}

JOB_WAITING, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)


class FdtdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fdtd error {code}: {msg}")
        self.code = code


class Binding:
    """Prototypes of one shared library exporting the ABI under `prefix`."""

    def __init__(self, dll, prefix, extra=True):
        self.dll = dll
        self.prefix = prefix
        table = dict(PROTOTYPES)
        if extra:
            table.update(GPU_PROTOTYPES)
        for name, (res, args) in table.items():
            fn = getattr(dll, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def check(self, rc):
        if rc != 0:
            raise FdtdError(rc, (self.last_error() or b"").decode(errors="replace"))
        return rc
