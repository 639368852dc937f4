"""ctypes mirror of include/fdtd2d.h (2D TEz / TMz stepper and batches)."""
import ctypes as C

from .abi import Drude2CP, FdtdError, Graphene, LorentzJp, Pulse  # noqa: F401

TEZ, TMZ = 0, 1
EX2, EY2, HZ2, DX2, DY2, EZ2, HX2, HY2, DZ2 = range(9)
COMP_NAMES_2D = ("Ex", "Ey", "Hz", "Dx", "Dy", "Ez", "Hx", "Hy", "Dz")
REGION_BOX, REGION_DISK = 0, 1
DFT_EX, DFT_HZ, DFT_EZ, DFT_HX = 1, 4, 8, 16


class Grid2D(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("precision", C.c_int32), ("periodic_x", C.c_int32),
                ("ds", C.c_double), ("courant", C.c_double), ("polarization", C.c_int32), ("reserved", C.c_int32)]


class Pml2D(C.Structure):
    _fields_ = [("npml", C.c_int32 * 2), ("order", C.c_double), ("R0", C.c_double), ("sigmam", C.c_double),
                ("eps_low", C.c_double), ("eps_high", C.c_double)]


class Region2D(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("lo", C.c_int32 * 2), ("hi", C.c_int32 * 2),
                ("center", C.c_double * 2), ("radius", C.c_double)]


class Tfsf2D(C.Structure):
    _fields_ = [("lo", C.c_int32), ("hi", C.c_int32), ("src_index", C.c_int32), ("npml", C.c_int32),
                ("eps", C.c_double), ("order", C.c_double), ("R0", C.c_double), ("pulse", Pulse)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PD = C.POINTER(C.c_double)

PROTOTYPES_2D = {
    "create": (_I, [C.POINTER(_P), C.POINTER(Grid2D)]),
    "destroy": (_I, [_P]),
    "set_pml": (_I, [_P, C.POINTER(Pml2D)]),
    "add_dielectric": (_I, [_P, C.POINTER(Region2D), _D]),
    "add_drude2cp": (_I, [_P, C.POINTER(Region2D), C.POINTER(Drude2CP)]),
    "add_graphene": (_I, [_P, C.POINTER(Region2D), C.POINTER(Graphene)]),
    "add_lorentz_jp": (_I, [_P, C.POINTER(Region2D), C.POINTER(LorentzJp)]),
    "add_line_source": (_I, [_P, _I, C.POINTER(Pulse)]),
    "add_dft_line": (_I, [_P, _I, _I, _PD, _I, _I, _I, C.POINTER(_I)]),
    "add_probe": (_I, [_P, _I, _I, C.POINTER(_I)]),
    "add_tfsf": (_I, [_P, C.POINTER(Tfsf2D), C.POINTER(_I)]),
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
}

# batches: the HIP library only
BATCH_PROTOTYPES_2D = {
    "batch_create": (_I, [C.POINTER(_P), C.POINTER(_P), _I]),
    "batch_run": (_I, [_P, C.POINTER(C.c_int32)]),
    "batch_sync": (_I, [_P]),
    "batch_destroy": (_I, [_P]),
}


class Binding2D:
    """2D prototypes of one shared library exporting include/fdtd2d.h under
    `prefix` (``fdtd2d_`` for the HIP library, ``ofdtd2d_`` for the oracle)."""

    def __init__(self, dll, prefix, last_error, batches=True):
        self.dll = dll
        self.prefix = prefix
        table = dict(PROTOTYPES_2D)
        if batches:
            table.update(BATCH_PROTOTYPES_2D)
        for name, (res, args) in table.items():
            fn = getattr(dll, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)
        self.last_error = last_error

    def check(self, rc):
        if rc != 0:
            raise FdtdError(rc, (self.last_error() or b"").decode(errors="replace"))
        return rc
