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
LIB = os.path.join(ORACLE_DIR, "liboracle.so")
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
