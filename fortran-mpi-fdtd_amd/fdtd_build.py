"""Build steps for the gfx950 library, the Fortran host and the test oracles.

``build_library`` compiles csrc/*.hip with hipcc for gfx950 into
libfdtd_amd.so next to this file (in-tree, so it travels with the tree).
``build_fortran_host`` compiles the iso_c_binding interface module and the
Fortran90 host program against it.  ``build_oracles`` runs oracle/Makefile.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfdtd_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("fdtd_api.hip", "fdtd_kernels.hip", "fdtd_pool.hip",
                                                   "fdtd2d.hip", "fdtd_noalloc.hip")]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no multiply-add contraction: every product and sum is rounded on its
    # own, as in the FP64 CPU builds (x86-64 without FMA) the parity bar
    # compares against; the stencil is bound by memory, not arithmetic
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return res


def build_library():
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *SOURCES])
    os.replace(tmp, LIB)
    return LIB


def build_fortran_host():
    """Fortran90 hosts (fortran/host3d.f90, fortran/host2d.f90) over the C
    ABI via iso_c_binding."""
    fdir = os.path.join(HERE, "fortran")
    mod = os.path.join(fdir, "_mod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for m in ("fdtd_c", "fdtd2d_c"):
        obj = os.path.join(mod, m + ".o")
        _run([FLANG, "-O2", "-c", os.path.join(fdir, m + ".f90"), "-module-dir", mod, "-I", mod, "-o", obj])
        objs.append(obj)
    outs = []
    for host, deps in (("host3d", objs[:1]), ("host2d", objs)):
        out = os.path.join(fdir, host)
        _run([FLANG, "-O2", "-I", mod, os.path.join(fdir, host + ".f90"), *deps,
              "-L", HERE, "-lfdtd_amd", f"-Wl,-rpath,{HERE}", "-o", out])
        outs.append(out)
    return outs[0]


def build_oracles(with_reference=True):
    odir = os.path.join(REPO, "oracle")
    _run(["make", "-C", odir, "all"])
    if with_reference and os.path.isdir("/root/reference") and shutil.which(FLANG):
        _run(["make", "-C", odir, "ref"])
        _run(["make", "-C", odir, "ref2d"])
        if os.path.isfile("/opt/conda/include/mpif.h"):  # the 1D TFSF driver links MPICH
            _run(["make", "-C", odir, "ref1d"])


if __name__ == "__main__":
    print(build_library())
