"""The C-ABI library loads and exports every entry point include/fdtd.h
declares; the oracle exports the same set under its prefix; the Fortran
interface module binds only names the header declares."""
import ctypes as C
import os
import re

import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "fdtd.h")
HEADER2D = os.path.join(ROOT, "include", "fdtd2d.h")


def declared(header=HEADER, prefix="fdtd_"):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_2d_header_library_oracle_and_prototypes_agree():
    from fdtd_amd import abi2d

    names = declared(HEADER2D, "fdtd2d_")
    assert "fdtd2d_batch_run" in names and "fdtd2d_step" in names
    dll = C.CDLL(os.path.join(PKG, "libfdtd_amd.so"))
    assert not [n for n in names if not hasattr(dll, n)]
    odll = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    assert not ["o" + n for n in names if "batch" not in n and not hasattr(odll, "o" + n)]
    have = {"fdtd2d_" + n for n in list(abi2d.PROTOTYPES_2D) + list(abi2d.BATCH_PROTOTYPES_2D)}
    assert set(names) == have
    src = open(os.path.join(PKG, "fortran", "fdtd2d_c.f90")).read()
    bound = set(re.findall(r'bind\(c,\s*name="(fdtd2d_[a-z0-9_]+)"\)', src))
    assert bound and bound <= set(names)


def test_header_declares_the_abi():
    names = declared()
    for must in ("fdtd_create", "fdtd_destroy", "fdtd_set_pml", "fdtd_add_dielectric", "fdtd_add_drude2cp",
                 "fdtd_add_graphene", "fdtd_add_plane_source", "fdtd_add_dft_plane", "fdtd_add_probe",
                 "fdtd_step", "fdtd_get_field", "fdtd_get_dft", "fdtd_get_flux", "fdtd_get_probe",
                 "fdtd_last_error", "fdtd_energy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = os.path.join(PKG, "libfdtd_amd.so")
    if not os.path.exists(lib):
        pytest.fail("libfdtd_amd.so not built (run __graft_entry__.build())")
    dll = C.CDLL(lib)
    missing = [n for n in declared() if not hasattr(dll, n)]
    assert not missing, missing


def test_oracle_exports_the_same_abi():
    dll = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    skip = {"fdtd_memory_bytes", "fdtd_set_profiling", "fdtd_kernel_stats", "fdtd_algorithmic_bytes",
            "fdtd_cell_counts", "fdtd_estimate_bytes", "fdtd_autotune", "fdtd_get_flux_noalloc",
            "fdtd_checkpoint_size", "fdtd_checkpoint_save", "fdtd_checkpoint_load"}
    skip |= {n for n in declared() if n.startswith("fdtd_pool_")}
    missing = ["o" + n for n in declared() if n not in skip and not hasattr(dll, "o" + n)]
    assert not missing, missing


def test_fortran_interface_binds_declared_names():
    src = open(os.path.join(PKG, "fortran", "fdtd_c.f90")).read()
    bound = set(re.findall(r'bind\(c,\s*name="(fdtd_[a-z0-9_]+)"\)', src))
    assert bound and bound <= set(declared())


def test_python_prototypes_cover_the_header():
    from fdtd_amd import abi

    have = {"fdtd_" + n for n in list(abi.PROTOTYPES) + list(abi.GPU_PROTOTYPES)}
    assert set(declared()) == have


def test_struct_layouts_match_the_header():
    """Sizes of the ctypes mirrors equal the C sizes (compiled probe)."""
    import subprocess
    import tempfile

    from fdtd_amd import abi

    from fdtd_amd import abi2d

    code = ('#include "fdtd2d.h"\n#include <stdio.h>\nint main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",'
            "sizeof(fdtd_grid),sizeof(fdtd_pml),sizeof(fdtd_region),sizeof(fdtd_drude2cp),"
            "sizeof(fdtd_graphene),sizeof(fdtd_pulse),sizeof(fdtd2d_grid),sizeof(fdtd2d_pml),"
            "sizeof(fdtd2d_region),sizeof(fdtd_lorentz_jp),sizeof(fdtd_tfsf),sizeof(fdtd2d_tfsf));return 0;}\n")
    with tempfile.TemporaryDirectory() as t:
        src = os.path.join(t, "p.c")
        open(src, "w").write(code)
        exe = os.path.join(t, "p")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        sizes = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    ours = [C.sizeof(t) for t in (abi.Grid, abi.Pml, abi.Region, abi.Drude2CP, abi.Graphene, abi.Pulse,
                                  abi2d.Grid2D, abi2d.Pml2D, abi2d.Region2D, abi.LorentzJp, abi.Tfsf, abi2d.Tfsf2D)]
    assert sizes == ours


def test_kernel_ids_mirror_the_header():
    """The ctypes mirror's kernel ids and names (fdtd_kernel_stats,
    fdtd_algorithmic_bytes) follow the header's FDTD_K_* enum one for one."""
    from fdtd_amd import abi

    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    ids = {m.group(1): int(m.group(2)) for m in re.finditer(r"\bFDTD_K_([A-Z0-9_]+)\s*=\s*(\d+)", text)}
    assert ids["COUNT"] == abi.K_COUNT == len(abi.KERNEL_NAMES)
    for name, value in ids.items():
        assert getattr(abi, "K_" + name) == value, name
