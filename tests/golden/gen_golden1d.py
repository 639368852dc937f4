"""Golden vectors of the reference's 1D TFSF run.

Runs oracle/_ref/ref1d (built by ``make -C oracle ref1d`` from
FDTDsimulations/1D/TFSF-air_SLG_SiO2_Sidisp/program.f90 and the 1D modules,
in this development container only) under ``mpirun -np 4`` in a scratch
directory and stores what it wrote: the incident-line trace
(transforms/r0infld.dat: n*dt, ex_inc, hz_inc at monitor_pti), the
reflection trace (r0rfld.dat: n*dt, ex, hz at monitor_ptr), both spectra
(r0pump_Syin.dat, r0R_Syr.dat) and the parameters the run printed.  The
traces are kept at every 4th step (the spectra cover every step).

Usage: python tests/golden/gen_golden1d.py
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_EXE = os.path.join(REPO, "oracle", "_ref", "ref1d")
MPIRUN = os.environ.get("MPIRUN", "/opt/conda/bin/mpirun")


def main():
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("transforms", "timesLine"):
            os.makedirs(os.path.join(tmp, d))
        res = subprocess.run([MPIRUN, "-np", "4", REF_EXE], cwd=tmp, capture_output=True, text=True, timeout=600)
        if res.returncode != 0:
            raise RuntimeError(res.stdout + res.stderr)
        T = os.path.join(tmp, "transforms")
        inc = np.loadtxt(os.path.join(T, "r0infld.dat"))
        rfl = np.loadtxt(os.path.join(T, "r0rfld.dat"))
        syin = np.loadtxt(os.path.join(T, "r0pump_Syin.dat"))
        syr = np.loadtxt(os.path.join(T, "r0R_Syr.dat"))
    log = res.stdout

    def num(key):
        return int(re.search(r"\b" + key + r"=\s*(-?\d+)", log).group(1))

    params = np.array([num("Nt"), num("tfinStop"), num("spos"), num("TFSF_bd"), num("monitor_pti"),
                       num("monitor_ptr"), num("Npos"), num("yBDlow"), num("yBDhigh")], dtype=np.int64)
    out = os.path.join(HERE, "ref1d_tfsf.npz")
    np.savez_compressed(out, inc=inc[::4], rfl=rfl[::4], syin=syin, syr=syr, params=params, nrows=len(inc))
    print(out, os.path.getsize(out), "bytes; params", params.tolist())


if __name__ == "__main__":
    sys.exit(main())
