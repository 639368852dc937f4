"""Golden vectors from a run of the reference's own shipped 3D program.

FDTDsimulations/3D/air-AuAntennasSiO2-Si/program.f90 is compiled where it
lies against the reference's 3D modules and MPICH (oracle/Makefile, target
sim3) and run as its run_local script runs it, under mpirun -np 6
(oracle/run_sim3.sh: 125 x 1266 x 25 cells, 13,696 steps, about 13 minutes on
this container's cores).  This script reads that run's output files and
writes sim3_run.npz:

  lam, Syin, Syr, Syt   transforms/r0Syin.dat, r0Syr.dat, r<rt>Syt.dat
                        (wavelength in nm, flux per wavelength; 100 rows)
  infld_head            transforms/r0infld.dat rows 1..800 (t, Ex..Hz)
  infld, rfld, tfld     every 4th row (from the first) of r0infld.dat,
                        r0rfld.dat and r<rt>tfld.dat, with the row's step n
                        as column 0 (then t = n*dt and Ex, Ey, Ez, Hx, Hy, Hz)
  timing                the run's wall time, ranks, rate and CPU (JSON text)

    python tests/golden/gen_golden_sim3.py [run folder]
"""
import io
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def _ftn(text):
    # Fortran Ew.d drops the exponent letter for three-digit exponents
    return re.sub(r"(?<=[0-9.])([+-][0-9]{3})(?![0-9])", r"E\1", text)


def _load(path):
    with open(path) as f:
        return np.loadtxt(io.StringIO(_ftn(f.read())), ndmin=2)


def main(run):
    T = os.path.join(run, "transforms")
    syt = [f for f in os.listdir(T) if f.endswith("Syt.dat")]
    tfl = [f for f in os.listdir(T) if f.endswith("tfld.dat")]
    assert len(syt) == 1 and len(tfl) == 1, (syt, tfl)
    rt = int(syt[0][1:-len("Syt.dat")])
    sin, sr, st = (_load(os.path.join(T, n)) for n in ("r0Syin.dat", "r0Syr.dat", syt[0]))
    assert sin.shape == sr.shape == st.shape == (100, 2)
    assert np.array_equal(sin[:, 0], sr[:, 0]) and np.array_equal(sin[:, 0], st[:, 0])
    inc, ref, tra = (_load(os.path.join(T, n)) for n in ("r0infld.dat", "r0rfld.dat", tfl[0]))
    timing = json.load(open(os.path.join(run, "timing.json")))
    steps = timing["steps"]
    tfin = len(inc) + 1  # infld holds n = 1 .. tfinStop-1
    assert len(ref) == steps - tfin and len(tra) == steps

    def every4(rows, n0):
        n = np.arange(n0, n0 + len(rows), dtype=np.float64)[:, None]
        return np.hstack([n, rows])[::4]

    out = os.path.join(HERE, "sim3_run.npz")
    np.savez_compressed(out, lam=sin[:, 0], Syin=sin[:, 1], Syr=sr[:, 1], Syt=st[:, 1], infld_head=inc[:800],
                        infld=every4(inc, 1), rfld=every4(ref, tfin + 1), tfld=every4(tra, 1),
                        tfin=np.int64(tfin), transmission_rank=np.int64(rt), timing=json.dumps(timing))
    print(out, os.path.getsize(out), "bytes; tfinStop", tfin, "Syt rank", rt)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(REPO, "oracle", "_ref", "sim3_run"))
