"""Write tests/golden/ref2d_shipped.npz from the output files the reference
ships for its 2D case (FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate/,
a 3-rank run of program.f90: 125 x 1266 cells, 13696 steps).

The files are data the reference holds (its transforms/, timesLiney/ and
timesGnuplotmap/ outputs); this script only parses them.  Run in a checkout
that has the reference at /root/reference:

    python tests/golden/gen_golden2d.py
"""
import os

import numpy as np

SIM2 = "/root/reference/FDTDsimulations/2D/air-nanoAntennas_Si02-Sisubstrate"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref2d_shipped.npz")
N_Y_RANK = 422  # rows per rank (Ymax = 3)
LINE_FRAMES = (5, 10, 20)
MAP_FRAME, MAP_RANK = 20, 2


def load(rel):
    return np.loadtxt(os.path.join(SIM2, rel))


def main():
    out = {}
    for name in ("r0Syin", "r0Syr", "r2Syt"):
        out[name] = load(f"transforms/{name}.dat")            # lambda [nm], Sy
    out["rta"] = load("transforms/RTA_2Dfdtd.data")            # lambda, R, T, A (calR_Sy.f90)
    out["infld"] = load("transforms/r0infld.dat")              # n*dt, Ex, Hz at (62,14), n < tfinStop
    out["rfld"] = load("transforms/r0rfld.dat")                # same point, n > tfinStop
    out["tfld"] = load("transforms/r2tfld.dat")                # (62,1170), every step
    for f in LINE_FRAMES:                                      # column x = ic over all ranks
        parts = [load(f"timesLiney/r{r}E_time{f}.dat") for r in range(3)]
        out[f"line{f}"] = np.concatenate(parts, axis=0)        # y [nm], Ex, Hz
    m = load(f"timesGnuplotmap/r{MAP_RANK}E_time{MAP_FRAME}.dat")  # i, j (global), Hz
    out["map_ij"] = m[:, :2].astype(np.int32)
    out["map_hz"] = m[:, 2]
    out["map_frame"] = np.array(MAP_FRAME)
    out["line_frames"] = np.array(LINE_FRAMES)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
