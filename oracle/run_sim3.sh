#!/bin/bash
# Runs the reference's shipped 3D program (built by `make sim3` into
# _ref/sim3) as its own run_local script does: mpirun -np 6, from a folder
# holding transforms/ and timesLiney/.  TEST INFRASTRUCTURE ONLY.
#
#   oracle/run_sim3.sh [run folder, default oracle/_ref/sim3_run]
#
# Prints the wall time of the whole mpirun and the rate in cell-updates/s
# (125 x 1266 x 25 cells x 13,696 steps; the set-up before the time loop
# takes well under a second of it).
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
RUN="${1:-$HERE/_ref/sim3_run}"
MPIRUN="${MPIRUN:-/opt/conda/bin/mpirun}"
NP=6  # Xmax*Ymax*Zmax of program.f90, fixed at compile time
[ -x "$HERE/_ref/sim3" ] || make -C "$HERE" sim3
mkdir -p "$RUN/transforms" "$RUN/timesLiney"
cd "$RUN"
t0=$(date +%s.%N)
"$MPIRUN" -np $NP "$HERE/_ref/sim3" > out.out 2>&1
t1=$(date +%s.%N)
python3 - "$t0" "$t1" "$NP" <<'EOF'
import json, os, platform, sys
t0, t1, np_ = float(sys.argv[1]), float(sys.argv[2]), int(sys.argv[3])
cells, steps = 125 * 1266 * 25, 13696
model = ""
for line in open("/proc/cpuinfo"):
    if line.startswith("model name"):
        model = line.split(":", 1)[1].strip()
        break
rec = {"seconds": t1 - t0, "ranks": np_, "cells": cells, "steps": steps,
       "mcells_per_s": cells * steps / (t1 - t0) / 1e6, "cpu_model": model,
       "host_cpus": os.cpu_count(), "machine": platform.machine()}
json.dump(rec, open("timing.json", "w"))
print(json.dumps(rec))
EOF
