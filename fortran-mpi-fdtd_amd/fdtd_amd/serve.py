"""Served job stream: continuous batching of independent 3D jobs on one GPU.

``Pool`` wraps the library's fdtd_pool (HBM-budget admission, one stream per
job, chunked progress).  ``job_stream`` draws the served workload from a
seed: Poisson arrivals of gold-nanoparticle jobs, grid side log-uniform in
64-320 cells (ds 1-2 nm), 10-cell PML, a sphere, rod or dimer with feature
sizes 10-60 nm, 1500-5000 steps, DFT at 16-64 wavelengths, FP32 fields with
FP64 DFT accumulation.  ``served_run`` replays the stream in real time
through a pool; ``sequential_run`` runs the same jobs one at a time.
"""
import ctypes as C
import math
import time

import numpy as np

from . import NANO, box, cylinder, gold_params_dict, library, make_simulation, sphere, wavelength_frequencies
from . import abi
from .abi import EX, EY, HX, HY
from .workloads import COURANT_3D, pulse_for_band


class Pool:
    def __init__(self, budget_bytes, max_active=16, chunk_steps=64, binding=None):
        self.b = binding or library()
        self.h = C.c_void_p()
        self.b.check(self.b.pool_create(C.byref(self.h), int(budget_bytes), int(max_active), int(chunk_steps)))

    def submit(self, sim, nsteps):
        t = C.c_int()
        self.b.check(self.b.pool_submit(self.h, sim.h, int(nsteps), C.byref(t)))
        return t.value

    def progress(self):
        r, w = C.c_int(), C.c_int()
        self.b.check(self.b.pool_progress(self.h, C.byref(r), C.byref(w)))
        return r.value, w.value

    def status(self, ticket):
        st, wm, rm = C.c_int(), C.c_double(), C.c_double()
        rc = self.b.pool_status(self.h, int(ticket), C.byref(st), C.byref(wm), C.byref(rm))
        return st.value, wm.value, rm.value, rc

    def cancel(self, ticket):
        self.b.check(self.b.pool_cancel(self.h, int(ticket)))

    def release(self, ticket):
        self.b.check(self.b.pool_release(self.h, int(ticket)))

    def usage(self):
        u, b = C.c_size_t(), C.c_size_t()
        self.b.check(self.b.pool_usage(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def close(self):
        if self.h:
            self.b.pool_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_spec(rng, pml=10, side=None, geom=None, nsteps=None):
    """One served job (draws in this order: side, ds, geometry, feature size,
    steps, nfreq).  `side`, `geom` and `nsteps` replace their draws (which
    are still taken, so the other draws stay those of the stream)."""
    side_d = int(round(math.exp(rng.uniform(math.log(64), math.log(320)))))
    side = side_d if side is None else int(side)
    ds = rng.uniform(1.0, 2.0) * NANO
    geom_d = ("sphere", "rod", "dimer")[int(rng.integers(0, 3))]
    geom = geom_d if geom is None else geom
    feat = rng.uniform(10.0, 60.0) * NANO
    nsteps_d = int(rng.integers(1500, 5001))
    nsteps = nsteps_d if nsteps is None else int(nsteps)
    nfreq = int(rng.integers(16, 65))
    n = side
    c = (n + 1) / 2.0
    room = (n - 2 * pml - 6) / 2.0  # cells from the centre to the PML, with margin
    f = min(feat / ds, 2 * room)
    mats = [{"type": "dielectric", "region": box((1, 1, 1), (n, n, n)), "eps": 1.0}]
    gold = gold_params_dict()
    if geom == "sphere":
        mats.append({"type": "drude2cp", "region": sphere((c, c, c), f / 2), "params": gold})
    elif geom == "rod":
        mats.append({"type": "drude2cp", "region": cylinder((c, c, c), max(f / 6, 1.0), f / 2, 0), "params": gold})
    else:
        r = f / 4
        g = max(f / 10, 1.0)
        for sgn in (-1, 1):
            mats.append({"type": "drude2cp", "region": sphere((c + sgn * (r + g / 2), c, c), r), "params": gold})
    pul = pulse_for_band(400e-9, 900e-9, 600e-9)
    om = list(wavelength_frequencies(400e-9, 900e-9, nfreq))
    return {
        "name": f"served_{geom}_{n}",
        "grid": {"nx": n, "ny": n, "nz": n, "ds": ds, "courant": COURANT_3D, "periodic": [0, 0, 0]},
        "pml": {"mode": "all_faces", "npml": [pml, pml, pml], "order": 3.0, "R0": 1e-6},
        "materials": mats,
        "sources": [dict(axis=2, index=pml + 2, comp=EX, amplitude=1.0, **pul)],
        "dft": [{"axis": 2, "index": n - pml - 2, "omega": om, "comps": [EX, EY, HX, HY]}],
        "probes": [[int(c), int(c), n - pml - 3]],
        "nsteps": nsteps,
    }


def job_stream(seed=0, njobs=24, rate_hz=20.0):
    """Poisson arrivals (exponential gaps at rate_hz) of served jobs."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(njobs):
        t += float(rng.exponential(1.0 / rate_hz))
        out.append((t, job_spec(rng)))
    return out


def _work(spec):
    g = spec["grid"]
    return g["nx"] * g["ny"] * g["nz"] * spec["nsteps"]


def _prepare(spec, precision, tune):
    """Handle of a job; with `tune`, fdtd_autotune picks its step form (the
    two-kernel step or one-pass tiles) from two timed steps of each, which
    give the same bits."""
    sim, ids, pids = make_simulation(spec, precision=precision)
    if tune:
        sim.autotune(2)
    return sim, ids, pids


def sequential_run(jobs, precision="fp32", tune=True):
    """Each job alone, back to back: create, tune, step all, read the flux."""
    t0 = time.perf_counter()
    res = []
    for _, spec in jobs:
        sim, ids, _ = _prepare(spec, precision, tune)
        sim.step(spec["nsteps"])
        res.append(sim.flux(ids[0]))
        sim.close()
    el = time.perf_counter() - t0
    work = sum(_work(s) for _, s in jobs)
    return {"seconds": el, "mcells_per_s": work / el / 1e6, "flux": res}


def served_run(jobs, precision="fp32", budget_bytes=200 << 30, max_active=16, chunk_steps=64, realtime=True,
               tune=True):
    """Replay the stream: submit each job at its arrival time (created and
    tuned as sequential_run does), drive the pool, read each result as soon
    as its job is done."""
    pool = Pool(budget_bytes, max_active, chunk_steps)
    t0 = time.perf_counter()
    live = {}
    lat = [None] * len(jobs)
    flux = [None] * len(jobs)
    nxt = 0
    while nxt < len(jobs) or live:
        now = time.perf_counter() - t0
        while nxt < len(jobs) and (not realtime or jobs[nxt][0] <= now):
            spec = jobs[nxt][1]
            sim, ids, _ = _prepare(spec, precision, tune)
            arrival = t0 + (jobs[nxt][0] if realtime else now)
            live[pool.submit(sim, spec["nsteps"])] = (nxt, sim, ids, arrival)
            nxt += 1
        pool.progress()
        for t in list(live):
            st, _, _, rc = pool.status(t)
            if st in (abi.JOB_DONE, abi.JOB_FAILED, abi.JOB_CANCELLED):
                k, sim, ids, arrival = live.pop(t)
                if st != abi.JOB_DONE:
                    raise RuntimeError(f"served job {k} ended in state {st} (rc {rc})")
                flux[k] = sim.flux(ids[0])
                lat[k] = (time.perf_counter() - arrival) * 1e3
                pool.release(t)
                sim.close()
    el = time.perf_counter() - t0
    pool.close()
    work = sum(_work(s) for _, s in jobs)
    lat_a = np.array(lat)
    return {"seconds": el, "mcells_per_s": work / el / 1e6, "p50_ms": float(np.percentile(lat_a, 50)),
            "p99_ms": float(np.percentile(lat_a, 99)), "latency_ms": lat, "flux": flux}
