"""Benchmark: Mcell-updates/s of the 3D Yee hot path at 512^3 (gold dimer,
12-cell split-field PML on all faces, 64-wavelength DFT on 3 planes, 8
probes), FP32 fields beside the plain FP64 stepper, on one MI355X.

One "step" is one full time step of the workload (E half step with the
dispersive cells and the PML, H half step, DFT monitors and probes) over the
whole 512^3 grid; inputs are resident in HBM before the timed region.

Prints ONE JSON line.  ``roofline`` is measured live: HIP events around every
launch of the dominant kernel (the fused E update) on the handle's stream,
its algorithmic bytes per launch divided by its mean duration.
``cpu_baseline`` times the FP64 CPU oracle (OpenMP) on 60 steps of the
same 512^3 workload on this host.  ``served_stream`` replays a seeded
Poisson stream of smaller gold jobs (64-320 cells a side) through the
library's job pool and reports throughput and p50/p99 job latency beside the
same jobs run one at a time.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fortran-mpi-fdtd_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s at 512³ gold+PML, FP32 vs FP64; %HBM roofline; p99 job latency"


def run_gpu(spec, precision, steps, warmup):
    import fdtd_amd
    from fdtd_amd import abi

    t0 = time.time()
    sim, ids, pids = fdtd_amd.make_simulation(spec, precision=precision)
    sim.commit()
    tuned = sim.autotune(3) if os.environ.get("FDTD_AUTOTUNE", "1") != "0" else None
    setup_s = time.time() - t0
    sim.step(warmup)
    sim.sync()
    sim.set_profiling(True)
    t1 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    elapsed = time.perf_counter() - t1
    nx, ny, nz = sim.shape
    cells = nx * ny * nz
    kern = {}
    for k in (abi.K_E, abi.K_H, abi.K_ADE, abi.K_MON, abi.K_GEN_E, abi.K_GEN_H, abi.K_XF):
        ms, n = sim.kernel_stats(k)
        if n:
            kern[abi.KERNEL_NAMES[k]] = {"mean_ms": ms / n, "launches": n,
                                         "bytes_per_launch": sim.algorithmic_bytes(k) if k != abi.K_MON else None}
    # the one-pass step (E and H of a step in one kernel) times everything
    # under the E timer; the two-kernel step has separate E and H kernels
    e = kern["e_update"]
    onepass = "h_update" not in kern
    ach = e["bytes_per_launch"] / (e["mean_ms"] * 1e-3) / 1e9
    h_upd = None
    if not onepass:
        h = kern["h_update"]
        ach_h = h["bytes_per_launch"] / (h["mean_ms"] * 1e-3) / 1e9
        h_upd = {"achieved_GBs": ach_h, "frac": ach_h / HBM_PEAK_GBS, "mean_launch_ms": h["mean_ms"]}
    out = {
        "value": cells * steps / elapsed / 1e6,
        "ms_per_step": elapsed * 1e3 / steps,
        "setup_s": setup_s,
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_march_plain" if onepass else "k_fused_e",
                     "bytes_per_launch": e["bytes_per_launch"], "mean_launch_ms": e["mean_ms"]},
        "h_update": h_upd,
        "kernels": kern,
        "step_form": None if tuned is None else ({"two_kernel_planes_per_tile": tuned[1]} if tuned[0] == 0 else
                                                 {"onepass_tile": [64, tuned[0]], "z_chunk": tuned[1]}),
        "device_bytes": sim.memory_bytes(),
        "cell_counts": dict(zip(("pml", "ade"), sim.cell_counts())),
    }
    flux = [float(v) for v in sim.flux(ids[2])[:4]]
    out["flux_sample"] = flux
    sim.close()
    return out


PROFILES = os.path.join(ROOT, "profiles")


def profiled_traffic(kernel, prec, summary):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    `summary` under profiles/ (FETCH_SIZE and WRITE_SIZE runs, each scaled
    by a 1 GiB copy's reading): (bytes, summary file) or (None, None)."""
    path = os.path.join(PROFILES, summary)
    try:
        lines = open(path).readlines()
    except OSError:
        return None, None
    for line in lines:
        w = line.split()
        if len(w) == 4 and w[0] == "traffic_bytes_per_launch" and w[1] == kernel and w[2] == prec:
            return float(w[3]), os.path.relpath(path, ROOT)
    return None, None


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (None where absent)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(spec, steps):
    """FP64 CPU oracle (OpenMP) on `steps` steps of the same workload."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    sim, _, _ = oracle.make_simulation(spec)
    sim.commit()
    sim.step(1)
    t0 = time.perf_counter()
    sim.step(steps)
    el = time.perf_counter() - t0
    nx, ny, nz = sim.shape
    sim.close()
    return {"value": nx * ny * nz * steps / el / 1e6, "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"FP64 oracle (fdtd_oracle.c, OpenMP) on the same {nx}x{ny}x{nz} workload, "
                      f"{steps} timed steps after 1 warm-up step ({el:.1f} s)"}


def served(args):
    """Served job stream: Poisson arrivals replayed in real time through the
    pool, then the same jobs one at a time (FP32, and the plain FP64 stepper)."""
    import numpy as np
    from fdtd_amd import serve

    jobs = serve.job_stream(seed=args.seed, njobs=args.jobs, rate_hz=args.rate)
    mix = serve.served_run(jobs, precision="fp32", max_active=args.max_active, chunk_steps=args.chunk)
    seq32 = serve.sequential_run(jobs, precision="fp32")
    seq64 = serve.sequential_run(jobs, precision="fp64") if not args.no_fp64 else None
    same = all(np.array_equal(a, b) for a, b in zip(mix["flux"], seq32["flux"]))
    out = {"jobs": len(jobs), "rate_hz": args.rate, "max_active": args.max_active, "chunk_steps": args.chunk,
           "seconds": mix["seconds"], "mcells_per_s": mix["mcells_per_s"], "p50_ms": mix["p50_ms"],
           "p99_ms": mix["p99_ms"], "sequential_fp32_mcells_per_s": seq32["mcells_per_s"],
           "speedup_vs_sequential_fp32": mix["mcells_per_s"] / seq32["mcells_per_s"],
           "results_identical_alone_and_mixed": bool(same)}
    if seq64:
        out["sequential_fp64_mcells_per_s"] = seq64["mcells_per_s"]
        out["speedup_vs_sequential_fp64"] = mix["mcells_per_s"] / seq64["mcells_per_s"]
    return out


def _batch_vs_separate(specs, steps, separate=True):
    """Time `steps` steps of `specs` stepped as one batch (fdtd2d_batch_run),
    then (unless `separate` is false) the same simulations as separate
    handles (one fdtd2d_step call per handle, each on its own stream, all
    enqueued before the first sync)."""
    import numpy as np
    from fdtd_amd import sim2d

    cells = sum(s["grid"]["nx"] * s["grid"]["ny"] for s in specs)
    work = cells * steps

    def build():
        sims = [sim2d.make_simulation2d(s, precision="fp32") for s in specs]
        for s, _, _ in sims:
            s.commit()
        return sims

    sims = build()
    batch = sim2d.Batch2D([s for s, _, _ in sims])
    batch.run([8] * len(sims))  # graph capture and first touch, untimed
    batch.sync()
    t0 = time.perf_counter()
    batch.run([steps] * len(sims))
    batch.sync()
    tb = time.perf_counter() - t0
    fb = [s.flux(ids[1]) for s, ids, _ in sims]
    batch.close()
    for s, _, _ in sims:
        s.close()
    # algorithmic words per main cell-update: TEz E 7 (Hz, Dx, Dy rw, Ex, Ey),
    # H 4; TMz E 5 (Hx, Hy, Dz rw, Ez), H 5 (Ez, Hx rw, Hy rw): 11 and 10
    words = sum((10 if s["grid"].get("polarization") == "tmz" else 11) * s["grid"]["nx"] * s["grid"]["ny"]
                for s in specs) / cells
    n_tmz = sum(1 for s in specs if s["grid"].get("polarization") == "tmz")
    out = {"instances": len(specs), "tmz_instances": n_tmz, "cells": cells, "steps": steps, "dtype": "f32",
           "batched_mcells_per_s": work / tb / 1e6, "achieved_GBs_all_in": work * words * 4 / tb / 1e9,
           "separate_handles_mcells_per_s": None, "speedup": None, "results_identical_batched_and_alone": None}
    if not separate:
        return out
    sims = build()
    for s, _, _ in sims:
        s.step(8)
        s.sync()
    t0 = time.perf_counter()
    for s, _, _ in sims:
        s.step(steps)
    for s, _, _ in sims:
        s.sync()
    ts = time.perf_counter() - t0
    same = all(np.array_equal(a, s.flux(ids[1])) for a, (s, ids, _) in zip(fb, sims))
    for s, _, _ in sims:
        s.close()
    out.update({"separate_handles_mcells_per_s": work / ts / 1e6, "speedup": ts / tb,
                "results_identical_batched_and_alone": bool(same)})
    return out


def batched_2d(args):
    """Batched 2D sweep: `--b2-count` TEz and TMz instances drawn from the
    seed (sides 256-2048 cells, gold disk r 10-80 nm, Kubo graphene line E_F
    0.1-0.6 eV, 16-cell PML, 32 wavelengths on 2 lines, FP32), all stepped
    `--b2-steps` steps in one batch, then as separate handles."""
    from fdtd_amd import sim2d

    specs = sim2d.sweep_stream(seed=args.seed, count=args.b2_count, nsteps=args.b2_steps)
    return _batch_vs_separate(specs, args.b2_steps, not args.b2_no_separate)


def batched_2d_small(args):
    """The same sweep draws at a fixed side of `--b2s-side` cells: instances
    too small to fill the GPU alone, where one batch replaces many launches."""
    from fdtd_amd import sim2d

    specs = sim2d.sweep_stream(seed=args.seed, count=args.b2_count, nsteps=args.b2s_steps, side=args.b2s_side)
    out = _batch_vs_separate(specs, args.b2s_steps, not args.b2_no_separate)
    out["side"] = args.b2s_side
    return out


def _kernel_breakdown(sim, steps):
    """Per-kernel mean launch time (HIP events on the handle's stream) over
    `steps` more steps, with each kernel's algorithmic bytes per launch."""
    from fdtd_amd import abi

    sim.set_profiling(True)
    sim.step(steps)
    sim.sync()
    kern = {}
    for k in range(abi.K_COUNT):
        ms, n = sim.kernel_stats(k)
        if n:
            # algorithmic bytes of one step (the ADE figure covers every
            # dispersive group, one launch each) over the kernel's time per step
            b = sim.algorithmic_bytes(k) if k != abi.K_MON else None
            kern[abi.KERNEL_NAMES[k]] = {"mean_ms": ms / n, "launches": n, "bytes_per_step": b,
                                         "ms_per_step": ms / steps,
                                         "frac": None if not b else b / (ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS}
    sim.set_profiling(False)
    return kern


def c3_mixed_precision(args, precisions=("fp32", "fp64")):
    """The 384x384x256 graphene-sheet + gold-disk case (Kubo graphene with
    the Raman term, 16 jittered disks, 12-cell PML, 128 wavelengths on two
    planes) in FP32 and in FP64 for `--c3-steps` steps: throughput of each,
    the per-kernel breakdown of 50 further FP32 steps, and the FP32
    spectra's error against the FP64 GPU run (which reproduces the FP64
    oracle to 1e-12 on the test cases; the tests hold reduced C3 FP32
    spectra to 1e-5 of the oracle itself)."""
    import numpy as np
    import fdtd_amd
    from fdtd_amd import workloads

    spec = workloads.c3_graphene_disks(seed=args.seed, nsteps=args.c3_steps)
    out = {"workload": spec["name"], "steps": args.c3_steps, "nfreq": len(spec["dft"][0]["omega"])}
    flux = {}
    for prec in precisions:
        sim, ids, _ = fdtd_amd.make_simulation(spec, precision=prec)
        sim.commit()
        tuned = sim.autotune(3) if os.environ.get("FDTD_AUTOTUNE", "1") != "0" else None
        sim.step(1)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(args.c3_steps - 1)
        sim.sync()
        el = time.perf_counter() - t0
        nx, ny, nz = sim.shape
        out[f"{prec}_mcells_per_s"] = nx * ny * nz * (args.c3_steps - 1) / el / 1e6
        out[f"{prec}_ms_per_step"] = el * 1e3 / (args.c3_steps - 1)
        flux[prec] = [sim.flux(i) for i in ids]
        if prec == "fp32":
            out["fp32_step_form"] = tuned
            out["fp32_cell_counts"] = dict(zip(("pml", "ade"), sim.cell_counts()))
            out["fp32_kernels"] = _kernel_breakdown(sim, 50)
        sim.close()
    if "fp64" in flux and "fp32" in flux:
        errs = [float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(flux["fp32"], flux["fp64"])]
        out["fp32_over_fp64"] = out["fp32_mcells_per_s"] / out["fp64_mcells_per_s"]
        out["fp32_spectra_max_rel_err_vs_gpu_fp64"] = errs
    return out


def fp32_spectra_error(args):
    """FP32 against FP64 GPU spectra, max over wavelengths and planes of
    |S32 - S64| / max|S64|: the 96^3 gold sphere (all-faces PML, 2000 steps,
    16 wavelengths) and the full 512^3 gold dimer run (4000 steps, 64
    wavelengths on 3 planes).  The FP64 GPU path reproduces the FP64 oracle
    to 1e-12 (bit for bit on the 96^3 sphere); the tests hold the FP32
    spectra of the full sphere and of reduced dimer and graphene cases to
    1e-5 of the oracle itself."""
    import numpy as np
    import fdtd_amd
    from fdtd_amd import workloads

    out = {}
    cases = [("gold_sphere_96", workloads.c1_gold_sphere())]
    if args.c2_spectra_steps > 0:
        cases.append((f"gold_dimer_{args.n}", workloads.c2_gold_dimer(seed=args.seed, n=args.n,
                                                                         nsteps=args.c2_spectra_steps)))
    for name, spec in cases:
        flux = {}
        for prec in ("fp64", "fp32"):
            sim, ids, _ = fdtd_amd.make_simulation(spec, precision=prec)
            sim.step(spec["nsteps"])
            flux[prec] = [sim.flux(i) for i in ids]
            sim.close()
        errs = [float(np.max(np.abs(a - b)) / np.max(np.abs(b))) for a, b in zip(flux["fp32"], flux["fp64"])]
        out[name] = {"steps": spec["nsteps"], "nfreq": len(spec["dft"][0]["omega"]), "max_rel_err": max(errs),
                     "per_plane": errs}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=60, help="timed oracle steps (about 15 s on 16 host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-2d", action="store_true")
    ap.add_argument("--no-spectra", action="store_true", help="skip the FP32-vs-FP64 spectra check")
    ap.add_argument("--no-c3", action="store_true", help="skip the graphene + gold-disk mixed-precision case")
    ap.add_argument("--c3-steps", type=int, default=6000)
    ap.add_argument("--only-c3", action="store_true",
                    help="run only the graphene + gold-disk case in FP32 and print its record (for profiling)")
    ap.add_argument("--c2-spectra-steps", type=int, default=4000,
                    help="steps of the 512^3 FP32-vs-FP64 spectra comparison (0: skip it)")
    ap.add_argument("--b2-count", type=int, default=256, help="batched 2D sweep: instances")
    ap.add_argument("--b2-steps", type=int, default=500, help="batched 2D sweep: timed steps")
    ap.add_argument("--b2s-side", type=int, default=128, help="small batched 2D sweep: grid side")
    ap.add_argument("--b2s-steps", type=int, default=1000, help="small batched 2D sweep: timed steps")
    ap.add_argument("--b2-no-separate", action="store_true",
                    help="batched 2D sweeps: time the batch only, not the separate handles")
    ap.add_argument("--jobs", type=int, default=24)
    ap.add_argument("--rate", type=float, default=20.0, help="served-stream arrival rate [jobs/s]")
    ap.add_argument("--max-active", type=int, default=4)
    ap.add_argument("--chunk", type=int, default=64)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run a plain device copy of this size (scales rocprof FETCH/WRITE counters)")
    args = ap.parse_args()
    if args.copy_calibration_mib:
        import torch

        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 2.0, out=dst)  # elementwise kernel: reads and writes exactly n*4 bytes
        torch.cuda.synchronize()
        del src, dst
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    from fdtd_amd import workloads

    if args.only_c3:
        print(json.dumps(c3_mixed_precision(args, precisions=("fp32",))))
        return
    spec = workloads.c2_gold_dimer(seed=args.seed, n=args.n)
    print(f"[bench] workload {spec['name']} gap {spec['gap_nm']:.3f} nm, fp32 {args.steps} steps", file=sys.stderr,
          flush=True)
    r32 = run_gpu(spec, "fp32", args.steps, args.warmup)
    print(f"[bench] fp32 {r32['value']:.0f} Mcell/s", file=sys.stderr, flush=True)
    r64 = None
    if not args.no_fp64:
        r64 = run_gpu(spec, "fp64", args.steps, args.warmup)
        print(f"[bench] fp64 {r64['value']:.0f} Mcell/s", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(spec, args.cpu_steps)
        print(f"[bench] cpu {cpu['value']:.1f} Mcell/s", file=sys.stderr, flush=True)
    srv = None
    if not args.no_serve:
        srv = served(args)
        print(f"[bench] served {srv['mcells_per_s']:.0f} Mcell/s p50 {srv['p50_ms']:.0f} ms p99 {srv['p99_ms']:.0f} ms",
              file=sys.stderr, flush=True)
    b2 = b2s = None
    if not args.no_2d:
        b2 = batched_2d(args)
        def vs(r):
            return "not timed" if r["speedup"] is None else f"{r['speedup']:.2f}x"

        print(f"[bench] batched 2d {b2['batched_mcells_per_s']:.0f} Mcell/s (separate handles: {vs(b2)})",
              file=sys.stderr, flush=True)
        b2s = batched_2d_small(args)
        print(f"[bench] batched 2d small {b2s['batched_mcells_per_s']:.0f} Mcell/s (separate handles: {vs(b2s)})",
              file=sys.stderr, flush=True)
    spec_err = None if args.no_spectra else fp32_spectra_error(args)
    if spec_err:
        print("[bench] fp32 spectra error " + ", ".join(f"{k} {v['max_rel_err']:.1e}" for k, v in spec_err.items()),
              file=sys.stderr, flush=True)
    c3 = None
    if not args.no_c3:
        c3 = c3_mixed_precision(args)
        print(f"[bench] c3 fp32 {c3['fp32_mcells_per_s']:.0f} fp64 {c3['fp64_mcells_per_s']:.0f} Mcell/s, "
              f"spectra err {max(c3['fp32_spectra_max_rel_err_vs_gpu_fp64']):.2e}", file=sys.stderr, flush=True)
    # the FP32 x-face tiles (plain tiles reaching into an x PML slab, a
    # launch of their own beside the dominant kernel): the same figures
    xf = r32["kernels"].get("x_face")
    xface = None
    if xf:
        ach = xf["bytes_per_launch"] / (xf["mean_ms"] * 1e-3) / 1e9
        xface = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                 "traffic": None, "kernel": "k_march_plain_xf", "bytes_per_launch": xf["bytes_per_launch"],
                 "mean_launch_ms": xf["mean_ms"]}
    if args.n == 512 and args.seed == 0:  # the workload the summary profiles
        tb, src = profiled_traffic(r32["roofline"]["kernel"], "fp32", "k_march_plain_gold_dimer_512.md")
        r32["roofline"]["traffic"] = tb
        if tb is not None:
            r32["roofline"]["traffic_source"] = src
        if xface:
            xface["traffic"], _ = profiled_traffic("k_march_plain_xf", "fp32", "k_march_plain_gold_dimer_512.md")
    line = {
        "metric": METRIC,
        "value": r32["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": spec["name"], "grid": [args.n] * 3, "ds_nm": 1.0, "pml_cells": 12,
                   "gap_nm": spec["gap_nm"], "seed": args.seed, "dft_planes": 3, "nfreq": 64, "probes": 8,
                   "engine": "fused", "dft_accumulation": "f64"},
        "roofline": r32["roofline"],
        "roofline_x_face_tiles": xface,
        "fp64": None if r64 is None else {"value": r64["value"], "ms_per_step": r64["ms_per_step"],
                                          "dtype": "f64", "roofline": r64["roofline"]},
        "fp32_over_fp64": None if r64 is None else r32["value"] / r64["value"],
        "fp32_spectra_error_vs_gpu_fp64": spec_err,
        "graphene_disks_mixed_precision": c3,
        "cpu_baseline": cpu,
        "p99_job_latency_ms": None if srv is None else srv["p99_ms"],
        "served_stream": srv,
        "batched_2d": b2,
        "batched_2d_small": b2s,
        "detail": {"fp32": {k: r32[k] for k in ("kernels", "h_update", "step_form", "device_bytes", "cell_counts",
                                                 "setup_s")},
                   "fp64": None if r64 is None else {k: r64[k] for k in ("kernels", "h_update", "step_form",
                                                                        "device_bytes")}},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
