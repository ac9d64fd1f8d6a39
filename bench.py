"""Throughput of the LU-SGS iteration on one MI355X.

A step is one full LU-SGS iteration (ghost fill, spectral radii, local time
step, MUSCL + Roe residual, forward and backward hyperplane sweeps, update,
residual norm, section integrals and the CFL ramp) over the whole synthetic
grid, with the state resident in HBM.  Prints one JSON line.

The workload is named in ``config.workload``.  ``cpu_baseline`` times the
reference solver itself (compiled from its sources into oracle/_ref) on the
host, single-threaded: with more than one OpenMP thread the reference's face
loops race on a shared variable (source/Upwind_Convective.cpp:7,29) and give
wrong fluxes.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "3d-fluid-numerical-simulation-software_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

# Algorithmic HBM bytes per cell (FP64): every array a kernel semantically reads
# or writes counted once, derived quantities (W, |S|, unit normals) as
# recomputed, metrics as stored.
BYTES_PER_CELL = {
    "residual": 176,          # Q 5, S 9, nodes 3 -> R 5
    "sweep": 232,             # one sweep: R|dQ* 5, Q 5, lamc 3, dt 1, vol 1, S 9 -> dQ 5
    "sweep_ns": 256,          # the same plus lamv 3 (Navier-Stokes)
    "update": 128,            # Q 5, dQ 5, vol 1 -> Q 5
    "bc_spectral_dt": 152,    # Q 5, S 9, vol 1 -> lamc 3, dt 1
}
# HBM bytes one forward-sweep launch actually moves on a workload, from
# rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (copy-calibrated); the
# summaries are under profiles/.  Keyed by (ni, nj, nk, sweep form, viscous).
MEASURED_FWD_SWEEP_TRAFFIC = {(64, 64, 256, "rows", False): 472.5e6, (64, 64, 256, "rows", True): 506.0e6}
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ni", type=int, default=64)
    ap.add_argument("--nj", type=int, default=64)
    ap.add_argument("--nk", type=int, default=256)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--viscous", action="store_true", help="laminar Navier-Stokes (Reynolds number ~2e3 per unit length)")
    ap.add_argument("--sweep", choices=["rows", "kba3", "kba2", "kba", "plane"], default="rows")
    ap.add_argument("--chunk", type=int, default=1)
    ap.add_argument("--lag", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=40)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run three plain 1 GiB device copies and exit (scales the HBM byte counters)")
    return ap.parse_args()


def calibrate_copy():
    """Three device-to-device copies of 1 GiB each.  Under rocprofv3 --pmc the
    FETCH_SIZE / WRITE_SIZE readings of this copy kernel, against the 1 GiB it
    must read and write, give the factor that turns those counters into bytes."""
    import torch
    n = 1 << 27
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 3}))


def cpu_baseline(iters: int):
    """Reference solver on the C0 bump channel (96x32x32 cells), one thread."""
    import oracle as O
    from fv3d import cases
    if not os.path.exists(O.REF_BIN):
        return None
    case = cases.bump_channel(ni=32, nj=32, nk=96, seed=1, max_iter=iters)
    with tempfile.TemporaryDirectory() as td:
        cdir, odir = os.path.join(td, "case"), os.path.join(td, "out")
        case.write(cdir)
        r = O.run_reference(cdir, iters, odir, threads=1, dump=False, timeout=600)
    t = r["timing"]
    return {"value": t["cells"] * t["iterations"] / t["loop_seconds"], "unit": "cell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference solver (oracle/_ref, -O2, 1 thread) on the 96x32x32 bump channel, "
                      f"{t['iterations']} LU-SGS iterations, {t['loop_seconds']:.2f} s"}


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    from fv3d import cases
    from fv3d.solver import GpuSolver

    case = cases.bump_channel(ni=a.ni, nj=a.nj, nk=a.nk, seed=1, max_iter=a.steps, viscous=a.viscous,
                              len_ref=1e-4 if a.viscous else 1.0)
    g = GpuSolver(case, device=0,
                  debug={"plane": 2, "kba": 4, "kba2": 8, "kba3": 16, "rows": 0}[a.sweep] | (a.chunk << 8) | (a.lag << 16))
    ncell = g.n_cells
    graph = not a.no_graph
    g.iterate(a.warmup, graph=graph)
    g.sync()
    t0 = time.perf_counter()
    hist = g.iterate(a.steps, graph=graph)
    g.sync()
    t1 = time.perf_counter()
    secs = t1 - t0
    value = ncell * a.steps / secs
    # per-kernel device times from HIP events on the library's stream
    g.set_timing(True)
    g.iterate(max(5, min(a.steps, 20)), graph=False)
    ms, _ = g.timing()
    g.set_timing(False)
    classes = ["residual", "lusgs", "update", "bc_spectral_dt", "finalize"]
    times = {cl: ms[i] for i, cl in enumerate(classes)}
    if ms[5] > 0:
        # dominant kernel: the forward sweep, timed alone with HIP events on the
        # library's stream
        kname, kms = "k_lusgs_sl<forward> (LU-SGS forward sweep)", ms[5]
        kbytes = BYTES_PER_CELL["sweep_ns" if a.viscous else "sweep"]
        times["forward_sweep"], times["backward_sweep"] = ms[5], ms[6]
        traffic = MEASURED_FWD_SWEEP_TRAFFIC.get((a.ni, a.nj, a.nk, a.sweep, a.viscous))
    else:
        kname, kms = "LU-SGS forward+backward sweeps", ms[1]
        kbytes = 2 * BYTES_PER_CELL["sweep_ns" if a.viscous else "sweep"]
        traffic = None
    achieved = kbytes * ncell / (kms * 1e-3) / 1e9
    cpu = None if a.no_cpu_baseline else cpu_baseline(a.cpu_iters)
    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * secs / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{'laminar Navier-Stokes' if a.viscous else 'inviscid'} bump channel "
                               f"{a.nk}x{a.nj}x{a.ni} cells (single block), Roe + MUSCL3 "
                               f"primitive, LU-SGS local dt CFL 5, plain FP64",
                   "cells": ncell, "hip_graph": graph, "sweep": a.sweep, "chunk": a.chunk, "lag": a.lag},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_ms_avg": kms,
                     "bytes_per_cell": kbytes, "class_ms": times},
        "cpu_baseline": cpu,
        "final_res_dens": float(hist[-1, 1]),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
