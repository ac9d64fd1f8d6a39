"""Throughput and time to a 1e-6 residual drop of the solver on one MI355X:
the chosen form against the plain FP64 drop-in.  Prints one JSON line.

Default workload (the learned-closure passage): a five-block annular passage
(hub 0.2 m, tip 0.25 m, 22 blades) with a turning vane in the middle block,
88 x 88 cells in cross-section and 24/48/144/72/40 along the axis
(2,540,032 cells), Navier-Stokes with Spalart-Allmaras, Roe + MUSCL3 on
primitive variables, LU-SGS with local time step at CFL 5, multigrid from the
uniform field on the coarse level, stationary, outlet at 0.95 Pt.

* plain FP64: the reference's method on it, FP64 throughout, its two-grid
  cycle (the reference's maximum), Spalart-Allmaras transport;
* chosen: FP32 defect correction (FP32 LU-SGS operator over the FP64
  residual, update and norms), the four-level V-cycle, and all blocks of
  every stage in one launch captured in a HIP graph with the residual tiles
  the autotuner picks on this workload (fixed, so that the counter profile
  applies; ``--autotune`` picks them per run), and the learned eddy-viscosity
  closure in place of the nu~ transport
  (the network of fv3d/closure_vane.npz written by ``--workload
  closure-train``, its mue_t under-relaxed by 0.2, the Spalart-Allmaras value
  of the frozen nu~ as the gated fallback).  ``--no-closure`` runs the chosen
  form with the Spalart-Allmaras transport instead; by default the drop runs
  include that variant too.

A step is one fine-level iteration (with multigrid: fine LU-SGS, the second
fine residual, the coarse levels and the prolonged correction) over the whole
grid with the state resident in HBM; ``value`` is the chosen form's
cell-updates/s (fine cells x iterations / s).  ``time_to_1e-6_drop`` runs both
forms from the uniform start until Res.dens (the reference's density residual,
source/ComputeAllBlockVars.cpp:340-363) has fallen to 1e-6 of its first
fine-level value and compares mass flow, pressure ratio and efficiency
(source/Ouput.cpp:282-300) there.

Other workloads: ``c2`` (64 independent passages of 0.1-2M cells batched in
one context), ``passage5`` (the 1.08M-cell Baldwin-Lomax passage),
``bump`` (the single-block bump channel), ``closure-train`` (the closure's
training and held-out evaluation; writes the network).

``cpu_baseline`` times the reference solver itself (compiled from its sources
into oracle/_ref) on the host, single-threaded, on a 64,512-cell member of
the same passage family: with more than one OpenMP thread the reference's face
loops race on a shared variable (source/Upwind_Convective.cpp:7,29).
``cpu_baseline_all_cores`` times the C restatement of the same solver (the
test oracle, bit-identical to the reference at one thread) on every OpenMP
thread on the same sample.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "3d-fluid-numerical-simulation-software_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(REPO, "oracle"))

METRIC = "cell-updates/s + time to 1e-6 residual drop, chosen vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0
PREC = {"fp64": 0, "fp32": 1}
CLOSURE_WEIGHTS = os.path.join(PKG, "fv3d", "closure_vane.npz")
PROFILES = os.path.join(REPO, "profiles")

# Timing slots of fv3d_timing (include/fv3d.h)
SLOTS = ["residual", "lusgs", "update", "bc_spectral_dt", "finalize", "forward_sweep", "backward_sweep",
         "coarse_correction", "convective", "viscous", "bc_exchange", "turbulence"]


def kernel_bytes(precision: str, sa_transport: bool):
    """Algorithmic HBM bytes per cell of one launch: every array a kernel
    semantically reads or writes counted once (derived quantities such as W,
    |S| and unit normals count as recomputed), FP64 words of 8 B, the sweep
    operator's words in its own precision."""
    w = 4 if precision == "fp32" else 8
    nu = 4 if sa_transport else 0    # nu~, its residual, its diagonal, dnu~
    return {
        # Q 5, S_I,J,K 9, nodes 3 -> R 5
        "convective": 22 * 8,
        # Q 5, mue_t 1, face coefficients 27, S 9 -> R 4 read and written (the
        # continuity component of the viscous flux is zero, R's density stays)
        "viscous": 50 * 8,
        # R 5, Q 5, lambda_c 3, dt 1, Vol 1, S 9, lambda_v 3 -> dQ* 5 (+ nu~)
        "forward_sweep": (32 + nu) * w,
        # dQ* 5, Q 5, lambda_c 3, dt 1, Vol 1, S 9, lambda_v 3 -> dQ 5 (+ nu~)
        "backward_sweep": (32 + nu) * w,
    }


KERNEL_NAMES = {"convective": "k_convective (MUSCL + Roe/Steger-Warming + wall flux + source -> R)",
                "viscous": "k_viscous (face gradients + laminar and eddy viscous flux -> R)",
                "forward_sweep": "k_lusgs_sl<forward> (LU-SGS forward wavefront sweep)",
                "backward_sweep": "k_lusgs_sl<backward> (LU-SGS backward wavefront sweep)"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["c3", "c2", "c4", "passage5", "bump", "closure-train"], default="c3")
    ap.add_argument("--levels", type=int, default=4, help="grid levels of the chosen form's V-cycle")
    ap.add_argument("--no-closure", action="store_true",
                    help="chosen form with the Spalart-Allmaras transport instead of the learned closure "
                         "(fv3d/closure_vane.npz)")
    ap.add_argument("--no-chosen-sa", action="store_true",
                    help="skip the chosen form's Spalart-Allmaras drop run beside the closure's")
    ap.add_argument("--instances", type=int, default=64, help="c2: number of passages")
    ap.add_argument("--batch-seed", type=int, default=3)
    ap.add_argument("--batch-rpm", type=float, default=0.0,
                    help="c2: design rotation speed (the synthetic passage is stationary by design)")
    ap.add_argument("--compare-alone", action="store_true", help="c2: also time every passage alone")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default residual tile shapes")
    ap.add_argument("--autotune", action="store_true",
                    help="vane passage: pick the residual tile shapes on the running state instead of the fixed "
                         "configuration the counter profile measured")
    ap.add_argument("--tiles", default=None,
                    help="fixed configuration numbers C,V,S of the convective, viscous and nu~ residual kernels "
                         "in place of the autotuner (GpuSolver.TILE_SHAPES / STREAM_SHAPES numbering)")
    ap.add_argument("--ni", type=int, default=64, help="bump: cells along i (span)")
    ap.add_argument("--nj", type=int, default=64, help="bump: cells along j")
    ap.add_argument("--nk", type=int, default=256, help="bump: cells along k (streamwise)")
    ap.add_argument("--viscous", action="store_true", help="bump: laminar Navier-Stokes")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-drop", action="store_true", help="skip the time-to-1e-6-drop runs")
    ap.add_argument("--no-plain", action="store_true", help="c3/c4: time the chosen form only")
    ap.add_argument("--plain-only", action="store_true",
                    help="c3/c4: time the plain FP64 form and its drop only (a long plain run in a call of its own)")
    ap.add_argument("--drop-max-iter", type=int, default=20000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-converged", action="store_true",
                    help="skip the converged-field error of both forms on C0 against the reference state")
    ap.add_argument("--cpu-iters", type=int, default=24)
    ap.add_argument("--train-instances", type=int, default=56, help="closure-train: passages (48 train + 8 held out)")
    ap.add_argument("--train-develop", type=int, default=2500, help="closure-train: V-cycle iterations per passage")
    ap.add_argument("--closure-out", default=None, help="closure-train: where to write the network")
    ap.add_argument("--no-heldout", action="store_true", help="closure-train: train only")
    ap.add_argument("--heldout-max-iter", type=int, default=6000, help="closure-train: iteration cap per held-out run")
    ap.add_argument("--closure-relax", type=float, default=0.2, help="under-relaxation of the closure's mue_t")
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


def host_cpu():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


CONVERGED_C0 = os.path.join(REPO, "tests", "golden", "c0_converged.npz")


def converged_field_error():
    """Max and L2 error (relative to each field's largest magnitude) of every
    converged conservative field of both GPU forms on C0 (the 96x32x32 bump
    channel, single grid) against the reference's FP64 converged state
    (tests/golden/c0_converged.npz, written by the C restatement, which is the
    compiled reference bit for bit), after the same number of iterations."""
    import numpy as np
    from fv3d import cases
    from fv3d.solver import GpuSolver
    if not os.path.exists(CONVERGED_C0):
        return None
    z = np.load(CONVERGED_C0)
    meta = json.loads(bytes(z["meta"]).decode())
    n = int(meta["iterations"])
    ref = z["cv"]
    out = {"workload": "C0 bump channel 96x32x32, Euler, Roe + MUSCL3 primitive, LU-SGS local dt CFL 5, single grid",
           "iterations": n, "reference_drop": meta["drop"]}
    for form, prec in (("plain_fp64", "fp64"), ("chosen_fp32_operator", "fp32")):
        case = cases.bump_channel(max_iter=n + 10, **meta["kwargs"])
        g = GpuSolver(case, device=0, precision=PREC[prec])
        done = 0
        while done < n:
            k = min(5000, n - done)
            g.iterate(k, graph=True)
            done += k
        cv = g.get(1, "cv")[1:-1, 1:-1, 1:-1]
        g.close()
        errs = {}
        # scales: the largest density, momentum magnitude |rho V| (for all three
        # components: rho u of the spanwise-uniform channel is round-off-sized)
        # and total energy
        mom = float(np.sqrt((ref[..., 1:4] ** 2).sum(axis=-1)).max())
        scales = (float(np.abs(ref[..., 0]).max()), mom, mom, mom, float(np.abs(ref[..., 4]).max()))
        for q, name in enumerate(("rho", "rho_u", "rho_v", "rho_w", "rho_E")):
            d = cv[..., q] - ref[..., q]
            errs[name] = {"max": float(np.abs(d).max()) / scales[q], "l2": float(np.sqrt(np.mean(d * d))) / scales[q]}
        out[form] = errs
        print(f"bench: converged C0 {form}: max error {max(e['max'] for e in errs.values()):.2e}",
              file=sys.stderr, flush=True)
    return out


def cpu_sample(wkey: str, iters: int):
    """A bounded member of the workload's family for the CPU legs, and its
    description: the vane passage at a 24 x 24 cross-section (64,512 cells), or
    the stage at 24 x (24, 18) cells (56,736 cells), Spalart-Allmaras, the
    reference's two-grid cycle."""
    from fv3d import cases
    if wkey == "c4":
        c = cases.rotor_stator_stage(ni=24, nj_rotor=24, nj_stator=18, rotor_nk=(4, 6, 8, 16, 8, 6, 4),
                                     stator_nk=(4, 6, 10, 18, 12, 8, 4), max_iter=iters)
        return c, f"{c.n_cells:,}-cell rotor + stator stage of the same family (24 x (24, 18) cross-section)"
    c = cases.vane_passage(ni=24, nk_blocks=(8, 16, 48, 24, 16), levels=0, max_iter=iters)
    return c, f"{c.n_cells:,}-cell vane passage of the same family (24x24 cross-section)"


def cpu_baseline(iters: int, wkey: str = "c3"):
    """The compiled reference solver, one thread, on the workload's CPU sample
    (Spalart-Allmaras, two-grid)."""
    import oracle as O
    if not os.path.exists(O.REF_BIN):
        return None
    case, desc = cpu_sample(wkey, iters)
    with tempfile.TemporaryDirectory() as td:
        cdir, odir = os.path.join(td, "case"), os.path.join(td, "out")
        case.write(cdir)
        r = O.run_reference(cdir, iters, odir, threads=1, dump=False, timeout=600)
    t = r["timing"]
    model, ncpu = host_cpu()
    return {"value": t["cells"] * t["iterations"] / t["loop_seconds"], "unit": "cell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"reference solver (oracle/_ref, g++ -O2, 1 thread) on a {desc}, Spalart-Allmaras, "
                      f"two-grid cycle, {t['iterations']} iterations in {t['loop_seconds']:.2f} s; host {model}, "
                      f"{ncpu} logical CPUs visible"}


def cpu_all_cores(iters: int, wkey: str = "c3"):
    """The C restatement (oracle/, the checker of the GPU path) on every thread
    OpenMP is given, on the same 64,512-cell sample: its cell and face loops
    and the cells of each LU-SGS hyperplane run in parallel, with the same bits
    as one thread.  The reference itself cannot run multi-threaded (its
    convective face loops share one flag variable across OpenMP threads)."""
    import oracle as O
    case, desc = cpu_sample(wkey, iters)
    threads = O.set_threads(0)
    o = O.Oracle(case)
    t0 = time.perf_counter()
    o.iterate(iters)
    secs = time.perf_counter() - t0
    del o
    model, ncpu = host_cpu()
    return {"value": case.n_cells * iters / secs, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"C restatement (oracle/fv3d_oracle.c, gcc -O2 -fopenmp, {threads} threads) on the same "
                      f"{desc}, {iters} iterations in {secs:.2f} s; host {model}"}


def load_closure(relax: float = 1.0):
    import numpy as np
    if not os.path.exists(CLOSURE_WEIGHTS):
        return None
    z = np.load(CLOSURE_WEIGHTS)
    return {"params": z["params"].astype(np.float32), "bounds": z["bounds"].astype(np.float32),
            "meta": json.loads(bytes(z["meta"]).decode()) if "meta" in z else {}, "relax": relax}


def solver(case, precision, closure=None):
    from fv3d.solver import GpuSolver
    g = GpuSolver(case, device=0, precision=PREC[precision])
    if closure is not None:
        g.mlp_set(closure["params"], closure["bounds"])
        g.mlp_closure(True)
        g.mlp_relax(closure.get("relax", 1.0))
    return g


def time_form(case, precision, a, closure=None):
    """Whole-job throughput of one form and its per-kernel device times."""
    from fv3d.solver import GpuSolver
    g = solver(case, precision, closure)
    graph = not a.no_graph
    tiles = None
    if a.tiles:
        sel = [int(x) for x in a.tiles.split(",")]
        g.set_tiles(sel)
        tiles = [GpuSolver.conv_shape(sel[0])] + [GpuSolver.TILE_SHAPES[q] for q in sel[1:]]
        g.iterate(a.warmup, graph=graph)
    elif a.no_autotune or a.warmup < 2:
        g.iterate(a.warmup, graph=graph)
    else:
        # residual tile shapes picked on the running state, then one more
        # warm-up iteration (the graphs are captured again)
        g.iterate(a.warmup - 1, graph=graph)
        sel = g.autotune(3)
        tiles = [GpuSolver.conv_shape(sel[0])] + [GpuSolver.TILE_SHAPES[q] if q >= 0 else None for q in sel[1:]]
        g.iterate(1, graph=graph)
    g.sync()
    t0 = time.perf_counter()
    g.iterate(a.steps, graph=graph)
    g.sync()
    secs = time.perf_counter() - t0
    ncell = g.n_cells
    # per-kernel device times from HIP events on the library's stream (the
    # only stream its kernels run on), in a second, eager pass
    g.set_timing(True)
    g.iterate(max(5, min(a.steps, 10)), graph=False)
    ms, _ = g.timing()
    g.set_timing(False)
    g.close()
    return {"value": ncell * a.steps / secs, "ms_per_step": 1e3 * secs / a.steps, "cells": ncell,
            "class_ms": {cl: float(ms[i]) for i, cl in enumerate(SLOTS)}, "tiles": tiles}


def time_to_drop(case, precision, target, max_iter, closure=None, chunk=25, tiles=None, near=4.0):
    """Wall time and iterations until Res.dens falls to ``target`` times the
    first fine-level value (iteration coarse_iter + 1 with multigrid), with
    graph replay; the host reads the history once per chunk, and after every
    iteration once the residual has come within ``near`` times the target, so
    that the clock stops at the crossing rather than at the end of its chunk.
    Returns the history row of the iteration that reached it.  tiles: "C,V,S"
    residual configurations (None: the defaults)."""
    import numpy as np
    skip = case.cfg.coarse_iter if case.cfg.use_multigrid.upper() == "Y" else 0
    g = solver(case, precision, closure)
    if tiles:
        g.set_tiles([int(x) for x in tiles.split(",")])
    h = g.iterate(skip + 1, graph=True)   # captures the graphs; the last residual is the reference level
    r0 = float(h[-1, 1])
    g.sync()
    t0 = time.perf_counter()
    it, hit, row = skip + 1, None, None
    last = t0
    step = chunk
    while it < max_iter and hit is None:
        h = g.iterate(step, graph=True)
        k = np.nonzero(h[:, 1] <= target * r0)[0]
        if len(k):
            hit = it + int(k[0]) + 1
            row = h[k[0]]
        it += step
        if float(np.min(h[:, 1])) <= near * target * r0:
            step = 1
        if time.perf_counter() - last > 60.0:   # a progress line a minute on long runs
            last = time.perf_counter()
            print(f"bench: drop run at {it} iterations, Res.dens / first {h[-1, 1] / r0:.3e}", file=sys.stderr,
                  flush=True)
    g.sync()
    secs = time.perf_counter() - t0
    tripped = g.mlp_tripped() if closure is not None else None
    # every cycle iteration evaluates the closure once on every fine cell (its
    # first residual; the one the restriction takes keeps that mue_t); the
    # coarse start-up iterations do not evaluate it
    evals = g.n_cells * (it - skip) if closure is not None else None
    g.close()
    val = lambda col: float(row[col]) if row is not None else None
    return {"iterations": hit, "seconds": secs if hit else None, "res0": r0, "mass_in": val(2), "mass_out": val(3),
            "pressure_ratio": val(4), "efficiency": val(5), "gate_fallbacks": tripped, "closure_evaluations": evals,
            "gate_fallback_fraction": tripped / evals if evals else None,
            "note": None if hit else f"not reached in {max_iter} iterations"}


def traffic_from_profiles(kernel: str, wkey: str):
    """HBM bytes per launch of a kernel on a workload, from the 'traffic:'
    lines of the rocprofv3 summaries in profiles/ (FETCH_SIZE and WRITE_SIZE,
    copy-calibrated)."""
    if not os.path.isdir(PROFILES):
        return None
    for name in sorted(os.listdir(PROFILES)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                # traffic: <kernel> <workload> <bytes>
                p = line.strip().strip("`").split()
                if len(p) == 4 and p[0] == "traffic:" and p[1] == kernel and p[2] == wkey:
                    return float(p[3])
    return None


def valu_from_profiles(kernel: str, wkey: str):
    """VALU wave instructions and FP64 flop per launch of a kernel on a
    workload, from the 'valu:' lines of the rocprofv3 summaries in profiles/."""
    if not os.path.isdir(PROFILES):
        return None
    for name in sorted(os.listdir(PROFILES)):
        if not name.endswith(".md"):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                # valu: <kernel> <workload> <wave instructions> <fp64 flop>
                p = line.strip().split()
                if len(p) == 5 and p[0] == "valu:" and p[1] == kernel and p[2] == wkey:
                    return float(p[3]), float(p[4])
    return None


# one VALU wave instruction per SIMD per 4 cycles: 256 CUs x 64 lanes x 2.4 GHz
VALU_LANE_OPS = 256 * 64 * 2.4e9
FP64_PEAK_TFLOPS = 78.6


def roofline(r, precision, sa_transport, wkey):
    byt = kernel_bytes(precision, sa_transport)
    ks = {}
    for k, b in byt.items():
        ms = r["class_ms"].get(k, 0.0)
        if ms > 0:
            ach = b * r["cells"] / (ms * 1e-3) / 1e9
            ks[k] = {"kernel": KERNEL_NAMES[k], "ms": ms, "bytes_per_cell": b, "achieved": ach,
                     "frac": ach / HBM_PEAK_GBS, "traffic": traffic_from_profiles(k, wkey)}
            v = valu_from_profiles(k, wkey)
            if v:
                # issue time of the counted VALU instructions at the full rate
                # (transcendentals not weighted here) against the measured time
                ks[k]["valu"] = {"wave_instructions": v[0], "issue_floor_ms": 1e3 * v[0] * 64 / VALU_LANE_OPS,
                                 "issue_frac": (v[0] * 64 / VALU_LANE_OPS) / (ms * 1e-3),
                                 "fp64_tflops": v[1] / (ms * 1e-3) / 1e12,
                                 "fp64_frac": v[1] / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    dom = max(ks, key=lambda k: ks[k]["ms"])
    d = ks[dom]
    return {"bound": "hbm", "kernel": d["kernel"], "achieved": d["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": d["frac"], "traffic": traffic_from_profiles(dom, wkey), "kernel_ms_avg": d["ms"],
            "bytes_per_cell": d["bytes_per_cell"], "dominant_by": "per-launch device time in this run",
            "valu": d.get("valu"),
            "kernels": ks, "class_ms": r["class_ms"]}


def c3_main(a):
    from fv3d import cases
    plain = cases.vane_passage(levels=0, max_iter=a.drop_max_iter)
    chosen = cases.vane_passage(levels=a.levels, max_iter=a.drop_max_iter)
    wname = (f"vane passage 88x88x(24,48,144,72,40) ({plain.n_cells:,} cells; blocks 0.19M-1.12M), hub 0.2 m "
             "tip 0.25 m, 22 blades, turning vane theta = 0.1 pitch t^2, Navier-Stokes + Spalart-Allmaras, "
             "Roe + MUSCL3 primitive, LU-SGS local dt CFL 5, outlet 0.95 Pt, stationary")
    passage_main(a, plain, chosen, wname, "c3")


def c4_main(a):
    from fv3d import cases
    plain = cases.rotor_stator_stage(levels=0, max_iter=a.drop_max_iter)
    chosen = cases.rotor_stator_stage(levels=a.levels, max_iter=a.drop_max_iter)
    wname = (f"rotor + stator stage, 14 blocks ({plain.n_cells:,} cells; blocks 0.29M-1.97M): rotor 22 blades "
             "128x128x(24,40,64,120,64,40,24) at 3000 rpm, blades on the relative-flow helix turning 0.04 pitch, "
             "stator 30 blades 128x96x(24,48,72,144,96,64,32) turning -0.03 pitch, mixing plane between the rows, "
             "hub 0.2 m tip 0.25 m, "
             "Navier-Stokes + Spalart-Allmaras, Roe + MUSCL3 primitive, LU-SGS local dt CFL 5, outlet 0.95 Pt")
    passage_main(a, plain, chosen, wname, "c4")


# Residual configurations of the vane passage: what the autotuner picks on it
# (convective 8x4x8, viscous 4x4x16, nu~ 4x8x8), fixed so that every run times
# the kernels the counter profile in profiles/lusgs_vcycle_vane_c3.md measured
FIXED_TILES = {"c3": "3,0,2"}


def passage_main(a, plain, chosen, wname, wkey):
    if a.tiles is None and not a.autotune and wkey in FIXED_TILES:
        a.tiles = FIXED_TILES[wkey]
    closure = None if a.no_closure else load_closure(a.closure_relax)
    cells = plain.n_cells
    if a.plain_only:
        p = time_form(plain, "fp64", a)
        print(f"bench: plain FP64 timed, {p['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
        dp = None if a.no_drop else time_to_drop(plain, "fp64", 1e-6, a.drop_max_iter, tiles=a.tiles)
        if dp:
            print(f"bench: plain FP64 drop {dp['iterations']} iterations, {dp['seconds']} s", file=sys.stderr, flush=True)
        print(json.dumps({"metric": METRIC, "value": p["value"], "unit": "cell-updates/s", "n_gpus": 1,
                          "steps": a.steps, "warmup": a.warmup, "ms_per_step": p["ms_per_step"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                          "data": "synthetic",
                          "config": {"workload": wname, "cells": cells, "form": "plain FP64 only, the reference's "
                                     "two-grid cycle, Spalart-Allmaras transport"},
                          "time_to_1e-6_drop": {"plain_fp64": dp},
                          "roofline_plain_fp64": roofline(p, "fp64", True, wkey + "_fp64")}))
        return
    r = time_form(chosen, "fp32", a, closure)
    print(f"bench: chosen form timed, {r['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    p = None
    if not a.no_plain:
        p = time_form(plain, "fp64", a)
        print(f"bench: plain FP64 timed, {p['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    drop = None
    if not a.no_drop:
        dc = time_to_drop(chosen, "fp32", 1e-6, a.drop_max_iter, closure, tiles=a.tiles)
        print(f"bench: chosen form drop {dc['iterations']} iterations, {dc['seconds']} s", file=sys.stderr, flush=True)
        ds = None
        if closure is not None and not a.no_chosen_sa:
            ds = time_to_drop(chosen, "fp32", 1e-6, a.drop_max_iter, tiles=a.tiles)
            print(f"bench: chosen form with Spalart-Allmaras drop {ds['iterations']} iterations, {ds['seconds']} s",
                  file=sys.stderr, flush=True)
        dp = None
        if not a.no_plain:
            dp = time_to_drop(plain, "fp64", 1e-6, a.drop_max_iter, tiles=a.tiles)
            print(f"bench: plain FP64 drop {dp['iterations']} iterations, {dp['seconds']} s", file=sys.stderr, flush=True)
        drop = {"chosen": dc, "chosen_spalart_allmaras": ds, "plain_fp64": dp,
                "workload": "the headline workload, from the uniform start"}
        if dp and dc["seconds"] and dp["seconds"]:
            drop["speedup_chosen_vs_plain"] = dp["seconds"] / dc["seconds"]
            # the vane is stationary: its 'efficiency' divides by a near-zero
            # work term and is not a meaningful comparison, so it is not checked
            for q in ("mass_in", "pressure_ratio"):
                drop[q + "_rel_err"] = abs(dc[q] - dp[q]) / abs(dp[q])
        if dp and ds and ds["seconds"] and dp["seconds"]:
            drop["speedup_chosen_spalart_allmaras_vs_plain"] = dp["seconds"] / ds["seconds"]
            for q in ("mass_in", "pressure_ratio"):
                drop["chosen_spalart_allmaras_" + q + "_rel_err"] = abs(ds[q] - dp[q]) / abs(dp[q])
    conv = None if (a.no_converged or a.no_drop) else converged_field_error()
    cpu = None if a.no_cpu_baseline else cpu_baseline(a.cpu_iters, wkey)
    cpu_all = None if a.no_cpu_baseline else cpu_all_cores(a.cpu_iters, wkey)
    tile_desc = ("autotuned tiles" if a.tiles is None else
                 f"tiles {a.tiles} (what the autotuner picks on this workload)" if wkey in FIXED_TILES and
                 a.tiles == FIXED_TILES[wkey] else f"tiles {a.tiles}")
    chosen_desc = (f"FP32 defect correction + {a.levels}-level V-cycle + one launch per stage over all blocks in a "
                   f"HIP graph, {tile_desc}, face-parallel convective residual on levels of at most 1M cells" +
                   (" + learned closure (Spalart-Allmaras fallback)" if closure else
                    " + Spalart-Allmaras transport (no trained closure present)"))
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/update/norms, f32 LU-SGS operator and coarse corrections",
        "data": "synthetic",
        "config": {"workload": wname, "cells": cells, "chosen_form": chosen_desc,
                   "plain_form": "plain FP64, the reference's two-grid cycle, Spalart-Allmaras transport",
                   "hip_graph": not a.no_graph, "tiles": r["tiles"],
                   "closure": closure["meta"] if closure else None},
        "plain_fp64": {"value": p["value"], "ms_per_step": p["ms_per_step"], "class_ms": p["class_ms"]} if p else None,
        "time_to_1e-6_drop": drop,
        "converged_field_error": conv,
        "roofline": roofline(r, "fp32", closure is None, wkey if closure is None else wkey + "_closure"),
        "roofline_plain_fp64": roofline(p, "fp64", True, wkey + "_fp64") if p else None,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(out))


def c2_main(a):
    from fv3d import cases
    cs = cases.passage_batch(n=a.instances, seed=a.batch_seed, design_rpm=a.batch_rpm, max_iter=a.steps)
    ncell = sum(c.n_cells for c in cs)
    wname = (f"{a.instances} independent five-block blade passages in one batched context ({ncell:,} cells; "
             f"sizes log-uniform in [0.1M, 2M] cells, seed {a.batch_seed}; back-pressure factor U[0.95, 1.25] "
             f"on 0.76 Pt; rotation U[0.7, 1.05] x {a.batch_rpm:g} rpm), Navier-Stokes + Spalart-Allmaras, "
             "Roe + MUSCL3 primitive, LU-SGS local dt CFL 5, single grid level")
    r = time_form(cs, "fp32", a)
    out = {"metric": "cell-updates/s of the batched passages, FP32 LU-SGS over FP64 residual", "value": r["value"],
           "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 residual/update, f32 LU-SGS operator", "data": "synthetic",
           "config": {"workload": wname, "cells": ncell, "tiles": r["tiles"]},
           "roofline": roofline(r, "fp32", True, f"c2_{a.instances}s{a.batch_seed}")}
    if a.compare_alone:
        total, cells = 0.0, 0
        for q, c in enumerate(cs):
            g = solver(c, "fp32")
            g.iterate(a.warmup, graph=True)
            g.sync()
            t0 = time.perf_counter()
            g.iterate(a.steps, graph=True)
            g.sync()
            total += time.perf_counter() - t0
            cells += g.n_cells
            g.close()
            print(f"alone: passage {q + 1}/{len(cs)} timed", file=sys.stderr, flush=True)
        alone = cells * a.steps / total
        out["batch_vs_alone"] = {"alone_value": alone, "speedup": r["value"] / alone}
    print(json.dumps(out))


def legacy_main(a):
    """The 1.08M-cell Baldwin-Lomax passage (two-grid) or the bump channel,
    FP32 defect correction against plain FP64."""
    from fv3d import cases
    if a.workload == "bump":
        case = cases.bump_channel(ni=a.ni, nj=a.nj, nk=a.nk, seed=1, max_iter=a.steps, viscous=a.viscous,
                                  len_ref=1e-4 if a.viscous else 1.0)
        wname = f"{'laminar' if a.viscous else 'inviscid'} bump channel {a.nk}x{a.nj}x{a.ni} cells"
        wkey = f"bump{a.nk}x{a.nj}x{a.ni}" + ("ns" if a.viscous else "")
    else:
        case = cases.annular_passage(ni=64, nj=64, nk_blocks=(12, 40, 128, 60, 24), seed=1, rpm=0.0,
                                     blade_blocks=(2,), camber=0.1, viscous=True, len_ref=1.0, max_iter=a.steps,
                                     z_extent=(-0.06, -0.035, -0.015, 0.035, 0.055, 0.08))
        case.cfg.turbulence = 100
        case.cfg.use_multigrid, case.cfg.reuse_solution, case.cfg.initial, case.cfg.coarse_iter = "Y", "N", "U", 2
        wname = ("five-block annular blade passage 64x64x(12,40,128,60,24) (1,081,344 cells), circular-arc blade, "
                 "Navier-Stokes + Baldwin-Lomax, two-grid cycle, stationary")
        wkey = "passage5mg"
    res = {pr: time_form(case, pr, a) for pr in ("fp32", "fp64")}
    r = res["fp32"]
    out = {"metric": METRIC, "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 residual/update, f32 LU-SGS operator", "data": "synthetic",
           "config": {"workload": wname, "cells": r["cells"], "tiles": r["tiles"]},
           "plain_fp64": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms_per_step"]},
           "roofline": roofline(r, "fp32", False, wkey)}
    print(json.dumps(out))


# the closure's training family: the vane passage at three resolutions (cells
# across, blocks along the axis; every size a multiple of 8)
CLOSURE_RES = {40: (16, 24, 64, 32, 16), 64: (16, 32, 104, 48, 32), 88: (24, 48, 144, 72, 40)}


def closure_train(a):
    """The closure's training set and held-out evaluation on the vane family:
    ``train_instances - 8`` training passages, each at one of three
    resolutions (40, 64 or 88 cells across, 0.24M / 0.96M / 2.54M cells,
    drawn in turn) and a back-pressure ratio drawn uniform in [0.93, 0.97] of
    Pt (which sets the through-flow speed), developed with Spalart-Allmaras in
    one batched context on the three-level V-cycle; the network is trained on
    the GPU on their mue_t fields; 8 held-out passages at the headline
    resolution (88 cells across, 2,540,032 cells) and back pressures of their
    own then run from the uniform start to a 1e-6 drop once with the closure
    and once with Spalart-Allmaras, in the chosen form, and mass flow and
    pressure ratio are compared, with the gate's fallback fraction.  Writes
    the network to fv3d/closure_vane.npz (or --closure-out)."""
    import numpy as np
    from fv3d import cases, closure
    rng = np.random.default_rng(a.batch_seed)
    n = a.train_instances
    prs = rng.uniform(0.93, 0.97, n)
    res_cycle = (40, 64, 88)

    def fam(q, ni, levels=3):
        return cases.vane_passage(ni=ni, nk_blocks=CLOSURE_RES[ni], seed=100 + q, levels=levels,
                                  p_ratio=float(prs[q]), max_iter=100000)

    train, held = list(range(n - 8)), list(range(n - 8, n))
    tcases = [fam(q, res_cycle[q % 3]) for q in train]
    t0 = time.perf_counter()
    g = solver(tcases, "fp32")
    done = 0
    while done < a.train_develop:   # in chunks, with a progress line each
        k = min(250, a.train_develop - done)
        g.iterate(k, graph=True)
        done += k
        print(f"closure: developing, {done} iterations ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    g.sync()
    dev_s = time.perf_counter() - t0
    ncell = sum(c.n_cells for c in tcases)
    print(f"closure: developed {len(train)} passages ({ncell:,} cells) in {dev_s:.1f} s", file=sys.stderr, flush=True)
    xs, ts = [], []
    for q in range(len(train)):
        for b in range(1, 6):
            xs.append(g.mlp_features(b, inst=q + 1).reshape(-1, closure.N_IN))
            ts.append(g.mlp_targets(b, inst=q + 1).ravel())
    x, t = np.concatenate(xs).astype(np.float32), np.concatenate(ts).astype(np.float32)
    del xs, ts
    perm = np.random.default_rng(a.batch_seed + 1).permutation(len(x))
    x, t = x[perm], t[perm]
    bounds = closure.gate_bounds(x, t)
    g.mlp_set(closure.init_params(a.batch_seed), bounds)
    t1 = time.perf_counter()
    steps = 4000
    hist = g.mlp_train(x, t, steps, 4096, 2e-3)
    train_s = time.perf_counter() - t1
    P = g.mlp_get()
    g.close()
    print(f"closure: trained on {len(x):,} cells in {train_s:.1f} s, loss {hist[0]:.4f} -> {hist[-1]:.5f}",
          file=sys.stderr, flush=True)
    meta = {"family": "vane passage at 40, 64 and 88 cells across (0.24M, 0.96M, 2.54M cells) in turn, back "
                      f"pressure U[0.93, 0.97] Pt, seed {a.batch_seed}", "train_passages": len(train),
            "train_cells": ncell, "samples": int(len(x)), "steps": steps, "batch": 4096, "lr": 2e-3,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]), "develop_iterations": a.train_develop}
    out_path = a.closure_out or CLOSURE_WEIGHTS
    np.savez(out_path, params=P, bounds=bounds, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8))
    if a.no_heldout:
        print(json.dumps({"training": meta}))
        return
    model = {"params": P, "bounds": bounds, "meta": meta, "relax": a.closure_relax}
    res = {}
    for mode in ("sa", "closure"):
        rows = []
        for q in held:
            d = time_to_drop(fam(q, 88, levels=a.levels), "fp32", 1e-6, a.heldout_max_iter,
                             model if mode == "closure" else None)
            rows.append(d)
            print(f"closure: held-out {q} {mode}: {d['iterations']} iterations", file=sys.stderr, flush=True)
        res[mode] = rows
    rel = lambda k: [abs(c[k] - s[k]) / abs(s[k]) if c[k] is not None and s[k] is not None else None
                     for c, s in zip(res["closure"], res["sa"])]
    out = {"metric": "learned closure vs Spalart-Allmaras on held-out passages (converged to a 1e-6 drop)",
           "training": meta, "develop_seconds": dev_s, "train_seconds": train_s,
           "held_out": {"cells": fam(held[0], 88).n_cells, "pressure_ratios": [float(prs[q]) for q in held],
                        "mass_in_rel_err": rel("mass_in"), "pr_rel_err": rel("pressure_ratio"),
                        "iterations_sa": [d["iterations"] for d in res["sa"]],
                        "iterations_closure": [d["iterations"] for d in res["closure"]],
                        "seconds_sa": [d["seconds"] for d in res["sa"]],
                        "seconds_closure": [d["seconds"] for d in res["closure"]],
                        "gate_fallback_fraction": [d["gate_fallback_fraction"] for d in res["closure"]]}}
    print(json.dumps(out))


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus must be 1 (the multi-GPU sharding is not implemented)",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "c3":
        c3_main(a)
    elif a.workload == "c4":
        c4_main(a)
    elif a.workload == "c2":
        c2_main(a)
    elif a.workload == "closure-train":
        closure_train(a)
    else:
        legacy_main(a)


if __name__ == "__main__":
    main()
