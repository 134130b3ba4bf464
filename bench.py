#!/usr/bin/env python3
"""Throughput of the per-cell explicit update on one MI355X.

Default workload: the 2D two-material stiffened-gas shock/inclusion
interaction on a single level of 2048 x 1024 cells in 64 x 64 boxes (MUSCL +
THINC/BVD, beta 1.6, HLLC, CFL 0.5), synthetic initial data generated here.
`--workload plate3d` runs the 3D elastoplastic copper/aluminium plate impact
on 256^3 cells in 64^3 boxes (Us-up Mie-Gruneisen, 5-wave flux, perfect
plasticity); `--workload explosion_rz` an axisymmetric (r, z) underwater
explosion on 1448^2 cells with the cylindrical geometric source;
`--workload reactive2d` a reactive mixture (MixtureEOS with rootFind in every
conservativeToPrimitive, pressure-based burn) on 2048 x 1024 cells;
`--workload plate2d_grad` the gradient of a probe-trace loss of the 2D plate
impact (512^2, 200 steps) with respect to six material parameters by the
tangent-linear form, against the forward run.  One "step" is one full time step of the reference's loop body
(main.cpp:124-171): CFL dt with the driver clamps, two RK stages, the RK
average, plasticUpdate and cleanUpV (solids), cleanUpAlpha, the NaN scan and
the next step's CFL reduction.  The state is resident in HBM before timing
starts.

Prints one JSON line:  metric/value (Mcell-updates/s = cells x steps / s),
roofline of the dominant kernel (HIP-event times on the library's stream over
the timed region), step roofline (8 (3 N_s + 2 N_c) bytes per cell-update) and
the CPU baseline (the oracle restatement on this host's cores, a bounded sample
of the same workload).  The full-size default run also reports, under
`companions`, the 3D plate impact's mixed-over-FP64 speed-up with its L1 error
and the 2D plate impact's gradient cost over the forward run, each from a
child run of this script (`--companions on|off` forces them either way),
and under `machine` the HBM and FMA rates that plain kernels reach on this
device (tenjin_amd/csrc/tj_peaks.hip; `--peaks` asks for them on any run).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell of each kernel (unique state reads + writes, FP64,
# two materials; halo re-reads and THINC history reads not counted).
FLUID_KERNEL_BYTES = {
    # stage kernels: x kernel = 8 primitives + 6 further conserved comps in, the
    # partially updated conserved state (8 comps) and 2 star alphas out; y kernel
    # = 8 primitives and the partial state (8) in, 2 star alphas and 15 comps
    # out, in stage 2 also 15 base comps in and the sound speed out
    "xstage1": 112 + 64 + 16,
    "xstage2": 112 + 64 + 16,
    "ystage1": 64 + 64 + 16 + 120,
    "ystage2_post": 64 + 64 + 16 + 120 + 120 + 8,
    "xstage1_mixed": 112 + 64 + 16,
    "xstage2_mixed": 112 + 64 + 16,
    "ystage1_mixed": 64 + 64 + 16 + 120,
    "ystage2_post_mixed": 64 + 64 + 16 + 120 + 120 + 8,
    "flux_x": 64 + 72 + 16,            # 8 primitives in, 9 flux comps out, 2 star alphas out
    "flux_y": 64 + 72 + 16,
    "flux_z": 64 + 72,
    "update_stage1": 64 + 2 * 72 + 120,          # 8 conserved in, 2 faces x 9 comps, 15 comps out
    "update_stage2_post": 64 + 2 * 72 + 120 + 120 + 8,  # + 15 base comps in, + sound speed out
    # y pass with the update folded in: 8 primitives + 6 further conserved
    # comps in, x faces in (9 comps), star alphas out, 15 comps out
    "flux_y_update1": 112 + 72 + 16 + 120,
    "flux_y_update2_post": 112 + 72 + 16 + 120 + 120 + 8,  # + base in, + sound speed out
    # mixed form: FP32 flux planes (9 comps x 4 B), FP64 state
    "flux_x_mixed": 64 + 36 + 16,
    "flux_y_mixed": 64 + 36 + 16,
    "flux_z_mixed": 64 + 36,
    "update_stage1_mixed": 64 + 2 * 36 + 120,
    "update_stage2_post_mixed": 64 + 2 * 36 + 120 + 120 + 8,
}
# two-material fluid with one mixture material: 9 primitives (lambda), 10
# flux comps (alpha rho lambda), 19 persistent comps written by the update
MIX_KERNEL_BYTES = {
    "flux_x": 72 + 80 + 16,
    "flux_y": 72 + 80 + 16,
    "update_stage1": 72 + 2 * 80 + 152,
    "update_stage2_post": 72 + 2 * 80 + 152 + 152 + 8,
    "flux_x_mixed": 72 + 40 + 16,
    "flux_y_mixed": 72 + 40 + 16,
    "update_stage1_mixed": 72 + 2 * 40 + 152,
    "update_stage2_post_mixed": 72 + 2 * 40 + 152 + 152 + 8,
}
SOLID_KERNEL_BYTES = {
    # 19 primitives in (alpha, rho_k, epsilon, u, p, V), 23 face comps out
    # (19 conserved fluxes, S*, the V* row of the face)
    "flux_x": 152 + 184,
    "flux_y": 152 + 184,
    "flux_z": 152 + 184,
    # 19 conserved in, 3 directions x 23 face comps, 29 state comps out
    "update_stage1": 152 + 3 * 184 + 232,
    # + 29 base comps in (RK average), + sound speed out
    "update_stage2_post": 152 + 3 * 184 + 232 + 232 + 8,
    # mixed form: FP32 flux planes (23 comps x 4 B), FP64 state
    "flux_x_mixed": 152 + 92,
    "flux_y_mixed": 152 + 92,
    "flux_z_mixed": 152 + 92,
    "update_stage1_mixed": 152 + 3 * 92 + 232,
    "update_stage2_post_mixed": 152 + 3 * 92 + 232 + 232 + 8,
}
# 8 (3 N_s + 2 N_c): N_s/N_c = 14/8 (two-material fluid), 27/19 (two-material solid)
# reactive2d: N_s = 14 + 4 (alpha rho lambda, lambda, two phase densities), N_c = 9
# stage forms of tj_set_tuning's fuse argument, and the workloads whose levels
# the stage kernels serve (2D fluid, no mixture)
FORMS = {"stage": 2, "planes": 0, "march": 1}
STAGE2D_WORKLOADS = ("inclusion2d", "explosion_rz")

STEP_BYTES = {"inclusion2d": 464, "plate3d": 952, "tube1d": 464, "ensemble": 464, "explosion_rz": 464,
              "reactive2d": 8 * (3 * 18 + 2 * 9)}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["inclusion2d", "plate3d", "tube1d", "ensemble", "explosion_rz",
                                           "reactive2d", "plate2d_grad", "inclusion2d_grad"], default="inclusion2d")
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 500 for inclusion2d, 40 for plate3d)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--prec", choices=["fp64", "mixed", "both"], default="both",
                    help="both: FP64 is the headline value, the mixed form is timed beside it with its L1 error")
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--n3", type=int, default=256, help="plate3d: cells per side")
    ap.add_argument("--nrz", type=int, default=1448, help="explosion_rz: cells per side")
    ap.add_argument("--n2g", type=int, default=512, help="plate2d_grad: cells per side")
    ap.add_argument("--box", type=int, default=64)
    ap.add_argument("--seglen", type=int, default=0,
                    help="rows per marching segment: > 0 fixed; -1 the library's own pick per pass; 0 (default) the "
                         "fastest of the library's pick, 8 and 12, timed on a scratch copy of the level first, "
                         "around the middle of the run")
    ap.add_argument("--form", choices=["stage", "planes", "march"], default="stage",
                    help="RK stage form of a 2D fluid level: stage = the x and y stage kernels with no flux planes "
                         "(default; levels with a mixture use planes), planes = flux planes and an update kernel, "
                         "march = flux planes with the update inside the y pass")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--companions", choices=["auto", "on", "off"], default="auto",
                    help="also run the 3D plate impact (FP64 and mixed) and the 2D plate impact's gradient as child "
                         "runs and report them in the JSON line; auto = on the full-size default workload with "
                         "the CPU baseline")
    ap.add_argument("--peaks", action="store_true",
                    help="measure the device's HBM and FMA rates with plain kernels and report them (always on the "
                         "full-size default run)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a plain 1 GiB device copy first (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = {"inclusion2d": 500, "plate3d": 40, "tube1d": 1280, "ensemble": 200,
                      "explosion_rz": 500, "reactive2d": 200, "plate2d_grad": 200,
                      "inclusion2d_grad": 100}[args.workload]
    return args


def require_clean_run(status: int, steps_done: int, steps_expected: int, what: str) -> None:
    """A timed run counts only if every loop body ran a real step: a status bit
    (NaN stop, fatal states) or a loop that ended early makes the later
    launches no-ops and the rate meaningless, so the bench stops instead."""
    if status != 0 or steps_done != steps_expected:
        raise SystemExit(f"bench: {what}: the timed run is invalid (status {status}, {steps_done} device steps "
                         f"of {steps_expected})")


def cpu_model() -> str:
    """The host CPU's model name (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_leg(pb, threads: int, budget_s: float, max_steps: int):
    """The oracle on `threads` OpenMP threads: one warm-up step, then steps
    until the time budget or max_steps is reached; (steps, seconds)."""
    from oracle.oracle import Oracle, set_threads

    set_threads(threads)
    o = Oracle(pb)
    o.step(0.2 * o.timestep())  # warm-up step (not timed)
    n = 0
    t0 = time.perf_counter()
    while True:
        dt = o.timestep()
        o.step(dt if n + 1 >= 5 else 0.2 * dt)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= max_steps:
            break
    del o
    return n, el


def cpu_baseline(pb, budget_s: float, max_steps: int = 20):
    """The oracle (CPU restatement) on this host's cores: first steps of the
    same workload from its initial condition, until the time budget is used,
    on the host's CPU share (OMP_NUM_THREADS, else all cores up to 16) and on
    one core.  The budget is split between the two legs."""
    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, nproc)
    n, el = _cpu_leg(pb, threads, 0.6 * budget_s, max_steps)
    n1, el1 = _cpu_leg(pb, 1, 0.4 * budget_s, max_steps)
    size = "x".join(str(v) for v in pb.n[:pb.dim])
    return {
        "value": pb.ncells * n / el / 1e6,
        "unit": "Mcell-updates/s",
        "cores": threads,
        "kind": "port",
        "one_core": {"value": pb.ncells * n1 / el1 / 1e6, "cores": 1, "steps": n1, "seconds": round(el1, 2)},
        "cpu_model": cpu_model(),
        "nproc": nproc,
        "sample": f"oracle restatement, {size} level, {n} steps after 1 warm-up step, {el:.1f} s, {threads} OpenMP "
                  f"threads over boxes (the host's CPU share; nproc {nproc}); one core: {n1} steps, {el1:.1f} s",
    }


TRAFFIC_FILES = {"inclusion2d": "pmc_traffic_shock_inclusion_2d.txt", "plate3d": "pmc_traffic_plate_impact_3d.txt",
                 "tube1d": "pmc_traffic_shock_tube_1d.txt", "explosion_rz": "pmc_traffic_explosion_rz.txt",
                 "reactive2d": "pmc_traffic_reactive_2d.txt"}


def read_traffic(workload: str, name: str):
    """HBM bytes per launch of `name` from the committed rocprofv3 PMC summary."""
    path = os.path.join(HERE, "profiles", TRAFFIC_FILES[workload])
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == "traffic_bytes_per_launch" and len(parts) >= 3 and parts[1] == name:
            return float(parts[2])
    return None


def read_valu_gflop(workload: str, name: str):
    """FP64 GFLOP per launch of `name` from the committed rocprofv3 VALU
    counter summary (profiles/valu_fp64_kernels.txt)."""
    path = os.path.join(HERE, "profiles", "valu_fp64_kernels.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 4 and parts[0] == "valu_gflop_per_launch" and parts[1] == workload and parts[2] == name:
            return float(parts[3])
    return None


FP64_VALU_PEAK_TFLOPS = 78.6


def bench_tube1d(args):
    """The 1D shock tube (1000 cells): latency-bound, so the loop bodies run as
    replays of one captured hipGraph of 64 bodies; the eager device loop is
    timed beside it."""
    import torch
    from tenjin_amd import GpuLevel, TJ_PREC_FP64, configs

    pb = configs.shock_tube_1d()
    group = 64
    nrep = max(1, args.steps // group)
    steps = nrep * group

    def run(graph):
        g = GpuLevel(pb)
        g.init_from_prims(pb.prims, pb.shift_pressure)
        g.begin_run(float("inf"))
        if graph:
            g.graph_capture(group, TJ_PREC_FP64)
            g.graph_launch()  # warm-up
        else:
            g.advance(group, TJ_PREC_FP64)
        g.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nrep):
            if graph:
                g.graph_launch()
            else:
                g.advance(group, TJ_PREC_FP64)
        g.sync()
        el = time.perf_counter() - t0
        require_clean_run(g.status(), g.clock()[2], group + steps, f"tube1d graph={graph}")
        del g
        return el

    el_graph = run(True)
    el_eager = run(False)
    value = pb.ncells * steps / el_graph / 1e6
    out = {
        "metric": "Mcell-updates/s on 1 GPU",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": group,
        "ms_per_step": round(1e3 * el_graph / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "1D two-material stiffened-gas shock tube, 1000 cells, MUSCL+HLLC, CFL 0.5",
                   "cells": pb.ncells, "precision": "fp64", "launch": f"hipGraph of {group} loop bodies",
                   "loop_body": "k_clock_dt, k_xstage1d stage 1, k_xstage1d stage 2 (+ clock advance)",
                   "parallelism": "single GPU"},
        "eager_ms_per_step": round(1e3 * el_eager / steps, 5),
        "graph_speedup": round(el_eager / el_graph, 3),
        # 1000 cells: each launch is a few waves, so the step is bound by launch
        # and dependency latency; the whole-step bandwidth shows how far from HBM
        "roofline": {"bound": "hbm", "kernel": "whole step (graph replays; launch- and latency-bound at 1000 cells)",
                     "achieved": round(value * 1e6 * STEP_BYTES["tube1d"] / 1e9, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(value * 1e6 * STEP_BYTES["tube1d"] / 1e9 / HBM_PEAK_GBS, 6),
                     "traffic": None},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(pb, args.cpu_seconds, 1000),
    }
    print(json.dumps(out), flush=True)


def bench_ensemble(args):
    """256 independent two-material stiffened-gas tubes (128 1D with 500-8000
    cells, 128 2D with sides 64-512, seed 6), MUSCL+THINC+HLLC, mixed
    precision with FP64 update, each with its own dt: the whole run is one
    captured hipGraph launched once, the 2D problems on the stage kernels with
    the y-tile shape autotuned per 2D size bucket.  Work = sum of cells x
    steps."""
    import torch
    from tenjin_amd import BATCH_TILES, GpuBatch, TJ_PREC_MIXED, autotune, configs

    probs = configs.ensemble()
    prec = TJ_PREC_MIXED
    t0 = time.perf_counter()
    table = autotune.tune_tiles(probs, prec)
    t_tune = time.perf_counter() - t0
    b = GpuBatch(probs)
    autotune.apply_tiles(b, probs, table)
    b.init_from_prims()
    b.begin_run()
    # the whole run is one captured hipGraph, launched once: every loop body of
    # every problem (per-problem dt, clamps and NaN stop stay on the device)
    steps = args.steps
    group = steps
    nrep = 1
    t0 = time.perf_counter()
    b.graph_capture(steps, prec)
    t_capture = time.perf_counter() - t0
    b.graph_launch()  # warm-up run (not timed)
    b.sync()
    b.init_from_prims()
    b.begin_run()
    b.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.graph_launch()
    b.sync()
    el = time.perf_counter() - t0
    t, n, st = b.clocks()
    for p_, (sp, np_) in enumerate(zip(st, n)):
        require_clean_run(int(sp), int(np_), steps, f"ensemble problem {p_}")
    cells = sum(pb.ncells for pb in probs)
    value = cells * steps / el / 1e6
    out = {
        "metric": "Mcell-updates/s on 1 GPU",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": group,
        "ms_per_step": round(1e3 * el / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-update",
        "data": "synthetic",
        "config": {"workload": "ensemble of 256 two-material stiffened-gas tubes (128 1D, 500-8000 cells; 128 2D, "
                               "sides 64-512), MUSCL+THINC+HLLC, per-problem dt",
                   "problems": len(probs), "cells": cells, "precision": "mixed",
                   "launch": f"one hipGraph launch for the whole run ({steps} loop bodies of every problem; capture and "
                             f"instantiation {t_capture:.2f} s, outside the timed region)",
                   "form": "stage kernels (2D), flux-plane passes (1D)",
                   "tile_by_bucket": {str(autotune.BUCKETS[k]): list(BATCH_TILES[v]) for k, v in table.items()},
                   "autotune_s": round(t_tune, 2),
                   "parallelism": "single GPU"},
        "steps_done_min_max": [min(n), max(n)],
        "status_nonzero": sum(1 for x in st if x),
        # the batch runs as graph replays, whose kernels are not timed one by
        # one: the roofline is the whole step's (464 algorithmic bytes per
        # cell-update, FP64 state; the FP32 flux planes are not counted)
        "roofline": {"bound": "hbm", "kernel": "whole step (graph replays of all k_b_* launches)",
                     "achieved": round(value * 1e6 * STEP_BYTES["ensemble"] / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(value * 1e6 * STEP_BYTES["ensemble"] / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_ensemble(probs, args.cpu_seconds),
    }
    print(json.dumps(out), flush=True)


def cpu_baseline_ensemble(probs, budget_s: float):
    """The oracle on every 16th problem of the ensemble (8 1D and 8 2D
    problems), each from its initial condition for as many steps as its share
    of the budget allows (at most 20), on the host's CPU share; the rate is
    the sample's cell-updates over its time."""
    from oracle.oracle import Oracle, set_threads

    nproc = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, nproc)
    set_threads(threads)
    sample = probs[::16]
    work, el, steps = 0, 0.0, 0
    for pb in sample:
        o = Oracle(pb)
        o.step(0.2 * o.timestep())  # warm-up step (not timed)
        t0 = time.perf_counter()
        n = 0
        while n < 20 and time.perf_counter() - t0 < budget_s / len(sample):
            dt = o.timestep()
            o.step(dt if n + 1 >= 5 else 0.2 * dt)
            n += 1
        el += time.perf_counter() - t0
        work += pb.ncells * n
        steps += n
        del o
    return {"value": work / el / 1e6, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": nproc,
            "sample": f"oracle restatement (FP64), {len(sample)} of the {len(probs)} problems (every 16th), "
                      f"{steps} steps in all after one warm-up step each, {el:.1f} s, {threads} OpenMP threads"}


def bench_plate2d_grad(args):
    """The differentiated 2D plate impact (512^2 Cu/Al, 200 steps, probe u_x
    at the aluminium's far face): the loss sum_n (u_n - u*_n)^2 against a
    target trace from perturbed parameters, and its gradient with respect to
    Gamma0, s and Y of both materials by the tangent-linear form (6
    directions; the direction-group size is tuned first).  Reports the
    gradient run's throughput (cells x steps / s of the gradient), the
    forward run's, their ratio, the gradient against central differences of
    the FP64 form (h = 1e-6 |theta| and 1e-4 |theta|), and the mixed forward
    run's time and error against the FP64 one."""
    import numpy as np
    import torch
    from tenjin_amd import GpuLevel, TJ_PREC_FP64, TJ_PREC_MIXED, configs, tangent

    n, steps = args.n2g, args.steps
    theta = configs.plate_params()
    pb = configs.plate_impact_differentiated(n=n, steps=steps)
    probe = configs.plate_probe(pb)
    target = tangent.forward_trace(configs.plate_impact_differentiated(n=n, theta=configs.plate_target_theta(),
                                                                       steps=steps), steps, probe)

    P0 = torch.from_numpy(pb.prims).to("cuda")

    def forward_once():
        """initialisation from the device-resident primitives + the time loop
        (what each direction group of the tangent run repeats)"""
        g = GpuLevel(pb)
        g.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.init_from_device(P0, pb.shift_pressure)
        g.begin_run(float("inf"))
        g.advance(steps)
        g.sync()
        el = time.perf_counter() - t0
        st = g.status()
        g.close()
        return el, st

    forward_once()  # warm-up
    t_fwd, st_fwd = min(forward_once() for _ in range(3))
    require_clean_run(st_fwd, steps, steps, "plate2d_grad forward")

    # the mixed forward beside it: time, and relative L1 of rho, p, alpha and
    # of the probe trace against the FP64 run after the same steps
    def run_state(prec, trace=False):
        g = GpuLevel(pb)
        g.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.init_from_device(P0, pb.shift_pressure)
        g.begin_run(float("inf"))
        tr = []
        if trace:
            comp, i, j, k = probe
            for _ in range(steps):
                g.advance(1, prec)
                g.sync()
                tr.append(float(g.S[comp, k, j, i].item()))
        else:
            g.advance(steps, prec)
        g.sync()
        el = time.perf_counter() - t0
        S = g.S.clone()
        st = g.status()
        g.close()
        return el, S, np.array(tr), st

    run_state(TJ_PREC_MIXED)  # warm-up
    t_mix = min(run_state(TJ_PREC_MIXED)[0] for _ in range(3))
    _, S64, tr64, _ = run_state(TJ_PREC_FP64, trace=True)
    _, S32, tr32, st_mix = run_state(TJ_PREC_MIXED, trace=True)
    L = pb.layout
    mixed_err = {}
    for name, c in (("rho", L.comp(configs.RHO)), ("p", L.comp(configs.P)), ("alpha", L.comp(configs.ALPHA, 0))):
        mixed_err[name] = float(f"{float((S32[c] - S64[c]).abs().sum() / S64[c].abs().sum()):.3e}")
    mixed_err["probe_trace"] = float(f"{np.abs(tr32 - tr64).sum() / max(np.abs(tr64).sum(), 1e-300):.3e}")
    del S64, S32
    # direction-group size: 10-step trial runs of each
    trial = {}
    for chunk in (1, 2, 3):
        tr = tangent.TangentRun(pb, tangent.ALL_DIRECTIONS, chunk=chunk)
        tr.run(2, probe)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr.run(10, probe)
        trial[chunk] = time.perf_counter() - t0
        tr.close()
    chunk = min(trial, key=trial.get)
    tr = tangent.TangentRun(pb, tangent.ALL_DIRECTIONS, chunk=chunk)
    tr.run(2, probe)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    trace, done = tr.run(steps, probe)
    t_grad = time.perf_counter() - t0
    st_grad = tr.status()
    tr.close()
    require_clean_run(st_grad, done, steps, "plate2d_grad gradient")
    r = trace[:, 0] - target
    loss = float(r @ r)
    grad = 2.0 * (r @ trace[:, 1:])
    # central differences of the FP64 form along each parameter, at two steps:
    # the loss (~1e-5 here) is a sum of squares of probe values that the wave
    # has barely reached, so the quotient's rounding noise grows as h shrinks
    fd_err = {}
    for rel in (1e-6, 1e-4):
        fd = np.zeros(6)
        for p in range(6):
            h = rel * abs(theta[p])
            lpm = []
            for sgn in (1.0, -1.0):
                th = list(theta)
                th[p] += sgn * h
                u = tangent.forward_trace(configs.plate_impact_differentiated(n=n, theta=th, steps=steps), steps,
                                          probe)
                lpm.append(tangent.probe_loss(u, target))
            fd[p] = (lpm[0] - lpm[1]) / (2 * h)
        fd_err[f"h={rel:.0e}|theta|"] = [float(f"{abs(a - b) / max(abs(b), 1e-300):.3e}") for a, b in zip(grad, fd)]
    value = pb.ncells * steps / t_grad / 1e6
    out = {
        "metric": "Mcell-updates/s on 1 GPU",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 2,
        "ms_per_step": round(1e3 * t_grad / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"differentiated 2D elastoplastic Cu/Al plate impact, {n}^2 cells, {steps} steps, "
                               f"impact {pb.speed:.1f} m/s; loss = sum of squared probe-trace misfits; gradient "
                               "w.r.t. Gamma0, s, Y of both materials by the tangent-linear form",
                   "cells": pb.ncells, "precision": "fp64", "directions": 6, "chunk": chunk,
                   "chunk_trial_s": {str(k): round(v, 4) for k, v in trial.items()},
                   "parallelism": "single GPU"},
        "adjoint_over_forward": round(t_grad / t_fwd, 3),
        "per_direction_over_forward": round(t_grad / t_fwd / 6, 3),
        "forward_value": round(pb.ncells * steps / t_fwd / 1e6, 2),
        "forward_ms": round(1e3 * t_fwd, 3),
        "mixed_forward": {"value": round(pb.ncells * steps / t_mix / 1e6, 2), "ms": round(1e3 * t_mix, 3),
                          "speedup_vs_fp64": round(t_fwd / t_mix, 3), "l1_rel_err_vs_fp64": mixed_err,
                          "status": st_mix},
        "gradient_ms": round(1e3 * t_grad, 3),
        "loss": loss,
        "gradient": [float(f"{v:.9e}") for v in grad],
        "gradient_rel_err_vs_central_differences": fd_err,
        "status": [st_fwd, st_grad],
        "steps_done": done,
        "roofline": None,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        # the gradient by central differences on the CPU costs 13 runs of the
        # oracle: its rate is the oracle's forward rate / 13
        cb = cpu_baseline(configs.plate_impact_differentiated(n=min(n, 256), steps=steps), args.cpu_seconds, 40)
        cb["value"] = cb["value"] / 13.0
        cb["sample"] = "gradient by central differences (13 forward runs): " + cb["sample"]
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


def bench_inclusion2d_grad(args):
    """The gradient of a probe-trace loss on the 2D shock/inclusion (THINC)
    with respect to the Gruneisen coefficients (adiabaticIndex - 1) of both
    gases, by the tangent-linear form (2 directions).  The target trace comes
    from both coefficients x 1.05; the probe is u_x at x = 0.53, mid-height,
    which the incoming shock crosses.
    Reports the gradient run's throughput, the forward run's, their ratio,
    and the gradient against central differences of the FP64 form."""
    import numpy as np
    import torch
    from tenjin_amd import GpuLevel, configs, tangent

    steps = args.steps

    def make(theta):
        pb = configs.shock_inclusion_2d(nx=args.nx, ny=args.ny, box=args.box)
        for m in range(2):
            pb.materials[m].gamma = theta[m] + 1.0
        return pb

    theta = [m.gamma - 1.0 for m in configs.shock_inclusion_2d(nx=64, ny=32, box=16).materials]
    pb = make(theta)
    # x = 0.53 at mid-height: the shock (from x = 0.5) crosses it within the run
    probe = (pb.layout.comp(configs.VELOCITY, 0, 0), int(0.53 * args.ny), args.ny // 2, 0)
    target = tangent.forward_trace(make([t * 1.05 for t in theta]), steps, probe)
    P0 = torch.from_numpy(pb.prims).to("cuda")

    def forward_once():
        g = GpuLevel(pb)
        g.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.init_from_device(P0, pb.shift_pressure)
        g.begin_run(float("inf"))
        g.advance(steps)
        g.sync()
        el = time.perf_counter() - t0
        g.close()
        return el

    forward_once()
    t_fwd = min(forward_once() for _ in range(3))
    dirs = [(0, "Gamma0"), (1, "Gamma0")]
    trial = {}
    for chunk in (1, 2):
        tr = tangent.TangentRun(pb, dirs, chunk=chunk)
        tr.run(2, probe)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr.run(10, probe)
        trial[chunk] = time.perf_counter() - t0
        tr.close()
    chunk = min(trial, key=trial.get)
    tr = tangent.TangentRun(pb, dirs, chunk=chunk)
    tr.run(2, probe)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    trace, done = tr.run(steps, probe)
    t_grad = time.perf_counter() - t0
    st = tr.status()
    tr.close()
    require_clean_run(st, done, steps, "inclusion2d_grad gradient")
    r = trace[:, 0] - target
    grad = 2.0 * (r @ trace[:, 1:])
    fd = []
    for p in range(2):
        h = 1e-6 * abs(theta[p])
        lpm = []
        for sgn in (1.0, -1.0):
            th = list(theta)
            th[p] += sgn * h
            lpm.append(tangent.probe_loss(tangent.forward_trace(make(th), steps, probe), target))
        fd.append((lpm[0] - lpm[1]) / (2 * h))
    out = {
        "metric": "Mcell-updates/s on 1 GPU",
        "value": round(pb.ncells * steps / t_grad / 1e6, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 2,
        "ms_per_step": round(1e3 * t_grad / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gradient of a probe-trace loss on the 2D shock/inclusion, {args.nx}x{args.ny} cells, "
                               f"{pb.max_grid_size}^2 boxes, MUSCL+THINC+HLLC, {steps} steps, w.r.t. the Gruneisen "
                               "coefficients of both gases by the tangent-linear form",
                   "cells": pb.ncells, "precision": "fp64", "directions": 2, "chunk": chunk,
                   "chunk_trial_s": {str(k): round(v, 4) for k, v in trial.items()}, "parallelism": "single GPU"},
        "adjoint_over_forward": round(t_grad / t_fwd, 3),
        "per_direction_over_forward": round(t_grad / t_fwd / 2, 3),
        "forward_value": round(pb.ncells * steps / t_fwd / 1e6, 2),
        "forward_ms": round(1e3 * t_fwd, 3),
        "gradient_ms": round(1e3 * t_grad, 3),
        "loss": float(r @ r),
        "gradient": [float(f"{v:.9e}") for v in grad],
        "gradient_rel_err_vs_central_differences": [float(f"{abs(a - b) / max(abs(b), 1e-300):.3e}")
                                                    for a, b in zip(grad, fd)],
        "status": st,
        "steps_done": done,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the domain decomposition over several GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    from tenjin_amd import GpuLevel, TJ_PREC_FP64, TJ_PREC_MIXED, autotune, configs

    if args.calib_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.workload == "tube1d":
        bench_tube1d(args)
        return
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "plate2d_grad":
        bench_plate2d_grad(args)
        return
    if args.workload == "inclusion2d_grad":
        bench_inclusion2d_grad(args)
        return
    if args.workload == "inclusion2d":
        pb = configs.shock_inclusion_2d(nx=args.nx, ny=args.ny, box=args.box)
        kbytes = FLUID_KERNEL_BYTES
        workload = (f"2D two-material shock/inclusion, {pb.n[0]}x{pb.n[1]} cells, {pb.max_grid_size}^2 boxes, "
                    "MUSCL+THINC(beta 1.6)+HLLC, CFL 0.5")
    elif args.workload == "reactive2d":
        pb = configs.reactive_tube(2, args.nx, box=args.box, thinc=1)
        kbytes = MIX_KERNEL_BYTES
        workload = (f"2D reactive mixture (MixtureEOS, pressure-based burn) beside air, {pb.n[0]}x{pb.n[1]} cells, "
                    f"{pb.max_grid_size}^2 boxes, MUSCL+THINC(beta 1.6)+HLLC, CFL 0.5")
    elif args.workload == "explosion_rz":
        pb = configs.underwater_explosion(2, args.nrz, box=args.box, thinc=1)
        kbytes = FLUID_KERNEL_BYTES
        workload = (f"2D axisymmetric (r, z) underwater explosion, {pb.n[0]}x{pb.n[1]} cells, "
                    f"{pb.max_grid_size}^2 boxes, MUSCL+THINC(beta 1.6)+HLLC, cylindrical geometric source, CFL 0.5")
    else:
        pb = configs.plate_impact(dim=3, n=args.n3, box=args.box)
        kbytes = SOLID_KERNEL_BYTES
        workload = (f"3D elastoplastic Cu/Al plate impact, {args.n3}^3 cells, {pb.max_grid_size}^3 boxes, "
                    "Us-up Mie-Gruneisen, MUSCL+5-wave HLLC, perfect plasticity, CFL 0.5")
    ncell = pb.ncells

    tuned = {}
    sampled = {}

    def timed(prec):
        """warm-up steps, then exactly args.steps timed steps of the device loop
        (segment length: --seglen, else the fastest of a few timed on a scratch
        copy of the level beforehand)"""
        fuse = FORMS[args.form]
        seglen = max(args.seglen, 0)
        if fuse == 2 and args.workload in STAGE2D_WORKLOADS:
            tuned[prec] = {"seglen": "unused by the stage kernels"}
        elif args.seglen == 0:
            # timed around the middle of the run (at most 250 loop bodies in)
            seglen, trial = autotune.tune_level_seglen(pb, prec, fuse,
                                                       at_step=min(args.warmup + args.steps // 2, 250))
            tuned[prec] = {"seglen": seglen, "trial_s": {str(k): round(v, 5) for k, v in trial.items()}}
        g = GpuLevel(pb)
        g.set_tuning(seglen, fuse)
        g.init_from_prims(pb.prims, pb.shift_pressure)
        g.begin_run(float("inf"))
        # warm-up: the first steps untimed (code loading), the rest with HIP
        # events on every launch for the per-kernel breakdown, which also
        # names the dominant kernel
        first = min(2, args.warmup)
        g.advance(first, prec)
        g.prof_enable(True)
        g.advance(args.warmup - first, prec)
        g.sync()
        warm = g.prof_report()
        # the timed steps: events around the dominant kernel's launches only
        # (each event record costs the stream a few microseconds)
        dname = max(warm.items(), key=lambda kv: kv[1][1])[0] if warm else None
        # events on about 50 of its launches: every launch in short runs, one
        # in ten over C2's 500 steps
        per_step = warm[dname][0] / max(1, args.warmup - first) if warm else 1.0
        every = max(1, int(args.steps * per_step) // 50)
        g.prof_enable(True)
        g.prof_only(dname)
        g.prof_sample(every)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.advance(args.steps, prec)
        g.sync()
        t1 = time.perf_counter()
        prof = g.prof_report()
        g.prof_sample(1)
        g.prof_only(None)
        g.prof_enable(False)
        sampled[prec] = every
        status = g.status()
        t_end, _, nsteps_done = g.clock()
        require_clean_run(status, nsteps_done, args.warmup + args.steps, f"{args.workload} prec {prec}")
        el = t1 - t0
        if dname is None:  # no profiled warm-up step: every timed launch carried events
            warm = prof
            dname = max(prof.items(), key=lambda kv: kv[1][1])[0]
        dcount, dms = prof[dname]
        davg_s = dms / dcount / 1e3
        dbytes = kbytes[dname] * ncell
        achieved = dbytes / davg_s / 1e9
        res = {
            "value": ncell * args.steps / el / 1e6,
            "ms_per_step": 1e3 * el / args.steps,
            "status": status,
            "device_steps": nsteps_done,
            "t_end": t_end,
            "kernel_avg_ms": {k: round(v[1] / v[0], 4) for k, v in warm.items()},
            "roofline": {
                "bound": "hbm",
                "kernel": dname,
                "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(args.workload, dname),
                "algorithmic_bytes_per_launch": dbytes,
                "avg_launch_ms": round(davg_s * 1e3, 4),
            },
        }
        gfl = read_valu_gflop(args.workload, dname) if prec == TJ_PREC_FP64 else None
        if gfl is not None:
            # the FP64 flux kernels are VALU/latency-bound: their compute roof
            tf = gfl / davg_s / 1e3
            res["roofline"]["valu"] = {"bound": "fp64-valu", "achieved": round(tf, 2), "peak": FP64_VALU_PEAK_TFLOPS,
                                       "unit": "TFLOP/s", "frac": round(tf / FP64_VALU_PEAK_TFLOPS, 4),
                                       "flops_per_launch": round(gfl * 1e9)}
        return g, res

    def rel_l1(a, b):
        """relative L1 error of density, pressure and the first volume fraction"""
        L = pb.layout
        out = {}
        for name, c in (("rho", L.comp(configs.RHO)), ("p", L.comp(configs.P)), ("alpha", L.comp(configs.ALPHA, 0))):
            x, y = a[c].double(), b[c].double()
            out[name] = float((x - y).abs().sum() / y.abs().sum().clamp_min(1e-300))
        return out

    mixed = None
    if args.prec in ("fp64", "both"):
        g, main_res = timed(TJ_PREC_FP64)
        prec = TJ_PREC_FP64
        if args.prec == "both":
            ref_state = g.S.clone()
            t_ref, _, n_ref = g.clock()
            del g
            torch.cuda.empty_cache()
            gm, mres = timed(TJ_PREC_MIXED)
            del gm
            torch.cuda.empty_cache()
            # accuracy: the mixed form from the same initial state to the FP64 run's final time
            ge = GpuLevel(pb)
            ge.init_from_prims(pb.prims, pb.shift_pressure)
            n_mix, t_mix = ge.run(t_ref, max_steps=n_ref + 1000, prec=TJ_PREC_MIXED, check_every=64)
            errs = rel_l1(ge.S, ref_state)
            del ge, ref_state
            torch.cuda.empty_cache()
            mixed = {
                "value": round(mres["value"], 2),
                "ms_per_step": round(mres["ms_per_step"], 4),
                "speedup_vs_fp64": round(mres["value"] / main_res["value"], 3),
                "dtype": "f32-flux/f64-update",
                "l1_rel_err_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs.items()},
                "compared_at": {"t": t_ref, "fp64_steps": n_ref, "mixed_steps": n_mix, "mixed_t": t_mix},
                "roofline": mres["roofline"],
                "kernel_avg_ms": mres["kernel_avg_ms"],
                "status": mres["status"],
            }
        else:
            del g
    else:
        g, main_res = timed(TJ_PREC_MIXED)
        prec = TJ_PREC_MIXED
        del g

    value = main_res["value"]
    ms_per_step = main_res["ms_per_step"]
    roofline = main_res["roofline"]
    kernel_ms = main_res["kernel_avg_ms"]
    status = main_res["status"]
    nsteps_done = main_res["device_steps"]
    step_bytes = STEP_BYTES[args.workload]
    step_gbs = value * 1e6 * step_bytes / 1e9

    out = {
        "metric": "Mcell-updates/s on 1 GPU",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == TJ_PREC_FP64 else "f32-flux/f64-update",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "cells": ncell,
            "precision": "fp64" if prec == TJ_PREC_FP64 else "mixed",
            "stage_form": args.form if args.workload in STAGE2D_WORKLOADS else "planes",
            "seglen": args.seglen if args.seglen else {("fp64" if k == TJ_PREC_FP64 else "mixed"): v
                                                        for k, v in tuned.items()},
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "step_roofline": {
            "bytes_per_cell_update": step_bytes,
            "achieved": round(step_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(step_gbs / HBM_PEAK_GBS, 4),
        },
        "kernel_avg_ms": kernel_ms,
        "kernel_avg_ms_over": ("warm-up steps after the first two (HIP events on every launch); the roofline "
                               "kernel's avg_launch_ms is over the timed steps, which carry events on that kernel alone"
                               f" (on one launch in {sampled.get(prec, 1)} of it, the first included)"
                               if args.warmup > 2 else "the timed steps (HIP events on every launch)"),
        "status": status,
        "device_steps": nsteps_done,
    }
    if mixed is not None:
        out["mixed"] = mixed
    if not args.no_cpu:
        if args.workload in ("inclusion2d", "explosion_rz", "reactive2d"):
            out["cpu_baseline"] = cpu_baseline(pb, args.cpu_seconds, 20)
        else:
            # bounded sample: the same problem on 128^3 cells (one eighth of the cells)
            out["cpu_baseline"] = cpu_baseline(configs.plate_impact(dim=3, n=min(args.n3, 128), box=args.box),
                                               args.cpu_seconds, 3)
    else:
        out["cpu_baseline"] = None
    full_default = (args.workload == "inclusion2d" and (args.nx, args.ny) == (2048, 1024) and not args.no_cpu and
                    not args.calib_copy)
    if args.peaks or full_default:
        peaks = machine_peaks()
        out["machine"] = peaks
        if isinstance(peaks.get("hbm_copy_gbs"), float):
            out["roofline"]["frac_of_measured_copy"] = round(out["roofline"]["achieved"] / peaks["hbm_copy_gbs"], 4)
            out["step_roofline"]["frac_of_measured_copy"] = round(step_gbs / peaks["hbm_copy_gbs"], 4)
        if isinstance(peaks.get("fp64_fma_tflops"), float) and "valu" in out["roofline"]:
            v = out["roofline"]["valu"]
            v["frac_of_measured_fma"] = round(v["achieved"] / peaks["fp64_fma_tflops"], 4)
    if args.companions == "on" or (args.companions == "auto" and full_default):
        out["companions"] = companion_runs(args)
    print(json.dumps(out), flush=True)


def machine_peaks(nbytes: int = 4 << 30):
    """What plain kernels reach on this device (tenjin_amd/libtenjin_peaks.so,
    kernels of csrc/tj_peaks.hip): HBM read, write and copy rates over a
    buffer of `nbytes` (copy counts the bytes read plus written), and the FMA
    throughput in FP64 and in FP32 without packed instructions.  The
    datasheet peaks stay the rooflines' `peak`; these give the same fractions
    against what the device delivers."""
    import ctypes

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tenjin_amd", "libtenjin_peaks.so")
    if not os.path.exists(path):
        return {"error": "tenjin_amd/libtenjin_peaks.so is not built"}
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.tjp_hbm.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, dp, dp]
    lib.tjp_fma.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
    lib.tjp_hbm.restype = lib.tjp_fma.restype = ctypes.c_int
    out = {"buffer_bytes": nbytes}
    for kind, name in ((0, "hbm_read_gbs"), (1, "hbm_write_gbs"), (2, "hbm_copy_gbs")):
        rate, ms = ctypes.c_double(), ctypes.c_double()
        rc = lib.tjp_hbm(kind, nbytes, 20, ctypes.byref(rate), ctypes.byref(ms))
        out[name] = round(rate.value, 1) if rc == 0 else f"error {rc}"
    for fp64, name in ((1, "fp64_fma_tflops"), (0, "fp32_fma_tflops")):
        rate, ms = ctypes.c_double(), ctypes.c_double()
        rc = lib.tjp_fma(fp64, 4096, 10, ctypes.byref(rate), ctypes.byref(ms))
        out[name] = round(rate.value, 2) if rc == 0 else f"error {rc}"
    out["kernels"] = ("read: one sum per block; write: fill; copy: read + write; 16-byte accesses, grid-stride, "
                      "2048 blocks of 256; FMA: 8 chains per lane, 8 waves per SIMD, FP32 unpacked")
    return out


def _child_line(cmd, timeout_s: float):
    """Runs bench.py as a child process (not an exec) and returns its JSON
    line, or an error record."""
    try:
        r = subprocess.run(cmd, cwd=os.path.dirname(os.path.abspath(__file__)), capture_output=True, text=True,
                           timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"timed out after {timeout_s:.0f} s"}
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit status {r.returncode}: {r.stderr.strip()[-300:]}"}
    return json.loads(lines[-1])


def companion_runs(args):
    """The two other headline figures beside the 2D level's rate, each from a
    bench.py child run with its own JSON line: the mixed form's speed-up
    over FP64 on the 3D elastoplastic plate impact (with the density, pressure
    and volume-fraction L1 error between the two forms), and the cost of the
    six-parameter gradient of the 2D plate impact's probe loss over one
    forward run (tangent-linear form)."""
    me = os.path.abspath(__file__)
    comp = {}
    t0 = time.perf_counter()
    d = _child_line([sys.executable, me, "--workload", "plate3d", "--n3", str(args.n3), "--box", str(args.box),
                     "--steps", "40", "--warmup", "2", "--no-cpu"], 600)
    if "error" in d:
        comp["plate3d"] = d
    else:
        m = d["mixed"]
        comp["plate3d"] = {"workload": d["config"]["workload"], "steps": d["steps"],
                           "fp64": d["value"], "mixed": m["value"], "mixed_over_fp64": m["speedup_vs_fp64"],
                           "l1_rel_err_vs_fp64": m["l1_rel_err_vs_fp64"], "status": [d["status"], m["status"]]}
    d = _child_line([sys.executable, me, "--workload", "plate2d_grad", "--n2g", str(args.n2g), "--no-cpu"], 600)
    if "error" in d:
        comp["plate2d_grad"] = d
    else:
        comp["plate2d_grad"] = {"workload": d["config"]["workload"], "steps": d["steps"],
                                "form": "tangent-linear (forward mode), one pass per group of directions",
                                "directions": d["config"]["directions"], "forward_ms": d["forward_ms"],
                                "gradient_ms": d["gradient_ms"], "gradient_over_forward": d["adjoint_over_forward"],
                                "per_direction_over_forward": d["per_direction_over_forward"],
                                "gradient_rel_err_vs_central_differences":
                                    d["gradient_rel_err_vs_central_differences"], "status": d["status"]}
    comp["wall_s"] = round(time.perf_counter() - t0, 1)
    return comp


if __name__ == "__main__":
    main()
