#!/usr/bin/env python3
"""Throughput of the fused D3Q19 time step on one MI355X, in the reference's metric.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload backstep|cylinder|cavity|box19|box27|bluff27|ensemble]
                    [--streaming two-lattice|aa]

Default workload: backward-facing step, D3Q19, Smagorinsky LES (Cs=0.1, wall damping),
1536 x 192 x 128 lattice (37.7M cells), FixedVelocity inlet with seeded 5 % Gaussian
turbulence, Geier outlet + sponge, Re_H = 5500 -- synthetic inputs built by
lbffs_amd.cases.backward_step.  One step = one fused stream-collide launch over the
whole lattice with the populations resident in HBM.

Printed: one JSON line.  value = lattice updates per second of the FP32-storage form
in MLUPS over all nx*ny*nz cells (the reference's own count, ns.cpp:607); the
FP16-deviation form is measured in the same run and reported beside it, with the
velocity error of both forms against an FP64 run of the same steps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import lbffs_amd as lb  # noqa: E402
from lbffs_amd import cases  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E bytes/s
PROFILE_PMC = os.path.join(ROOT, "profiles", "step19_backstep_pmc.txt")
PROFILE_ENS_PMC = os.path.join(ROOT, "profiles", "batch_ensemble_pmc.txt")
PROFILE_BLUFF_PMC = os.path.join(ROOT, "profiles", "step27_bluff_pmc.txt")
PMC_OF_WORKLOAD = {"backstep": PROFILE_PMC, "bluff27": PROFILE_BLUFF_PMC}
# bytes one ensemble cell update must move: 19 FP16 populations in and out, and the
# WALE model's FP32 velocity (its gradient reads the neighbours' velocities of the previous
# step) written once and read once
ENS_BYTES = 2 * 19 * 2 + 2 * 3 * 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="backstep", choices=["backstep", "cylinder", "cavity", "box19", "box27", "bluff27", "ensemble"])
    ap.add_argument("--domains", type=int, default=512, help="ensemble: number of lattices")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--only", default="", help="comma list of storage forms to run (f32,f16,bf16)")
    ap.add_argument("--forces", type=int, default=0,
                    help="cylinder, bluff27: also run this many steps per storage form (and FP64) and report "
                         "Cd, Cl, St")
    ap.add_argument("--streaming", default="two-lattice", choices=["two-lattice", "aa"],
                    help="population streaming: two arrays, or in place (AA; D3Q19, Cyclic/BounceBack faces)")
    ap.add_argument("--rounding", default="stochastic", choices=["stochastic", "nearest"],
                    help="rounding of the 16-bit storage forms")
    ap.add_argument("--forces-ref", default="f64", choices=["f64", "f32"],
                    help="storage form the other forms' coefficients are compared with")
    ap.add_argument("--tune-trials", type=int, default=30, help="steps the autotuner times per launch shape")
    ap.add_argument("--launch", default="",
                    help="fixed launch shapes instead of the autotuner, e.g. f32=256:2,f16=256:34 (block:options)")
    return ap.parse_args()


def build_case(name):
    if name == "backstep":
        return cases.backward_step()
    if name == "cylinder":
        # channel past a cylinder: 400 x 160 x 16, D = 40 at (100, 80), inlet 0.05 with 1e-3
        # noise, outlet + 20-cell sponge, Re = 100, interpolated bounce-back with analytic q
        return cases.cylinder_channel(ibb=True)
    if name == "box19":
        # D3Q19 on a periodic 384 x 384 x 256 box (37.7M cells, the backstep's count), no
        # sub-grid model: the lattice in-place streaming can run
        return cases.periodic_box(384, 384, 256, tau=0.8, ini=(0.02, 0.01, 0.0), sgs=lb.SGS_NONE)
    if name == "box27":
        # D3Q27 cumulant step on a periodic 250^3 box (15.6M cells), no sub-grid model
        return cases.periodic_box(250, 250, 250, tau=0.8, sgs=lb.SGS_NONE, q=27)
    if name == "bluff27":
        # D3Q27 wind tunnel 1024 x 384 x 320 (126M cells) around a seeded procedural body of
        # 199,200 triangles (D = 80), inlet 0.04, Re 300, WALE, FP64 q-distances from the mesh
        return cases.bluff_body()
    return cases.cavity(64)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


ROUNDING = {"stochastic": lb.ROUND_STOCHASTIC, "nearest": lb.ROUND_NEAREST}


STREAMING = {"two-lattice": lb.STREAM_TWO_LATTICE, "aa": lb.STREAM_AA}


def run_form(case, storage, warmup, steps, tune=True, rounding=lb.ROUND_STOCHASTIC, streaming=lb.STREAM_TWO_LATTICE,
             trials=30, launch=None):
    """Warm up, then time exactly `steps` steps bracketed by device synchronisation.  The
    launch shape is the autotuner's choice (trials steps per candidate) or `launch`
    (block, options)."""
    lat = lb.Lattice(case.nx, case.ny, case.nz, q=case.q, storage=storage, sgs=case.sgs, collision=case.coll,
                     omega0=case.omega0, rounding=rounding, streaming=streaming)
    case.upload_to(lat)
    if launch is not None:
        lat.set_launch(*launch)
        lat.tuned = {"block": launch[0], "tune": launch[1], "fixed": True}
    else:
        lat.tuned = lat.autotune(trials, **case.params) if tune else None
    if warmup:
        lat.step(warmup, **case.params)
    lat.synchronize()
    t0 = time.perf_counter()
    lat.step(steps, **case.params)
    lat.synchronize()
    t1 = time.perf_counter()
    dev_ms = lat.last_step_ms()
    return lat, (t1 - t0), dev_ms


def velocity(lat):
    return np.concatenate([lat.download(k).astype(np.float64) for k in ("u", "v", "w")])


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def read_pmc(path=PROFILE_PMC, key="traffic_bytes_per_launch"):
    """HBM bytes per launch (or per step) from a committed counter summary (read and write
    request counters of the step kernel, calibrated per request size), or None."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as fh:
        for line in fh:
            if "=" in line and not line.lstrip().startswith("#"):
                k, v = line.split("=", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals[key])
    except (KeyError, ValueError):
        return None


def hbm_probes():
    """Practical HBM rates of plain streaming kernels (GB/s): copy, read-only and write-only
    over 4 GiB in the launch shapes that reach the highest rates on this part (one float4
    per work-item, non-temporal hints for copy and read)."""
    nb = 4 << 30
    out = {}
    for kind, name, unroll, nt, block, mult in ((0, "copy", 1, 1, 512, 2), (1, "read", 1, 1, 1024, 1),
                                                  (2, "write", 1, 0, 256, 1)):
        ms = lb.probe_stream(kind, unroll, nt, block, 0, nb, reps=10)
        out[name] = round(mult * nb / (ms * 1e-3) / 1e9, 1)
    log(f"[bench] HBM probes GB/s: {out}")
    return out


CPU_SAMPLE_S = 15.0  # seconds of oracle work the CPU baseline aims at (the timing step
# includes the run's set-up, so the measured run comes out shorter)


def cpu_steps(run_one):
    """Steps that fill about CPU_SAMPLE_S seconds, from the time of one (at least 1, at most 200)."""
    t0 = time.perf_counter()
    run_one()
    t1 = time.perf_counter() - t0
    return max(1, min(200, int(round(CPU_SAMPLE_S / max(t1, 1e-6)))))


def cpu_baseline(case):
    """The oracle (CPU restatement of kernel.cl) on the host cores, same lattice; one step
    times the rest, then a run of about CPU_SAMPLE_S seconds is measured."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po  # noqa: E402
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nsteps = cpu_steps(lambda: po.run19(case, 1, order=po.JACOBI))
    t0 = time.perf_counter()
    po.run19(case, nsteps, order=po.JACOBI)
    dt = time.perf_counter() - t0
    return {"value": nsteps * case.N / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{nsteps} steps of the full {case.nx}x{case.ny}x{case.nz} lattice, oracle FP32 "
                      f"(Jacobi order, OpenMP {threads} threads), {dt:.1f} s; the oracle is a C "
                      f"restatement of kernel.cl, not the reference's own code, whose OpenCL kernel "
                      f"has no CPU runtime here"}


def ensemble_main(a):
    """Batched ensemble: many cubic cavity and cylinder-channel lattices (edges 24-96, Re
    50-400, seeded), D3Q19 FP16-deviation storage, advanced by one launch per step over a
    packed block table; compared with one launch per lattice per step."""
    specs = cases.ensemble_specs(a.domains)
    log(f"[bench] ensemble of {len(specs)} lattices, {sum(n ** 3 for _, n, _ in specs) / 1e6:.1f}M cells")
    t0 = time.perf_counter()
    b = lb.Batch([(sp[1], sp[1], sp[1], cases.ensemble_omega0(sp)) for sp in specs], q=19, storage=lb.STORE_F16DEV,
                 sgs=cases.SGS_WALE)
    params = []
    fluid = 0
    sample = []  # lattices for the CPU baseline
    for d, sp in enumerate(specs):  # build and upload one lattice at a time
        c = cases.ensemble_case(sp, d)
        c.upload_to(b.lattice[d])
        params.append(c.params)
        fluid += c.N - int(c.arrays["solid"].sum())
        if d < 8:
            sample.append(c)
        if d % 64 == 63:
            log(f"[bench] built and uploaded {d + 1} lattices")
    log(f"[bench] built and uploaded in {time.perf_counter() - t0:.0f} s")
    N = b.N
    # launch shape: the engine times one work-group per block against persistent work-groups,
    # plain launches against captured graphs, and keeps the fastest
    grid, chunk, ms = b.autotune(params, 32)
    best = (ms, grid, chunk)
    log(f"[bench] batch autotune: grid={grid} graph_chunk={chunk}, {ms:.3f} ms/step, {N / ms / 1e3:.0f} MLUPS")
    if a.warmup:
        b.step(a.warmup, params)
    b.lattice[0].synchronize()
    t0 = time.perf_counter()
    b.step(a.steps, params)
    b.lattice[0].synchronize()
    wall = time.perf_counter() - t0
    dev_ms = b.last_step_ms()
    # the same lattices, one launch per lattice per step (each lattice's own lbffs_step)
    ref_steps = max(4, a.steps // 10)
    for lat, p in zip(b.lattice, params):
        lat.step(1, **p)
    b.lattice[0].synchronize()
    t1 = time.perf_counter()
    for _ in range(ref_steps):
        for lat, p in zip(b.lattice, params):
            lat.step(1, **p)
    b.lattice[0].synchronize()
    wall_single = (time.perf_counter() - t1) / ref_steps
    b.close()
    mlups = a.steps * N / wall / 1e6
    per_step = dev_ms / a.steps * 1e-3
    traffic = read_pmc(PROFILE_ENS_PMC, "traffic_bytes_per_step")
    cpu = None
    if not a.no_cpu_baseline:
        log("[bench] CPU baseline (oracle on host cores, the first 8 lattices)")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po  # noqa: E402
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        nsteps = cpu_steps(lambda: [po.run19(c, 1, order=po.JACOBI) for c in sample])
        t0 = time.perf_counter()
        for c in sample:
            po.run19(c, nsteps, order=po.JACOBI)
        dt = time.perf_counter() - t0
        cells = sum(c.N for c in sample)
        cpu = {"value": nsteps * cells / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": f"{nsteps} steps of the first 8 lattices ({cells} cells), oracle FP32 (the oracle has no "
                         f"16-bit storage form; Jacobi order, OpenMP {threads} threads), {dt:.1f} s; a C restatement "
                         f"of kernel.cl, not the reference's own code"}
    out = {
        "metric": "aggregate MLUPS of a batched ensemble, D3Q19 FP16-deviation storage",
        "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f16-storage/f32-math", "data": "synthetic",
        "config": {"workload": f"ensemble of {len(specs)} cubic lattices (cavity / cylinder channel, edge 24-96, "
                               f"Re 50-400, seeded)", "cells": N, "storage": "f16",
                   "launch": {"grid": best[1], "graph_chunk": best[2]}, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(fluid * ENS_BYTES / per_step / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(fluid * ENS_BYTES / per_step / HBM_PEAK, 4), "traffic": traffic,
                     "bytes_per_cell_update": ENS_BYTES,
                     "bytes_per_cell_terms": "19 FP16 populations read and written (76) + WALE's FP32 "
                                             "velocity, written for the next step and read back (24)",
                     "fluid_cells": fluid, "kernel": "batch_step_kernel",
                     "step_ms": round(per_step * 1e3, 4)},
        "one_launch_per_lattice": {"ms_per_step": round(wall_single * 1e3, 3),
                                   "mlups": round(N / wall_single / 1e6, 1),
                                   "batch_speedup": round(wall_single / (wall / a.steps), 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def force_history(case, storage, nsteps, every=50):
    """Drag and lift coefficients every `every` steps, from the FP64 device sums of the
    solid-link momentum: C = F / (0.5 rho u_in^2 A), A = D nz for the cylinder, the
    frontal area of the mesh for a bluff body."""
    coll = lb.COLL_CUMULANT if case.q == 27 else lb.COLL_RR
    lat = lb.Lattice(case.nx, case.ny, case.nz, q=case.q, storage=storage, sgs=case.sgs, collision=coll,
                     omega0=case.omega0)
    case.upload_to(lat)
    D, u = case.meta["D"], case.meta["u_in"]
    norm = 0.5 * u * u * case.meta.get("frontal_area", D * case.nz)
    cd, cl = [], []
    t0 = time.perf_counter()
    t_log = t0
    for n in range(nsteps // every):
        lat.step(every, **case.params)
        s = lat.reduce()
        cd.append(s[0] / norm)
        cl.append(s[1] / norm)
        if time.perf_counter() - t_log > 30:
            t_log = time.perf_counter()
            log(f"[bench]   step {(n + 1) * every}: Cd {cd[-1]:.4f} Cl {cl[-1]:.4f} ({t_log - t0:.0f} s)")
    lat.close()
    return np.array(cd), np.array(cl)


def shedding(cd, cl, every, D, u):
    """Mean Cd, rms Cl and the Strouhal number of the second half of the history (the
    dominant frequency of Cl, by FFT with a parabolic peak refinement)."""
    h = len(cl) // 2
    c = cl[h:] - cl[h:].mean()
    spec = np.abs(np.fft.rfft(c * np.hanning(len(c))))
    k = int(np.argmax(spec[1:])) + 1
    if 1 <= k < len(spec) - 1:
        a, b, g = spec[k - 1], spec[k], spec[k + 1]
        k = k + 0.5 * (a - g) / (a - 2 * b + g)
    f = k / (len(c) * every)
    return {"Cd": float(cd[h:].mean()), "Cl_rms": float(np.sqrt(np.mean(c * c))), "St": float(f * D / u)}


def forces_report(case, nsteps, forms, ref_name="f64"):
    every = 50
    D, u = case.meta["D"], case.meta["u_in"]
    out = {}
    runs = dict(forms)
    runs.setdefault(ref_name, lb.STORE_F64 if ref_name == "f64" else lb.STORE_F32)
    for name, st in runs.items():
        log(f"[bench] forces {name}: {nsteps} steps")
        cd, cl = force_history(case, st, nsteps, every)
        out[name] = shedding(cd, cl, every, D, u)
    ref = out[ref_name]
    for name in runs:
        if name == ref_name:
            continue
        r = out[name]
        r[f"Cd_rel_err_vs_{ref_name}"] = abs(r["Cd"] - ref["Cd"]) / abs(ref["Cd"])
        r[f"St_rel_err_vs_{ref_name}"] = abs(r["St"] - ref["St"]) / abs(ref["St"]) if ref["St"] else None
        r[f"Cl_rms_rel_err_vs_{ref_name}"] = abs(r["Cl_rms"] - ref["Cl_rms"]) / max(abs(ref["Cl_rms"]), 1e-12)
    out["steps"] = nsteps
    out["sample_every"] = every
    out["reference_form"] = ref_name
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "ensemble":
        return ensemble_main(a)
    log(f"[bench] building workload {a.workload}")
    t_build = time.perf_counter()
    case = build_case(a.workload)
    log(f"[bench] built {case.nx}x{case.ny}x{case.nz} in {time.perf_counter() - t_build:.1f} s")
    N = case.N
    fluid = int(N - int(case.arrays["solid"].sum()))
    forms = {"f32": lb.STORE_F32, "f16": lb.STORE_F16DEV, "bf16": lb.STORE_BF16DEV}
    default_forms = ["f32", "bf16"] if a.workload == "bluff27" else ["f32", "f16"]
    run = [f for f in (a.only.split(",") if a.only else default_forms) if f]
    res = {}
    vel = {}
    fixed = {}
    for item in (x for x in a.launch.split(",") if x):
        k, v = item.split("=")
        b, t = v.split(":")
        fixed[k] = (int(b), int(t))
    for name in run:
        log(f"[bench] {name}: warmup {a.warmup}, timed {a.steps}")
        lat, wall, dev_ms = run_form(case, forms[name], a.warmup, a.steps, rounding=ROUNDING[a.rounding],
                                     streaming=STREAMING[a.streaming], trials=a.tune_trials,
                                     launch=fixed.get(name))
        # populations in and out, and with WALE the FP32 velocity its gradient reads from the
        # previous step: written once and read once per cell update
        bytes_per_cell = 2 * case.q * (4 if name == "f32" else 2) + (24 if case.sgs == lb.SGS_WALE else 0)
        launch_ms = dev_ms / a.steps
        res[name] = {"wall_s": wall, "dev_ms": dev_ms, "launch_ms": launch_ms, "launch": lat.tuned,
                     "mlups": a.steps * N / wall / 1e6, "bytes_per_cell": bytes_per_cell,
                     "achieved_Bps": fluid * bytes_per_cell / (launch_ms * 1e-3)}
        if not a.no_error:
            vel[name] = velocity(lat)
        lat.close()
        log(f"[bench] {name}: {res[name]['mlups']:.0f} MLUPS, {launch_ms:.3f} ms/launch, "
            f"{res[name]['achieved_Bps'] / 1e9:.0f} GB/s")
    err = {}
    if not a.no_error and vel:
        log("[bench] f64 run for the error reference")
        lat, _, _ = run_form(case, lb.STORE_F64, a.warmup, a.steps, tune=False)
        v64 = velocity(lat)
        lat.close()
        for k, v in vel.items():
            err[k] = rel(v, v64)
    probes = hbm_probes()
    # the memory skeleton of this lattice's step (the same pulls and stores, no collision):
    # the rate the access pattern itself reaches on this GPU
    skel = {}
    if N % 256 == 0:
        for name in res:
            sb = 4 if name == "f32" else 2
            ms = lb.probe_lattice(case.q, sb, 0, case.nx, case.ny, case.nz, reps=10)
            skel[name] = {"ms": round(ms, 4), "kernel_over_skeleton": round(res[name]["launch_ms"] / ms, 3)}
        log(f"[bench] memory skeleton: {skel}")
    # the bluff body is quoted in its 16-bit form (BF16 deviations, 108 + 24 B per cell update)
    main_form = "bf16" if a.workload == "bluff27" and "bf16" in res else ("f32" if "f32" in res else run[0])
    r = res[main_form]
    pmc = PMC_OF_WORKLOAD.get(a.workload)
    traffic = read_pmc(pmc, "traffic_bytes_per_launch" + ("" if main_form == "f32" else "_" + main_form)) if pmc else None
    out = {
        "metric": "MLUPS + % of HBM roofline, FP32 vs FP16-storage D3Q19 at stated error",
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["wall_s"] / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_form == "f32" else ("f16-storage/f32-math" if main_form == "f16" else "bf16-storage/f32-math"),
        "data": "synthetic",
        "config": {"workload": {"backstep": "backward-facing step D3Q19 Smagorinsky LES Re_H=5500",
                                "cylinder": "channel past cylinder D3Q19 Re=100, interpolated bounce-back with analytic q",
                                "cavity": "lid-driven cavity D3Q19 Re=100",
                                "box19": "periodic box D3Q19 recursive-regularized, no SGS",
                                "box27": "periodic box D3Q27 cumulant, no SGS",
                                "bluff27": "bluff body D3Q27 cumulant WALE Re=300, procedural STL (199,200 "
                                           "triangles), FP64 q-distances, interpolated bounce-back"}[a.workload],
                   "lattice": [case.nx, case.ny, case.nz], "cells": N, "fluid_cells": fluid,
                   "collision": "cumulant" if case.q == 27 else "recursive-regularized",
                   "sgs": {0: "WALE", 1: "Smagorinsky", 2: "none"}[case.sgs],
                   "storage": main_form, "rounding_16bit": a.rounding, "streaming": a.streaming,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved_Bps"] / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(r["achieved_Bps"] / HBM_PEAK, 4),
                     "traffic": traffic,
                     "bytes_per_cell_update": r["bytes_per_cell"],
                     "kernel": "step27_kernel" if case.q == 27 else "step19_kernel",
                     "launch_ms": round(r["launch_ms"], 4), "probes_GBps": probes,
                     "memory_skeleton": skel.get(main_form)},
        "forms": {k: {"mlups": round(v["mlups"], 1), "launch_ms": round(v["launch_ms"], 4),
                      "hbm_frac": round(v["achieved_Bps"] / HBM_PEAK, 4),
                      "velocity_rel_l2_vs_f64": (round(err[k], 8) if k in err else None),
                      "launch": v["launch"],
                      "traffic": (read_pmc(pmc, "traffic_bytes_per_launch" + ("" if k == "f32" else "_" + k))
                                  if pmc else None),
                      "memory_skeleton_ms": skel.get(k, {}).get("ms"),
                      "error_steps": a.warmup + a.steps}
                  for k, v in res.items()},
    }
    if "f16" in res and "f32" in res:
        out["forms"]["f16_over_f32"] = round(res["f16"]["mlups"] / res["f32"]["mlups"], 3)
    if "bf16" in res and "f32" in res:
        out["forms"]["bf16_over_f32"] = round(res["bf16"]["mlups"] / res["f32"]["mlups"], 3)
    if case.surface is not None or "qlink_misses" in case.meta:
        out["config"]["qlink_misses"] = case.meta.get("qlink_misses")
        out["config"]["triangles"] = case.meta.get("triangles")
    if a.forces and a.workload in ("cylinder", "bluff27"):
        out["forces"] = forces_report(case, a.forces, {k: forms[k] for k in run}, a.forces_ref)
    if not a.no_cpu_baseline:
        log("[bench] CPU baseline (oracle on host cores)")
        if a.workload == "bluff27":
            # the same wind tunnel scaled to 256 x 96 x 80 cells (body D = 20), q-distances
            # as a host array for the oracle
            small = cases.bluff_body(nx=256, ny=96, nz=80, D=20.0, n_lat=80, n_lon=128, spz_cells=16,
                                     qlinks_on_device=False)
            out["cpu_baseline"] = cpu_baseline(small)
        else:
            out["cpu_baseline"] = cpu_baseline(case)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
