#!/usr/bin/env python3
"""SPH + self-gravity steps/s on one MI355X: the 1,000,000-particle giant
impact (900k-particle target of iron core + dunite mantle, 100k-particle
impactor of the same structure, ANEOS iron / dunite tables regridded to
200 x 200 log-spaced cells, 1.1 v_esc at 45 deg, theta = 0.5).

One step is src/main.cpp:125-173: kick/drift/predict, tree + interaction
lists, 4 density passes, EoS, derivative (+ Balsara), hydro force, tree
gravity, external force, final kick, dt reduction -- all on the GPU with the
particles resident in HBM.  The FP64 path is timed first (`value`), then the
mixed-precision path on the same initial state; both report their dominant
kernel against the MI355X roofline.  The CPU baseline is the oracle (the
reference's functors restated in C, OpenMP over i-groups, the reference's
compiler flags) timed for one step of the same workload on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32 (published; includes packed FP32)
GRAV_FLOPS = 38            # flops per P2P / P2M interaction (acc + potential)
# algorithmic bytes per i-particle (read + result) and per j-list entry, FP64 fields
SPH_BYTES = {"dens": (40 + 16, 32), "drvt": (64 + 40, 56), "hydro": (88 + 40, 104)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_tables(ctx, grid):
    from fdps_sph_amd import eos_tables as E
    ctx.load_table(0, E.PATHS["granite"])
    ctx.set_table(1, E.table("iron", grid))
    ctx.set_table(2, E.table("dunite", grid))


def energy(sim):
    m = sim.get("MASS")
    v = sim.get("VEL")
    u = sim.get("ENG")
    pot = sim.get("POT")
    return float(np.sum(m * (u + 0.5 * np.sum(v * v, 1))) + 0.5 * np.sum(m * pot))


def kernel_roofline(stats, mixed):
    """Achieved throughput of every interaction kernel from its live stats."""
    out = {}
    for k, s in stats.items():
        if s["launches"] == 0 or s["ms"] <= 0:
            continue
        sec = s["ms"] * 1e-3
        if k == "grav":
            flops = GRAV_FLOPS * s["pairs"]
            peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
            ach = flops / sec / 1e12
            out[k] = {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                      "avg_ms": s["ms"] / s["launches"], "per_launch": {"interactions": s["pairs"] / s["launches"],
                                                                        "flops": flops / s["launches"]}}
        else:
            bi, bj = SPH_BYTES[k]
            nbytes = bi * s["n_i"] + bj * s["n_j"]
            ach = nbytes / sec / 1e9
            out[k] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                      "avg_ms": s["ms"] / s["launches"], "per_launch": {"bytes": nbytes / s["launches"],
                                                                        "n_i": s["n_i"] / s["launches"],
                                                                        "n_j": s["n_j"] / s["launches"],
                                                                        "pairs": s["pairs"] / s["launches"]}}
    return out


def run_gpu(ic, precision, steps, warmup, theta, grid):
    from fdps_sph_amd import Context, Simulation
    ctx = Context(0, precision)
    load_tables(ctx, grid)
    t0 = time.perf_counter()
    sim = Simulation(ctx, ic, theta=theta)
    log(f"[{precision}] init (upload + first force evaluation) {time.perf_counter() - t0:.2f} s")
    init_state = {k: sim.get(k) for k in ("ID", "DENS", "ACC", "GRAV", "ENGDOT", "SMTH")}
    e0 = energy(sim)
    if warmup:
        sim.step(warmup)
    sim.set_stats(True)
    # sphgpu_sim_step synchronises the context's stream before it returns, so the
    # wall clock around it brackets exactly the K steps
    t0 = time.perf_counter()
    sim.step(steps)
    wall = time.perf_counter() - t0
    stats = sim.kernel_stats()
    stages = sim.stage_ms()
    e1 = energy(sim)
    dt, t = sim.time()
    log(f"[{precision}] {steps} steps in {wall:.3f} s ({1e3 * wall / steps:.2f} ms/step), dt {dt[0]:.4g} s")
    res = {"wall": wall, "ms_per_step": 1e3 * wall / steps, "steps_per_s": steps / wall, "stats": stats,
           "stages": stages, "init": init_state, "E0": e0, "E1": e1, "time": float(t[0]), "ngroups": None}
    ctx.close()
    return res


def rel_errors(a, b, ids_a, ids_b):
    """Per-particle relative error of a against b (matched by particle id)."""
    ia = np.argsort(ids_a)
    ib = np.argsort(ids_b)
    a = a[ia]
    b = b[ib]
    if a.ndim == 2:
        num = np.linalg.norm(a - b, axis=1)
        den = np.linalg.norm(b, axis=1)
    else:
        num = np.abs(a - b)
        den = np.abs(b)
    e = num / np.maximum(den, 1e-300)
    return {"max": float(e.max()), "p99": float(np.percentile(e, 99)), "rms": float(np.sqrt(np.mean(e * e)))}


def cpu_baseline(ic, grid, theta):
    """One step of the same workload on the CPU oracle (reference flags, OpenMP)."""
    from fdps_sph_amd import eos_tables as E
    from oracle import oracle as O
    n = len(ic["mass"])
    s = O.State(n)
    s.x[:], s.y[:], s.z[:] = ic["pos"].T
    s.vx[:], s.vy[:], s.vz[:] = ic["vel"].T
    s.m[:] = ic["mass"]
    s.rho[:] = ic["dens"]
    s.u[:] = ic["eng"]
    s.mat[:] = ic["mat"]
    tabs = [O.Table(E.PATHS["granite"], kind="fast"), O.Table(cells=E.table("iron", grid), kind="fast"),
            O.Table(cells=E.table("dunite", grid), kind="fast")]
    sim = O.Sim(s, tabs, O.params(theta=theta), kind="fast")
    t0 = time.perf_counter()
    sim.init()
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    sim.step()
    t_step = time.perf_counter() - t0
    log(f"[cpu oracle] {O.num_threads()} threads: init {t_init:.2f} s, one step {t_step:.2f} s")
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"one full step of the same {n:,}-particle workload after the initial force evaluation "
                      f"({t_step:.2f} s); CPU oracle = the reference's functors restated in C, -O3 -ffast-math "
                      "-funroll-loops -fopenmp, OpenMP over i-groups"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000, help="total particles (90/10 target/impactor)")
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--grid", type=int, default=200, help="EoS table size (grid x grid)")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is a design in DESIGN.md, not a measured configuration", file=sys.stderr)
        sys.exit(2)

    from fdps_sph_amd import ic as IC
    nt = int(round(0.9 * args.n))
    t0 = time.perf_counter()
    ic = IC.iron_dunite_impact(nt, args.n - nt)
    log(f"initial conditions: {args.n:,} particles in {time.perf_counter() - t0:.1f} s")

    runs = {}
    for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        runs[prec] = run_gpu(ic, prec, args.steps, args.warmup, args.theta, args.grid)

    main_prec = "fp64" if "fp64" in runs else "mixed"
    R = runs[main_prec]
    roof = kernel_roofline(R["stats"], main_prec == "mixed")
    dominant = max(roof, key=lambda k: roof[k]["avg_ms"] * R["stats"][k]["launches"])
    rl = roof[dominant]
    roofline = {"bound": rl["bound"], "achieved": rl["achieved"], "peak": rl["peak"], "unit": rl["unit"],
                "frac": rl["frac"], "traffic": None, "kernel": dominant, "avg_ms": rl["avg_ms"],
                "per_launch": rl["per_launch"]}
    stage_names = ["kick_drift", "sort_tree", "density_x4", "eos", "derivative", "hydro", "gravity", "final_dt"]
    out = {
        "metric": "SPH+gravity steps/s at 1M particles: mixed-precision vs FP64, with err vs FP64",
        "value": R["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32",
        "data": "synthetic (seeded lattice bodies)",
        "config": {"workload": f"giant impact, {args.n:,} particles: {nt:,}-particle iron-core/dunite-mantle "
                               f"target + {args.n - nt:,}-particle impactor, ANEOS {args.grid}x{args.grid} tables, "
                               f"1.1 v_esc at 45 deg, theta={args.theta}", "precision": main_prec,
                   "n_group": 64, "n_leaf": 16},
        "roofline": roofline,
        "kernels": {k: {kk: vv for kk, vv in v.items() if kk != "per_launch"} for k, v in roof.items()},
        "stages_ms": dict(zip(stage_names, [round(float(x), 3) for x in R["stages"]])),
    }
    if "mixed" in runs and "fp64" in runs:
        M, F = runs["mixed"], runs["fp64"]
        ids_m, ids_f = M["init"]["ID"], F["init"]["ID"]
        acc_m = M["init"]["ACC"] + M["init"]["GRAV"]
        acc_f = F["init"]["ACC"] + F["init"]["GRAV"]
        err = {"dens": rel_errors(M["init"]["DENS"], F["init"]["DENS"], ids_m, ids_f),
               "acc": rel_errors(acc_m, acc_f, ids_m, ids_f),
               "acc_hydro": rel_errors(M["init"]["ACC"], F["init"]["ACC"], ids_m, ids_f),
               "acc_grav": rel_errors(M["init"]["GRAV"], F["init"]["GRAV"], ids_m, ids_f),
               "du_dt": rel_errors(M["init"]["ENGDOT"], F["init"]["ENGDOT"], ids_m, ids_f)}
        mroof = kernel_roofline(M["stats"], True)
        out["mixed"] = {
            "steps_per_s": M["steps_per_s"], "ms_per_step": M["ms_per_step"],
            "speedup_vs_fp64": M["steps_per_s"] / F["steps_per_s"],
            "err_vs_fp64_first_force_eval": err,
            "energy_drift_vs_fp64": abs(M["E1"] - F["E1"]) / abs(F["E1"]),
            "energy_fp64_rel_change": abs(F["E1"] - F["E0"]) / abs(F["E0"]),
            "kernels": {k: {kk: vv for kk, vv in v.items() if kk != "per_launch"} for k, v in mroof.items()},
            "stages_ms": dict(zip(stage_names, [round(float(x), 3) for x in M["stages"]])),
        }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ic, args.grid, args.theta)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
