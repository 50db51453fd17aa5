#!/usr/bin/env python3
"""SPH + self-gravity steps/s on one MI355X, plain FP64 path and the
mixed-precision form, with the mixed form's error against FP64.

One step is src/main.cpp:125-173: kick/drift/predict, tree + interaction
lists, 4 density passes, EoS, derivative (+ Balsara), hydro force, tree
gravity, external force, final kick, dt reduction -- all on the GPU with the
particles resident in HBM.  Each step is timed on its own (the step call
returns after the context's stream has drained); `value` is the FP64 path's
steps/s over the K timed steps, `mixed` carries the same figures for the
mixed-precision form.

Workloads (--workload):
  impact1m     1,000,000 particles: 900k iron-core/dunite-mantle target + 100k
               impactor of the same structure, ANEOS tables regridded to
               200 x 200 log-spaced cells, 1.1 v_esc at 45 deg, theta 0.5,
               10 warm-up + 90 timed steps (default)
  granite10k   9,000 + 1,000 granite particles, 9 km/s at 45 deg, 50 steps;
               the CPU baseline runs all 50 steps
  theta-sweep  impact1m at theta 0.3, 0.5 and 0.7, 100 steps each precision
  batch64      64 independent granite impacts, log-uniform sizes in
               [5k, 200k] (3.1M particles), speeds U[5,15] km/s, angles
               U[0,75] deg, one launch per stage for all systems; n_group and
               n_leaf picked by fdps_sph_amd.autotune
  impact16m    12M iron/dunite target + 4M granite impactor, 20 steps
  functor1m    the drop-in boundary: the density pass of impact1m through the
               multiwalk dispatch / retrieve entry points (host EPI/EPJ arrays
               in the reference's layout, 200 walks per call, two tags in
               flight), PCIe included; reported beside the device-resident rate

Mixed-precision error (`mixed.err`): per-particle relative error of density,
acceleration (hydro, gravity, total) and du/dt against the FP64 path from the
same state, at the initial state and at the FP64 state after the last step;
`energy_drift` is |E_mixed - E_fp64| / |E_fp64| after the same number of
steps.  The CPU baseline is the oracle (the reference's functors restated in
C with the reference's compiler flags, OpenMP over i-groups) on the same
workload, on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "SPH+gravity steps/s at 1M particles: mixed-precision vs FP64, with err vs FP64"
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X vector FP64
FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32
GRAV_FLOPS = 38            # flops per P2P / P2M interaction (acc + potential)
# algorithmic bytes per i-particle (read + result) and per j-list entry, FP64
# fields.  hydro: 12 inputs, acc, du and dt out (dt only when it forms v_sig_max
# itself); vsig: position, velocity, c_s and h in, dt out; per entry position,
# velocity and c_s
SPH_BYTES = {"dens": (40 + 16, 32), "drvt": (64 + 40, 56), "hydro": (96 + 40, 104), "dens_fp64": (40 + 16, 32),
             "vsig": (64 + 8, 56)}
HYDRO_BYTES_SPLIT = (96 + 32, 104)
# kernel -> (symbol prefix, FP64 kernel even in the precision form)
KERNEL_SYMBOL = {"dens": ("k_dens", False), "drvt": ("k_drvt", False), "hydro": ("k_hydro", False),
                 "grav": ("k_grav", False), "dens_fp64": ("k_dens", True), "vsig": ("k_vsig", True)}


def kernel_symbol(kernel, mixed):
    """The device symbol of an interaction kernel (k_hydro, k_hydro_mixed, ...),
    as rocprofv3 and profiles/hbm_traffic.json name it."""
    sym, always64 = KERNEL_SYMBOL[kernel]
    return f"{sym}_mixed" if mixed and not always64 else sym
STAGES = ["kick_drift", "sort_tree", "density_x4", "eos", "derivative", "gravity", "hydro", "final_dt"]
# rocprofv3 counter summary of the interaction kernels (per launch): HBM
# traffic from FETCH_SIZE / WRITE_SIZE (copy-calibrated), VALU issue rate,
# counted FP32 / FP64 VALU flops; the commands that made it are in the file
COUNTERS_FILE = os.path.join(ROOT, "profiles", "counters_1M_iron_dunite.json")

DEFAULTS = {  # workload: (warmup, steps)
    "impact1m": (10, 90), "granite10k": (5, 45), "theta-sweep": (10, 90), "batch64": (5, 20),
    "impact16m": (2, 18), "functor1m": (1, 3)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ workloads
def make_workload(name, n=None):
    from fdps_sph_amd import ic as IC
    t0 = time.perf_counter()
    if name in ("impact1m", "theta-sweep", "functor1m"):
        n = n or 1_000_000
        nt = int(round(0.9 * n))
        ic = IC.iron_dunite_impact(nt, n - nt)
        desc = (f"giant impact, {n:,} particles: {nt:,}-particle iron-core/dunite-mantle target + {n - nt:,}-particle "
                "impactor, ANEOS iron/dunite tables regridded to 200x200, 1.1 v_esc at 45 deg")
        tables = {"granite": None, "iron": 200, "dunite": 200}
    elif name == "granite10k":
        ic = IC.granite_impact(9000, 1000)
        desc = ("giant impact, 10,000 granite particles: 9,000-particle target (R 6,371 km) + 1,000-particle "
                "impactor (R 3,390 km), granite table, 9 km/s at 45 deg")
        tables = {"granite": None, "iron": None, "dunite": None}
    elif name == "batch64":
        ic = IC.batch_of_impacts(64)
        desc = (f"64 independent granite impacts, {int(ic['counts'].sum()):,} particles (log-uniform "
                f"{int(ic['counts'].min()):,}..{int(ic['counts'].max()):,} per system), speeds U[5,15] km/s, "
                "angles U[0,75] deg")
        tables = {"granite": None, "iron": None, "dunite": None}
    elif name == "impact16m":
        n = n or 16_000_000
        nt = int(round(0.75 * n))
        ic = IC.three_material_impact(nt, n - nt)
        desc = (f"giant impact, {n:,} particles: {nt:,}-particle iron-core/dunite-mantle target + {n - nt:,}-particle "
                "granite impactor, ANEOS tables regridded to 200x200, 1.1 v_esc at 45 deg")
        tables = {"granite": 200, "iron": 200, "dunite": 200}
    else:
        raise SystemExit(f"unknown workload {name}")
    log(f"[{name}] initial conditions: {len(ic['mass']):,} particles in {time.perf_counter() - t0:.1f} s")
    return ic, desc, tables


def load_tables(ctx, tables):
    from fdps_sph_amd import eos_tables as E
    for mat, name in enumerate(("granite", "iron", "dunite")):
        ctx.set_table(mat, E.table(name, tables[name]))


def energy_parts(sim):
    """Kinetic, internal and potential (1/2 sum m phi) energy of all particles."""
    m = sim.get("MASS")
    v = sim.get("VEL")
    u = sim.get("ENG")
    pot = sim.get("POT")
    return {"kinetic": float(np.sum(m * 0.5 * np.sum(v * v, 1))), "internal": float(np.sum(m * u)),
            "potential": float(0.5 * np.sum(m * pot))}


def energy(sim):
    return sum(energy_parts(sim).values())


def snapshot(sim):
    """The particle state of a simulation as initial conditions."""
    return {"pos": sim.get("POS"), "vel": sim.get("VEL"), "mass": sim.get("MASS"), "dens": sim.get("DENS"),
            "eng": sim.get("ENG"), "mat": sim.get("MAT"), "sys": sim.get("SYS"), "id": sim.get("ID")}


FORCE_FIELDS = ("ID", "DENS", "ACC", "GRAV", "ENGDOT", "PRES", "DT", "DIVV", "SMTH")


def force_fields(sim):
    return {k: sim.get(k) for k in FORCE_FIELDS}


# ------------------------------------------------------------------ roofline
def kernel_roofline(stats, mixed):
    """Achieved throughput of every interaction kernel from its live HIP-event stats."""
    out = {}
    for k, s in stats.items():
        if s["launches"] == 0 or s["ms"] <= 0:
            continue
        sec = s["ms"] * 1e-3
        L = s["launches"]
        if k == "grav":
            flops = GRAV_FLOPS * s["pairs"]
            peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
            ach = flops / sec / 1e12
            out[k] = {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                      "avg_ms": s["ms"] / L, "launches": L,
                      "per_launch": {"interactions": s["pairs"] / L, "flops": flops / L, "n_i": s["n_i"] / L,
                                     "n_epj": s["n_j"] / L, "n_spj": s["n_s"] / L}}
        else:
            split = k == "hydro" and stats.get("vsig", {}).get("launches", 0) > 0
            bi, bj = HYDRO_BYTES_SPLIT if split else SPH_BYTES[k]
            nbytes = bi * s["n_i"] + bj * s["n_j"]
            ach = nbytes / sec / 1e9
            out[k] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": ach / HBM_PEAK_GBS, "avg_ms": s["ms"] / L, "launches": L,
                      "per_launch": {"bytes": nbytes / L, "n_i": s["n_i"] / L, "n_j": s["n_j"] / L,
                                     "pairs_in_tiles": s["pairs"] / L}}
    return out


def measured_counters(kernel, precision, workload):
    """Counter figures per launch of a kernel from the committed rocprofv3
    --pmc summary of the same workload: HBM bytes (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy-kernel calibration), VALU instructions per CU per cycle
    and counted VALU flops.  None where the summary has no such entry."""
    try:
        with open(COUNTERS_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("workload") != workload:
        return None
    return d.get("kernels", {}).get(kernel_symbol(kernel, precision == "mixed"))


# ------------------------------------------------------------------ runs
def run(ctx, ic, warmup, steps, label, precision, **params):
    """Simulation from `ic` in the given precision form: warm-up steps, then
    `steps` steps timed as one call (the library's stream drains at its end),
    then a diagnostics pass of a few more steps timed one by one with the
    per-kernel statistics on."""
    from fdps_sph_amd import Simulation
    ctx.set_precision(precision)
    t0 = time.perf_counter()
    sim = Simulation(ctx, ic, **params)
    t_init = time.perf_counter() - t0
    init = force_fields(sim)
    parts0 = energy_parts(sim)
    e0 = sum(parts0.values())
    dts = []
    for _ in range(warmup):
        sim.step(1)
        dts.append(float(np.min(sim.time()[0])))
    g0 = sim.graph_stats()
    t_all = time.perf_counter()
    sim.step(steps)  # returns after the context's stream has drained
    wall = time.perf_counter() - t_all
    g1 = sim.graph_stats()
    graph = {"replayed_steps_timed": g1[0] - g0[0], "captures_timed": g1[1] - g0[1], "captures_total": g1[1]}
    stages_timed = sim.stage_ms()  # the last timed step (list walks overlapping kernels)
    parts1 = energy_parts(sim)
    e1 = sum(parts1.values())
    dt, t = sim.time()
    gb, fe = sim.list_builds()
    state = snapshot(sim)
    # diagnostics pass: per-step wall times and per-kernel HIP-event statistics
    nstat = min(steps, 20)
    sim.set_stats(True)
    ms = []
    for _ in range(nstat):
        t0 = time.perf_counter()
        sim.step(1)
        ms.append(1e3 * (time.perf_counter() - t0))
    stats = sim.kernel_stats()
    stages = sim.stage_ms()
    ms = np.array(ms)
    log(f"[{label}] init {t_init:.2f} s; {steps} steps in {wall:.3f} s: {1e3 * wall / steps:.2f} ms/step; "
        f"{nstat} more steps one by one with kernel statistics: median {np.median(ms):.2f}, min {ms.min():.2f}; "
        f"dt {float(np.min(dt)):.4g} s, t {float(np.max(t)):.4g} s")
    return {"sim": sim, "state": state, "wall": wall, "ms_per_step": 1e3 * wall / steps,
            "median_ms": float(np.median(ms)),
            "steps_per_s": steps / wall, "stats": stats, "stages": stages, "stages_timed": stages_timed,
            "init": init, "E0": e0, "E1": e1,
            "parts0": parts0, "parts1": parts1, "warmup_dts": dts, "gather_builds_per_force_eval": gb / max(fe, 1),
            "t": float(np.max(t)), "steps_total": warmup + steps, "graph": graph}


def err_stats(a, b, ids_a, ids_b, scale=None):
    """Per-particle relative error of a against b, matched by particle id:
    max / p99 / rms of |a - b| / |b|, the norm-wise error ||a - b|| / ||b||
    over all particles and, with `scale`, max |a - b| / scale."""
    ia, ib = np.argsort(ids_a), np.argsort(ids_b)
    a, b = a[ia], b[ib]
    if a.ndim == 2:
        num = np.linalg.norm(a - b, axis=1)
        den = np.linalg.norm(b, axis=1)
    else:
        num = np.abs(a - b)
        den = np.abs(b)
    ok = den > 0
    e = np.divide(num[ok], den[ok], out=np.zeros(int(ok.sum())), where=num[ok] > 0)
    if e.size == 0:
        e = np.zeros(1)
    out = {"max": float(e.max()), "p99": float(np.percentile(e, 99)), "rms": float(np.sqrt(np.mean(e * e))),
           "norm": float(np.sqrt(np.sum(num * num)) / max(np.sqrt(np.sum(den * den)), 1e-300)),
           "nonzero_where_fp64_zero": int(np.sum((num > 0) & ~ok))}
    if scale is not None:
        s = scale[ib]
        out["max_on_force_scale"] = float(np.max(np.divide(num, s, out=np.zeros_like(num), where=num > 0)))
    return out


def force_errors(m, f):
    """Errors of the mixed force evaluation m against the FP64 one f."""
    ids_m, ids_f = m["ID"], f["ID"]
    # |a_hydro| + |a_grav|: the size of the two forces that sum to the acceleration
    fscale = np.linalg.norm(f["ACC"], axis=1) + np.linalg.norm(f["GRAV"], axis=1)
    return {"dens": err_stats(m["DENS"], f["DENS"], ids_m, ids_f),
            "acc": err_stats(m["ACC"] + m["GRAV"], f["ACC"] + f["GRAV"], ids_m, ids_f, fscale),
            "acc_hydro": err_stats(m["ACC"], f["ACC"], ids_m, ids_f),
            "acc_grav": err_stats(m["GRAV"], f["GRAV"], ids_m, ids_f),
            "du_dt": err_stats(m["ENGDOT"], f["ENGDOT"], ids_m, ids_f),
            "pres": err_stats(m["PRES"], f["PRES"], ids_m, ids_f),
            "dt": err_stats(m["DT"], f["DT"], ids_m, ids_f),
            "div_v": err_stats(m["DIVV"], f["DIVV"], ids_m, ids_f)}


def du_term_scale(sim, ngroups=200, seed=7):
    """Size of the summed terms of du/dt, sum_j m_j (|A_i| + |AV_ij|/2) |dv| |gradW_bar|
    (src/force.h:126-145), for the particles of `ngroups` sampled i-groups of a
    simulation's current state over its exported symmetric lists: the scale
    on which a du/dt whose terms cancel is compared.  Returns (ids, scale)."""
    st = sim.state()
    L = sim.lists(1)
    gs, joff, jidx = L["group_start"], L["j_off"], L["j_idx"]
    G = len(gs) - 1
    groups = np.random.default_rng(seed).choice(G, min(ngroups, G), replace=False)
    x, v, h, rho, p, c = st["pos"], st["vel"], st["smth"], st["dens"], st["pres"], st["snds"]
    f, bal, m = st["gradh"], st["bal"], st["mass"]
    A = f * p / rho ** 2
    norm3 = 1365.0 / (64.0 * np.pi)
    ids, out = [], []

    def dwdr(r, hh):  # |dW/dr| of the Wendland C6 kernel, support 2.5 h
        H = 2.5 * hh
        q = np.minimum(r / H, 1.0)
        return norm3 / H ** 4 * 22.0 * q * (1 + 7 * q + 16 * q * q) * (1 - q) ** 7
    for g in groups:
        ii = np.arange(gs[g], gs[g + 1])
        jj = jidx[joff[g]:joff[g + 1]]
        dr = x[ii][:, None, :] - x[jj][None]
        dv = v[ii][:, None, :] - v[jj][None]
        r = np.linalg.norm(dr, axis=2)
        dvdr = np.sum(dv * dr, axis=2)
        w = np.where(dvdr < 0, dvdr / np.maximum(r, 1e-300), 0.0)
        vsig = c[ii][:, None] + c[jj][None] - 3.0 * w
        av = np.abs(0.5 * vsig * w / (0.5 * (rho[ii][:, None] + rho[jj][None])) * 0.5 * (bal[ii][:, None] + bal[jj][None]))
        gbar = 0.5 * (f[ii][:, None] * dwdr(r, h[ii][:, None]) + f[jj][None] * dwdr(r, h[jj][None]))
        sc = np.sum(m[jj][None] * (np.abs(A[ii])[:, None] + 0.5 * av) * np.linalg.norm(dv, axis=2) * gbar, axis=1)
        ids.append(st["id"][ii])
        out.append(sc)
    return np.concatenate(ids), np.concatenate(out)


def compare_precisions(ctx, ic, warmup, steps, **params):
    """FP64 and mixed runs of the same workload plus the mixed form's errors.
    One context runs both forms: a process has few hardware queues, and a
    second context's two streams can end up sharing one (measured: the list
    walks then no longer overlap the force kernels)."""
    F = run(ctx, ic, warmup, steps, "fp64", "fp64", **params)
    M = run(ctx, ic, warmup, steps, "mixed", "mixed", **params)
    err0 = force_errors(M["init"], F["init"])
    # one force evaluation of each precision from the FP64 state after the last step
    state = F["state"]
    from fdps_sph_amd import Simulation
    ctx.set_precision("fp64")
    sf = Simulation(ctx, state, **params)
    fe = force_fields(sf)
    sids, sscale = du_term_scale(sf)
    ctx.set_precision("mixed")
    mf = force_fields(Simulation(ctx, state, **params))
    me = force_errors(mf, fe)
    # du/dt of the sampled groups' particles on the size of its summed terms
    pf = {k: v for k, v in zip(fe["ID"], fe["ENGDOT"])}
    pm = {k: v for k, v in zip(mf["ID"], mf["ENGDOT"])}
    num = np.array([abs(pm[k] - pf[k]) for k in sids])
    den = np.maximum(np.abs(np.array([pf[k] for k in sids])), sscale)
    e = np.divide(num, den, out=np.zeros_like(num), where=num > 0)
    me["du_dt"]["on_term_scale"] = {"particles": int(len(sids)), "max": float(e.max()),
                                    "p99": float(np.percentile(e, 99)), "rms": float(np.sqrt(np.mean(e * e)))}
    drift = abs(M["E1"] - F["E1"]) / abs(F["E1"])
    log(f"[mixed vs fp64] speedup {M['steps_per_s'] / F['steps_per_s']:.2f}x; acc err at the evolved state "
        f"max {me['acc']['max']:.2e} p99 {me['acc']['p99']:.2e} norm {me['acc']['norm']:.2e}; energy drift over "
        f"{F['steps_total']} steps {drift:.2e}")
    return F, M, {"initial_state": err0, "after_last_step": me}, drift


def stage_dict(stages):
    return dict(zip(STAGES, [round(float(x), 3) for x in stages]))


def precision_block(R, mixed, workload):
    roof = kernel_roofline(R["stats"], mixed)
    for k, v in roof.items():
        c = measured_counters(k, "mixed" if mixed else "fp64", workload)
        v["traffic"] = None if c is None else c.get("hbm_bytes_per_launch")
        # the other roof: fraction of the CU's one-wave64-VALU-instruction-per-cycle issue rate
        v["valu_issue_frac"] = None if c is None else c.get("valu_issue")
        if k == "grav" and c is not None and c.get("valu_flops_per_launch"):
            # flops the VALU counters count per launch (FP64 or FP32), over this run's kernel time
            cf = c["valu_flops_per_launch"] / (v["avg_ms"] * 1e-3) / 1e12
            v["counter_tflops"] = cf
            v["counter_frac"] = cf / v["peak"]
    return {"steps_per_s": R["steps_per_s"], "ms_per_step": R["ms_per_step"], "median_ms_per_step_stats_on": R["median_ms"],
            "graph": R["graph"],
            "gather_builds_per_force_eval": R["gather_builds_per_force_eval"],
            "kernels": {k: {kk: vv for kk, vv in v.items() if kk != "per_launch"} for k, v in roof.items()},
            "per_launch": {k: v["per_launch"] for k, v in roof.items()},
            "stages_ms": stage_dict(R["stages_timed"]), "stages_ms_stats_on": stage_dict(R["stages"])}, roof


def targets_block(mb, R0, theta, workload):
    """Pass/fail of the precision form's targets on this run: >= 2x the FP64
    steps/s, <= 1e-5 relative acceleration error against FP64, <= 1e-6
    relative total-energy drift against FP64 over 100 steps."""
    a = mb["err_vs_fp64"]["after_last_step"]["acc"]
    du = mb["err_vs_fp64"]["after_last_step"]["du_dt"]
    sp = mb["speedup_vs_fp64"]
    nsteps = R0["fp64"]["steps_total"]
    drift = mb["energy_drift_vs_fp64"]
    return {
        "workload": workload, "theta": theta,
        "speedup_ge_2": {"value": sp, "pass": sp >= 2.0},
        "acc_rel_err_le_1e-5": {"max": a["max"], "p99": a["p99"], "norm": a["norm"],
                                "max_on_force_scale": a.get("max_on_force_scale"),
                                "pass_max": a["max"] <= 1e-5, "pass_p99": a["p99"] <= 1e-5,
                                "pass_norm": a["norm"] <= 1e-5},
        "du_dt_rel_err_on_term_scale": du.get("on_term_scale"),
        # the target is stated over 100 steps: a shorter run reports its drift
        # without a verdict
        "energy_drift_le_1e-6": {"value": drift, "steps": nsteps,
                                 "pass": (drift <= 1e-6) if nsteps >= 100 else None},
    }


# ------------------------------------------------------------------ CPU baseline
def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ic, tables, theta, steps, nsys=1, warm=0):
    """The oracle on the same workload: the initial force evaluation and `warm`
    untimed steps, then `steps` timed steps."""
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
    if "sys" in ic:
        s.sys[:] = ic["sys"]
    tabs = [O.Table(cells=E.table(name, tables[name]), kind="fast") for name in ("granite", "iron", "dunite")]
    sim = O.Sim(s, tabs, O.params(theta=theta, nsys=nsys), kind="fast")
    t0 = time.perf_counter()
    sim.init()
    for _ in range(warm):
        sim.step()
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    t_step = (time.perf_counter() - t0) / max(steps, 1)
    model = cpu_model()
    log(f"[cpu oracle] {O.num_threads()} threads on {model}: init + {warm} warm-up step(s) {t_init:.2f} s, "
        f"{steps} timed step(s) {t_step:.3f} s/step")
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": O.num_threads(), "kind": "port", "cpu_model": model,
            "sample": f"{steps} full step(s) of the same {n:,}-particle workload, timed after the initial force "
                      f"evaluation and {warm} warm-up step(s) ({t_step:.3f} s/step); CPU oracle = the reference's "
                      "functors restated in C, -O3 -ffast-math -funroll-loops -fopenmp, OpenMP over i-groups"}


def cpu_force_eval(ic, tables, theta):
    """The oracle's initial force evaluation (4 density passes, EoS, derivative,
    hydro, gravity: a step without its O(N) kick/drift) -- the bounded CPU
    sample for workloads whose full CPU step would take minutes."""
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
    tabs = [O.Table(cells=E.table(name, tables[name]), kind="fast") for name in ("granite", "iron", "dunite")]
    sim = O.Sim(s, tabs, O.params(theta=theta), kind="fast")
    t0 = time.perf_counter()
    sim.init()
    t = time.perf_counter() - t0
    log(f"[cpu oracle] {O.num_threads()} threads: one force evaluation {t:.2f} s")
    return {"value": 1.0 / t, "unit": "steps/s", "cores": O.num_threads(), "kind": "port", "cpu_model": cpu_model(),
            "sample": f"one force evaluation (the initial one: 4 density passes, EoS, derivative, hydro, gravity) "
                      f"of the same {n:,}-particle workload ({t:.2f} s), timed as one step"}


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain copy kernel moving `nbytes` in and `nbytes` out per launch."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads nbytes, writes nbytes
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": nbytes, "bytes_written": nbytes, "launches": reps}}))


def functor_pass(precisions, tables, ic, warmup, steps, theta):
    """Density pass through sphgpu_dens_dispatch / _retrieve, as FDPS's
    calcForceAllAndWriteBackMultiWalk drives it (src/main.cpp:139, multiwalk
    form): the walks are the device driver's own exact gather lists of the
    initial state, copied into host arrays of the reference's EPI/EPJ layout
    before the clock starts."""
    import ctypes as C
    from fdps_sph_amd import Context, Simulation
    from fdps_sph_amd import structs as S
    from fdps_sph_amd._lib import check, lib
    out = {}
    for prec in precisions:
        ctx = Context(0, prec)
        load_tables(ctx, tables)
        sim = Simulation(ctx, ic, theta=theta)
        pos, m, h = sim.get("POS"), sim.get("MASS"), sim.get("SMTH")
        L = sim.lists(0)
        gs, joff, jidx = L["group_start"], L["j_off"], L["j_idx"]
        G = len(gs) - 1
        # the library's groups and lists, as host arrays of the boundary layout
        epi = np.zeros(len(m), S.EPI_DENS)
        epi["pos"], epi["mass"], epi["smth"], epi["id"] = pos, m, h, np.arange(len(m))
        epj = np.zeros(len(jidx), S.EPJ_DENS)
        epj["mass"], epj["pos"], epj["smth"] = m[jidx], pos[jidx], h[jidx]
        res = np.zeros(len(m), S.RES_DENS)
        ni = np.diff(gs).astype(np.int32)
        nj = np.diff(joff).astype(np.int32)
        pi = (C.c_void_p * G)(*[epi.ctypes.data + int(gs[g]) * epi.itemsize for g in range(G)])
        pj = (C.c_void_p * G)(*[epj.ctypes.data + int(joff[g]) * epj.itemsize for g in range(G)])
        pr = (C.c_void_p * G)(*[res.ctypes.data + int(gs[g]) * res.itemsize for g in range(G)])
        LB = lib()
        W = 200  # walks per call (FDPS n_walk_limit)
        batches = [(b, min(G, b + W)) for b in range(0, G, W)]
        sz = C.sizeof(C.c_void_p)

        def at(arr, k):  # pointer to element k of a pointer array
            return C.cast(C.addressof(arr) + k * sz, C.POINTER(C.c_void_p))

        def one_pass():
            pend = None
            for k, (b, e) in enumerate(batches):
                tag = k & 1
                check(LB.sphgpu_dens_dispatch(ctx.handle, tag, e - b, at(pi, b),
                                              ni[b:e].ctypes.data_as(C.POINTER(C.c_int32)), at(pj, b),
                                              nj[b:e].ctypes.data_as(C.POINTER(C.c_int32))))
                if pend is not None:
                    pt, pb, pe = pend
                    check(LB.sphgpu_dens_retrieve(ctx.handle, pt, pe - pb, ni[pb:pe].ctypes.data_as(
                        C.POINTER(C.c_int32)), at(pr, pb)))
                pend = (tag, b, e)
            pt, pb, pe = pend
            check(LB.sphgpu_dens_retrieve(ctx.handle, pt, pe - pb, ni[pb:pe].ctypes.data_as(C.POINTER(C.c_int32)),
                                          at(pr, pb)))
        for _ in range(warmup):
            one_pass()
        t0 = time.perf_counter()
        for _ in range(steps):
            one_pass()
        sec = (time.perf_counter() - t0) / steps
        host_bytes = epi.nbytes + epj.nbytes + res.nbytes
        log(f"[functor {prec}] density pass over {G:,} walks, {len(jidx):,} j entries: {1e3 * sec:.2f} ms "
            f"({host_bytes / sec / 1e9:.1f} GB/s of boundary arrays)")
        out[prec] = {"ms_per_pass": 1e3 * sec, "walks": G, "j_entries": int(len(jidx)), "boundary_bytes": host_bytes,
                     "boundary_GBps": host_bytes / sec / 1e9, "walks_per_call": W,
                     "results_finite": bool(np.all(np.isfinite(res["dens"])) and np.all(res["dens"] >= 5.0))}
        ctx.close()
    return out


# ------------------------------------------------------------------ main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=list(DEFAULTS), default="impact1m")
    ap.add_argument("--n", type=int, default=None, help="total particles (impact1m / impact16m)")
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--persistent", type=int, default=0, metavar="MASK",
                    help="interaction kernels run as one persistent launch: bitmask 1 density, 2 derivative, "
                         "4 hydro, 8 gravity (batch64 picks it by autotuning)")
    ap.add_argument("--graph", type=int, choices=[0, 1], default=1,
                    help="1 (default): every step after the first replays a recorded HIP graph of the whole step; "
                         "0: steps launched kernel by kernel")
    ap.add_argument("--vsig-split", type=int, choices=[0, 1, 2, 3], default=1, metavar="MASK",
                    help="precision forms (1 FP64, 2 mixed) whose hydro dt runs as a kernel of its own")
    ap.add_argument("--hydro-compact", type=int, choices=[0, 1], default=1,
                    help="1 (default): where dt runs in a kernel of its own, the hydro force terms read the "
                         "symmetric lists compacted to the entries within reach of either support; 0: whole lists")
    ap.add_argument("--walk-waves", type=int, default=0,
                    help="persistent waves of each interaction-list walk (0: the library's default, 4096)")
    ap.add_argument("--eos-density-fp32", action="store_true",
                    help="precision form: run the last density pass (the EoS input) in FP32 too")
    ap.add_argument("--sensitivity", type=float, default=0.0, metavar="EPS",
                    help="also run FP64 from positions perturbed by a seeded relative EPS (e.g. 1e-7, the FP32 "
                         "rounding level) and report its energy difference to the unperturbed FP64 run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB elementwise copy kernel 5 times (the known-size reference "
                         "for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding of this "
              "path is a design, not a measured configuration", file=sys.stderr)
        sys.exit(2)
    warmup, steps = DEFAULTS[args.workload]
    warmup = warmup if args.warmup is None else args.warmup
    steps = steps if args.steps is None else args.steps
    if steps < 1:
        raise SystemExit("--steps must be >= 1")

    from fdps_sph_amd import Context
    ic, desc, tables = make_workload(args.workload, args.n)
    if args.workload == "functor1m":
        precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
        fr = functor_pass(precs, tables, ic, warmup, steps, args.theta)
        main_p = precs[0]
        out = {"metric": METRIC, "value": None, "unit": "steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
               "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f64" if main_p == "fp64" else "f32", "data": "synthetic (seeded lattice bodies)",
               "config": {"workload": f"functor1m: density pass of {desc} through the multiwalk boundary"},
               "roofline": None, "functor_density_pass": fr}
        print(json.dumps(out), flush=True)
        return
    n = len(ic["mass"])
    nsys = int(ic["sys"].max()) + 1 if "sys" in ic else 1
    precs = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    C = Context(0, precs[0])
    load_tables(C, tables)

    params = {"theta": args.theta, "nsys": nsys, "eos_density_fp64": 0 if args.eos_density_fp32 else 1,
              "persistent": args.persistent, "graph": args.graph, "walk_waves": args.walk_waves,
              "vsig_split": args.vsig_split, "hydro_compact": args.hydro_compact}
    extra = {}
    if args.workload == "batch64":
        from fdps_sph_amd import autotune
        tprec = "mixed" if "mixed" in precs else "fp64"
        C.set_precision(tprec)
        tr = autotune.tune(C, ic, warmup=1, steps=3, log=log, **params)
        params.update(tr.params())
        extra["autotune"] = {"precision": tprec,
                             "best": dict(tr.params(), ms_per_step=tr.ms_per_step), "table": tr.table}

    sweep = [0.3, 0.5, 0.7] if args.workload == "theta-sweep" else [args.theta]
    results = {}
    for th in sweep:
        params["theta"] = th
        if len(precs) == 2:
            F, M, err, drift = compare_precisions(C, ic, warmup, steps, **params)
            results[th] = {"fp64": F, "mixed": M, "err": err, "drift": drift}
        else:
            R = run(C, ic, warmup, steps, precs[0], precs[0], **params)
            results[th] = {precs[0]: R}

    th0 = 0.5 if 0.5 in results else sweep[0]
    R0 = results[th0]
    main_prec = "fp64" if "fp64" in R0 else "mixed"
    R = R0[main_prec]
    block, roof = precision_block(R, main_prec == "mixed", args.workload)
    dominant = max(roof, key=lambda k: roof[k]["avg_ms"] * roof[k]["launches"])
    rl = roof[dominant]
    roofline = {"bound": rl["bound"], "achieved": rl["achieved"], "peak": rl["peak"], "unit": rl["unit"],
                "frac": rl["frac"], "traffic": rl["traffic"], "kernel": kernel_symbol(dominant, main_prec == "mixed"),
                "avg_ms": rl["avg_ms"], "per_launch": rl["per_launch"], "valu_issue_frac": rl["valu_issue_frac"]}
    for k in ("counter_tflops", "counter_frac"):
        if k in rl:
            roofline[k] = rl[k]
    if rl["bound"] == "valu":  # gravity: FMAs and a reciprocal square root per pair on the vector ALU, no MFMA
        roofline["note"] = ("vector-ALU bound: peak = MI355X's published " + ("FP32" if main_prec == "mixed" else "FP64")
                            + " vector rate; achieved = 38 flops per interaction x interactions / kernel time")
    out = {
        "metric": METRIC,
        "value": R["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": R["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32",
        "data": "synthetic (seeded lattice bodies)",
        "config": {"workload": f"{args.workload}: {desc}", "particles": n, "systems": nsys, "theta": th0,
                   "precision": main_prec, "n_group": params.get("n_group", 64),
                   "n_leaf": params.get("n_leaf", 16), "persistent": params.get("persistent", 0),
                   "list_front_lds": params.get("list_front_lds", 512), "graph": args.graph,
                   "walk_waves": params.get("walk_waves", 0) or 4096, "vsig_split": params.get("vsig_split", 1),
                   "hydro_compact": params.get("hydro_compact", 1),
                   "eos_density": "fp32" if args.eos_density_fp32 else "fp64"},
        "roofline": roofline,
        main_prec: block,
    }
    if nsys > 1:
        out["particle_steps_per_s"] = R["steps_per_s"] * n
    if "mixed" in R0 and "fp64" in R0:
        mb, _ = precision_block(R0["mixed"], True, args.workload)
        mb["dtype"] = "f32 pair arithmetic relative to the group centre, Kahan-summed; f64 state, EoS and integrator"
        mb["speedup_vs_fp64"] = R0["mixed"]["steps_per_s"] / R0["fp64"]["steps_per_s"]
        mb["err_vs_fp64"] = R0["err"]
        mb["energy_drift_vs_fp64"] = R0["drift"]
        mb["energy_drift_steps"] = R0["fp64"]["steps_total"]
        mb["energy_fp64_rel_change"] = abs(R0["fp64"]["E1"] - R0["fp64"]["E0"]) / abs(R0["fp64"]["E0"])
        F0, M0 = R0["fp64"], R0["mixed"]
        mb["energy_diagnostics"] = {
            "initial_offset": abs(M0["E0"] - F0["E0"]) / abs(F0["E0"]),
            "change_difference": abs((M0["E1"] - M0["E0"]) - (F0["E1"] - F0["E0"])) / abs(F0["E0"]),
            "fp64_parts_initial": F0["parts0"], "fp64_parts_final": F0["parts1"],
            "mixed_minus_fp64_parts_initial": {k: M0["parts0"][k] - F0["parts0"][k] for k in F0["parts0"]},
            "mixed_minus_fp64_parts_final": {k: M0["parts1"][k] - F0["parts1"][k] for k in F0["parts1"]},
            "warmup_dt_rel_diff_max": max([abs(a - b) / b for a, b in zip(M0["warmup_dts"], F0["warmup_dts"])],
                                          default=0.0),
            "final_time": {"fp64": F0["t"], "mixed": M0["t"]}}
        if nsys > 1:
            mb["particle_steps_per_s"] = R0["mixed"]["steps_per_s"] * n
        out["mixed"] = mb
        out["mixed_steps_per_s"] = mb["steps_per_s"]
        out["speedup_mixed_vs_fp64"] = mb["speedup_vs_fp64"]
        out["targets"] = targets_block(mb, R0, th0, args.workload)
    if len(sweep) > 1:
        out["theta_sweep"] = []
        for th in sweep:
            r = results[th]
            e = {"theta": th, "fp64_steps_per_s": r["fp64"]["steps_per_s"],
                 "mixed_steps_per_s": r["mixed"]["steps_per_s"],
                 "speedup": r["mixed"]["steps_per_s"] / r["fp64"]["steps_per_s"],
                 "energy_drift_vs_fp64": r["drift"], "steps": r["fp64"]["steps_total"],
                 "energy_drift_le_1e-6": r["drift"] <= 1e-6,
                 "err_vs_fp64": r["err"]}
            out["theta_sweep"].append(e)
    if args.sensitivity > 0 and "fp64" in R0:
        from fdps_sph_amd.ic import SplitMix64
        pic = dict(ic)
        rnd = SplitMix64(4242).uniform(3 * n).reshape(n, 3) * 2.0 - 1.0
        pic["pos"] = ic["pos"] * (1.0 + args.sensitivity * rnd)
        sens = []
        for th in sweep:
            params["theta"] = th
            P = run(C, pic, warmup, steps, f"fp64 theta {th}, positions perturbed by {args.sensitivity:g}", "fp64",
                    **params)
            F0 = results[th]["fp64"]
            e = {"theta": th, "perturbation": args.sensitivity,
                 "energy_rel_diff": abs(P["E1"] - F0["E1"]) / abs(F0["E1"]),
                 "initial_offset": abs(P["E0"] - F0["E0"]) / abs(F0["E0"]),
                 "parts_diff_final": {k: P["parts1"][k] - F0["parts1"][k] for k in F0["parts1"]},
                 "steps": F0["steps_total"]}
            sens.append(e)
            log(f"[fp64 sensitivity] theta {th}, perturbation {args.sensitivity:g}: energy difference after "
                f"{F0['steps_total']} steps {e['energy_rel_diff']:.2e}")
        out["fp64_sensitivity"] = sens if len(sens) > 1 else sens[0]
        params["theta"] = th0
    out.update(extra)
    # every timed step a replay of a recorded graph (the warm-up steps record
    # them), except the direct step that precedes each recording again after a
    # capacity grew (captures_timed counts those recordings)
    g = R["graph"]
    out["graph_captured"] = bool(args.graph) and g["replayed_steps_timed"] > 0 and \
        g["replayed_steps_timed"] + g["captures_timed"] >= steps

    if not args.no_cpu_baseline:
        if args.workload == "granite10k":
            out["cpu_baseline"] = cpu_baseline(ic, tables, th0, warmup + steps, nsys)
        elif args.workload == "impact16m":
            out["cpu_baseline"] = cpu_force_eval(ic, tables, th0)
        else:  # 1M, batch: the initial force evaluation, one warm-up step, three timed steps
            out["cpu_baseline"] = cpu_baseline(ic, tables, th0, 3, nsys, warm=1)
    print(json.dumps(out), flush=True)
    C.close()


if __name__ == "__main__":
    main()
