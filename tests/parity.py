"""Shared machinery of the GPU parity tests: the oracle fed with the GPU's own
state, per-particle errors of every output of a step, and the step-by-step
runner.

Every comparison starts both sides from the same bits: the initial force
evaluation from the same initial conditions, and every later step from the
GPU's state at the end of the previous one (all particle fields, each
system's dt and time), loaded into the oracle.  Positions and velocities of
the force evaluation are then identical on both sides, so sums whose terms
cancel are measured on the size of their summed terms with no allowance for
differences of the velocities."""
from __future__ import annotations

import numpy as np

from . import helpers as H

TOL = 1.0e-12


def _print(*a):
    print(*a, flush=True)


def oracle_state(oracle, st):
    """An oracle state holding every field of a GPU state (sim.state())."""
    n = len(st["mass"])
    s = oracle.State(n)
    s.x[:], s.y[:], s.z[:] = st["pos"].T
    s.vx[:], s.vy[:], s.vz[:] = st["vel"].T
    s.ax[:], s.ay[:], s.az[:] = st["acc"].T
    s.gx[:], s.gy[:], s.gz[:] = st["grav"].T
    s.vhx[:], s.vhy[:], s.vhz[:] = st["velhalf"].T
    s.rotx[:], s.roty[:], s.rotz[:] = st["rotv"].T
    for a, b in [("pot", "pot"), ("m", "mass"), ("h", "smth"), ("rho", "dens"), ("u", "eng"), ("p", "pres"),
                 ("cs", "snds"), ("divv", "divv"), ("gs", "gradh"), ("bal", "bal"), ("du", "engdot"),
                 ("dti", "dt"), ("uh", "enghalf")]:
        getattr(s, a)[:] = st[b]
    s.mat[:] = st["mat"]
    s.sys[:] = st["sys"]
    s.ngb[:] = st["ngb"]
    s.id[:] = st["id"]
    return s


def ic_state(oracle, ic):
    n = len(ic["mass"])
    s = oracle.State(n)
    s.x[:], s.y[:], s.z[:] = ic["pos"].T
    s.vx[:], s.vy[:], s.vz[:] = ic["vel"].T
    s.m[:] = ic["mass"]
    s.rho[:] = ic["dens"]
    s.u[:] = ic["eng"]
    s.mat[:] = ic["mat"]
    if "sys" in ic:
        s.sys[:] = ic["sys"]
    if "id" in ic:
        s.id[:] = ic["id"]
    return s


def pres_scale(tables, rho, u, mat, rel=1e-7):
    """|rho dP/drho| + |u dP/du| of the EoS tables at each particle's (rho, u)
    (central differences of the oracle's bilinear lookup): how far P moves
    per unit of relative change in its inputs."""
    out = np.zeros(len(rho))
    for k in range(len(rho)):
        t, r, e = tables[int(mat[k])], float(rho[k]), float(u[k])
        dr = (t.pressure(r * (1 + rel), e) - t.pressure(r * (1 - rel), e)) / (2 * rel)
        de = (t.pressure(r, e * (1 + rel)) - t.pressure(r, e * (1 - rel))) / (2 * rel) if e != 0 else 0.0
        out[k] = abs(dr) + abs(de)
    return out


def _rel(x, ref, scale=None):
    """Per-particle relative error (max over vector components) and the
    matching numerator/denominator for a norm-wise figure."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    num = np.abs(x - ref)
    den = np.abs(ref)
    if scale is not None:
        s = np.asarray(scale, dtype=np.float64)
        den = np.maximum(den, s[:, None] if (den.ndim == 2 and s.ndim == 1) else s)
    if num.ndim == 2:
        num, den = num.max(1), den.max(1)
    e = np.divide(num, den, out=np.zeros_like(num), where=num > 0)
    e = np.where((num > 0) & (den == 0), np.inf, e)
    return e, num, den


def field_errors(st, os_, tables, prm):
    """Per-particle errors of every output of a force evaluation and step.
    Sums whose terms cancel (acc, du/dt, gravity, div v, rot v) are measured
    against the size of their summed terms where that exceeds the value;
    pressure on the table's own sensitivity to its inputs and on the bulk
    modulus rho c^2 (its inputs rho and u carry rounding-level differences)."""
    from oracle import oracle as O
    assert np.array_equal(st["id"], os_.id), "particle order differs"
    sc = O.state_term_scales(os_, prm)
    acc = np.stack([os_.ax, os_.ay, os_.az], 1)
    grav = np.stack([os_.gx, os_.gy, os_.gz], 1)
    psc = np.maximum(os_.rho * os_.cs ** 2, pres_scale(tables, os_.rho, os_.u, os_.mat))
    return {
        "dens": _rel(st["dens"], os_.rho),
        "smth": _rel(st["smth"], os_.h),
        "acc": _rel(st["acc"], acc, sc["acc"]),
        "grav": _rel(st["grav"], grav, sc["grav"]),
        "du": _rel(st["engdot"], os_.du, sc["du"]),
        "divv": _rel(st["divv"], os_.divv, sc["div"]),
        "rotv": _rel(st["rotv"], np.stack([os_.rotx, os_.roty, os_.rotz], 1), sc["rot"]),
        "gradh": _rel(st["gradh"], os_.gs),
        "pot": _rel(st["pot"], os_.pot),
        "dt": _rel(st["dt"], os_.dti),
        "pos": _rel(st["pos"], np.stack([os_.x, os_.y, os_.z], 1)),
        # v = v_half + dt/2 (a + g), u = u_half + dt/2 du: their errors are the
        # forces' errors times dt/2 (dt_i bounds the global dt)
        "vel": _rel(st["vel"], np.stack([os_.vx, os_.vy, os_.vz], 1), os_.dti * (sc["acc"] + sc["grav"])),
        "eng": _rel(st["eng"], os_.u, os_.dti * sc["du"]),
        "pres": _rel(st["pres"], os_.p, psc),
    }


def stats(errs):
    out = {}
    for k, (e, num, den) in errs.items():
        nd = np.linalg.norm(den)
        out[k] = {"max": float(e.max()) if e.size else 0.0,
                  "p99": float(np.percentile(e, 99)) if e.size else 0.0,
                  "norm": float(np.linalg.norm(num) / nd) if nd > 0 else 0.0}
    return out


def compare(st, os_, tables, prm, tol=TOL, exact_ngb=True):
    """FP64 bar: every output within tol per particle, neighbour counts equal."""
    if exact_ngb:
        assert np.array_equal(st["ngb"], os_.ngb), "neighbour counts differ"
    errs = field_errors(st, os_, tables, prm)
    worst = {k: float(v[0].max()) if v[0].size else 0.0 for k, v in errs.items()}
    _print("   max rel err:", {k: f"{v:.2e}" for k, v in worst.items()})
    for k, v in worst.items():
        assert v <= tol, (k, v)
    return worst


def load_gpu_state(oracle, sim, tables, prm):
    """An oracle simulation started from the GPU's current state."""
    os_ = oracle_state(oracle, sim.state())
    osim = oracle.Sim(os_, tables, prm)
    dt, t = sim.time()
    osim.dt[:] = dt
    osim.time[:] = t
    return os_, osim


def run_per_step(sim, oracle, ic, tables, prm, nsteps, check, dt_tol=TOL):
    """Initial force evaluation from the same initial conditions, then nsteps
    steps, each from the GPU's state loaded into the oracle.  check(st, os_,
    label) judges the particle fields of each; every system's dt (and, after
    a step, its time) must agree within dt_tol."""
    os_ = ic_state(oracle, ic)
    osim = oracle.Sim(os_, tables, prm)
    osim.init()
    check(sim.state(), os_, "init")
    dt, _ = sim.time()
    assert H.rel_err(dt, osim.dt) <= dt_tol, (dt, osim.dt)
    for k in range(nsteps):
        os_, osim = load_gpu_state(oracle, sim, tables, prm)
        sim.step(1)
        osim.step()
        check(sim.state(), os_, f"step {k + 1}")
        dt, t = sim.time()
        _print(f"   dt gpu {dt.tolist()} oracle {osim.dt.tolist()}")
        assert H.rel_err(dt, osim.dt) <= dt_tol, (dt, osim.dt)
        assert H.rel_err(t, osim.time) <= TOL, (t, osim.time)


def fp64_check(tables, prm, tol=TOL):
    def check(st, os_, label):
        _print(f"  [{label}]")
        compare(st, os_, tables, prm, tol)
    return check


def mixed_check(tables, prm, tol=1e-5, dens_tol=1e-6, log=None):
    """The precision form's bar against the FP64 oracle on the same inputs:
    density and smoothing length per particle within dens_tol; accelerations
    (hydro + external and gravity), du/dt and the velocity derivatives within
    tol norm-wise and at the 99th percentile (the per-particle maximum is
    reported: it sits at particles whose hydro and gravity terms cancel)."""
    def check(st, os_, label):
        s = stats(field_errors(st, os_, tables, prm))
        _print(f"  [{label}]", {k: {kk: f"{vv:.1e}" for kk, vv in v.items()} for k, v in s.items()
                                if k in ("dens", "acc", "grav", "du", "divv", "pres", "dt")})
        if log is not None:
            log.append((label, s))
        assert s["dens"]["max"] <= dens_tol and s["smth"]["max"] <= dens_tol, (s["dens"], s["smth"])
        for k in ("acc", "grav", "du", "divv", "rotv", "pot"):
            assert s[k]["norm"] <= tol and s[k]["p99"] <= tol, (k, s[k])
    return check
