"""GPU parity on the benchmark workloads' shapes, at sizes the oracle runs in
seconds, with the 200 x 200 EoS tables converted from the ANEOS sources:

* iron core + dunite mantle target and impactor (the 1M workload's
  structure, 20k particles) in FP64 against the oracle at theta 0.3, 0.5 and
  0.7, init + 2 steps, each step from the GPU's own state;
* the three-material impact (iron/dunite target, granite impactor; the 16M
  workload's structure, 20k particles) the same way;
* the precision form on the iron/dunite impact against the FP64 oracle at
  each theta, and its total-energy drift against the FP64 GPU run;
* the full 1M iron/dunite workload: the derivative, hydro and gravity
  results of 100 sampled i-groups against the oracle's functors over the
  GPU's own exported lists, at 1e-12.
"""
import numpy as np
import pytest

from fdps_sph_amd import Context, Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from . import helpers as H
from .parity import TOL, fp64_check, mixed_check, run_per_step

pytestmark = pytest.mark.gpu
NAMES = ("granite", "iron", "dunite")


def _ctx(precision):
    c = Context(0, precision)
    for m, k in enumerate(NAMES):
        c.set_table(m, E.table(k, 200))
    return c


@pytest.fixture(scope="module")
def ctx200():
    c = _ctx("fp64")
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx200mix():
    c = _ctx("mixed")
    yield c
    c.close()


@pytest.fixture(scope="module")
def tabs200(oracle):
    return [oracle.Table(cells=E.table(k, 200)) for k in NAMES]


@pytest.fixture(scope="module")
def iron_dunite():
    return IC.iron_dunite_impact(18000, 2000, seed=1101)


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_iron_dunite_fp64(ctx200, oracle, tabs200, iron_dunite, theta):
    print(f"[iron/dunite 20k fp64 theta={theta}]", flush=True)
    prm = oracle.params(theta=theta)
    sim = Simulation(ctx200, iron_dunite, theta=theta)
    run_per_step(sim, oracle, iron_dunite, tabs200, prm, 2, fp64_check(tabs200, prm))


def test_three_material_fp64(ctx200, oracle, tabs200):
    ic = IC.three_material_impact(15000, 5000, seed=1104)
    print(f"[three materials 20k fp64, materials {np.bincount(ic['mat']).tolist()}]", flush=True)
    assert set(ic["mat"].tolist()) == {0, 1, 2}
    prm = oracle.params()
    run_per_step(Simulation(ctx200, ic), oracle, ic, tabs200, prm, 2, fp64_check(tabs200, prm))


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_iron_dunite_precision_form(ctx200mix, oracle, tabs200, iron_dunite, theta):
    print(f"[iron/dunite 20k precision form vs fp64 oracle, theta={theta}]", flush=True)
    prm = oracle.params(theta=theta)
    sim = Simulation(ctx200mix, iron_dunite, theta=theta)
    run_per_step(sim, oracle, iron_dunite, tabs200, prm, 2, mixed_check(tabs200, prm), dt_tol=1e-5)


def _energy(sim):
    st = sim.state()
    m = st["mass"]
    kin = 0.5 * np.sum(m * np.sum(st["vel"] ** 2, 1))
    return kin + np.sum(m * st["eng"]) + 0.5 * np.sum(m * st["pot"])


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_precision_form_energy_drift(ctx200, ctx200mix, iron_dunite, theta):
    """Total energy (kinetic + internal + gravitational) of the precision
    form against the FP64 run from the same initial state after 100 steps:
    |E_mixed - E_fp64| / |E_fp64| <= 1e-6 (the two start apart by the FP32
    gravity's potential error, which this includes)."""
    nsteps = 100
    out = {}
    for name, c in (("fp64", ctx200), ("mixed", ctx200mix)):
        sim = Simulation(c, iron_dunite, theta=theta)
        e0 = _energy(sim)
        sim.step(nsteps)
        out[name] = (e0, _energy(sim), sim.time()[1][0])
    (a0, a1, ta), (b0, b1, tb) = out["fp64"], out["mixed"]
    drift = abs(b1 - a1) / abs(a1)
    print(f"[energy drift theta={theta}] fp64 {a0:.6e} -> {a1:.6e} (t={ta:.4f}), mixed {b0:.6e} -> {b1:.6e} "
          f"(t={tb:.4f}): initial offset {abs(b0 - a0) / abs(a0):.2e}, after {nsteps} steps {drift:.2e}", flush=True)
    assert drift <= 1e-6


def _epi(kind, st, ii):
    from fdps_sph_amd import structs as S
    if kind == "drvt":
        a = np.zeros(len(ii), S.EPI_DRVT)
        a["pos"], a["vel"], a["smth"], a["dens"] = st["pos"][ii], st["vel"][ii], st["smth"][ii], st["dens"][ii]
    elif kind == "hydro":
        a = np.zeros(len(ii), S.EPI_HYDRO)
        a["pos"], a["vel"], a["smth"], a["dens"] = st["pos"][ii], st["vel"][ii], st["smth"][ii], st["dens"][ii]
        a["pres"], a["snds"], a["grad_smth"], a["Bal"] = st["pres"][ii], st["snds"][ii], st["gradh"][ii], st["bal"][ii]
        a["AVa"], a["id"] = 1.0, ii
    else:
        a = np.zeros(len(ii), S.EPI_GRAV)
        a["pos"], a["eps2"], a["id"] = st["pos"][ii], st["smth"][ii] ** 2, ii
    return a


def _epj(kind, st, jj):
    from fdps_sph_amd import structs as S
    if kind == "drvt":
        a = np.zeros(len(jj), S.EPJ_DRVT)
        a["mass"], a["pos"], a["vel"], a["smth"] = st["mass"][jj], st["pos"][jj], st["vel"][jj], st["smth"][jj]
    elif kind == "hydro":
        a = np.zeros(len(jj), S.EPJ_HYDRO)
        for f, g in (("pos", "pos"), ("vel", "vel"), ("dens", "dens"), ("mass", "mass"), ("smth", "smth"),
                     ("pres", "pres"), ("grad_smth", "gradh"), ("snds", "snds"), ("Bal", "bal")):
            a[f] = st[g][jj]
        a["AVa"], a["id"] = 1.0, jj
    else:
        a = np.zeros(len(jj), S.EPJ_GRAV)
        a["pos"], a["mass"], a["id"] = st["pos"][jj], st["mass"][jj], jj
    return a


def test_1m_sampled_groups_fp64(ctx200, oracle):
    """The 1M iron/dunite workload (900k + 100k, 200 x 200 tables) after its
    initial force evaluation (external force off, so the acceleration is the
    hydro functor's): 100 sampled i-groups' derivative, Balsara, hydro and
    gravity results against the oracle's functors fed with the GPU's state
    and its exported lists, at 1e-12 relative."""
    ic = IC.iron_dunite_impact(900_000, 100_000, seed=1001)
    print("[1M iron/dunite: sampled groups vs oracle functors]", flush=True)
    sim = Simulation(ctx200, ic, ext_force_until=0.0)
    st = sim.state()
    Ld, Lh, Lg = sim.lists(0), sim.lists(1), sim.lists(2)
    nodes = sim.nodes()
    gs = Ld["group_start"]
    G = len(gs) - 1
    assert np.array_equal(gs, Lh["group_start"]) and np.array_equal(gs, Lg["group_start"])
    groups = np.random.default_rng(7).choice(G, 100, replace=False)
    worst = {}

    def upd(k, v):
        worst[k] = max(worst.get(k, 0.0), v)
    G_ = 6.67e-11
    for g in groups:
        ii = np.arange(gs[g], gs[g + 1])
        jj = Ld["j_idx"][Ld["j_off"][g]:Ld["j_off"][g + 1]]
        a, b = _epi("drvt", st, ii), _epj("drvt", st, jj)
        ref = oracle.calc_derivative(a, b)
        sd, sr = oracle.drvt_term_scale(a, b)
        upd("divv", H.rel_err(st["divv"][ii], ref["div_v"], sd))
        upd("rotv", H.rel_err(st["rotv"][ii], ref["rot_v"], sr[:, None]))
        upd("gradh", H.rel_err(st["gradh"][ii], ref["grad_smth"]))
        bal = np.array([oracle.balsara(ref["div_v"][k], ref["rot_v"][k], st["snds"][i], st["smth"][i])
                        for k, i in enumerate(ii)])
        # Bal = |div|/D, D = |div| + |rot| + 1e-4 c/h: it moves by at most
        # (d div + d rot)/D, so it is measured on (sd + sr)/D
        D = np.abs(ref["div_v"]) + np.linalg.norm(ref["rot_v"], axis=1) + 1e-4 * st["snds"][ii] / st["smth"][ii]
        upd("bal", H.rel_err(st["bal"][ii], bal, (sd + sr) / D))
        jj = Lh["j_idx"][Lh["j_off"][g]:Lh["j_off"][g + 1]]
        a, b = _epi("hydro", st, ii), _epj("hydro", st, jj)
        ref = oracle.calc_hydro(a, b)
        sa, su = oracle.hydro_term_scale(a, b)
        upd("acc", H.rel_err(st["acc"][ii], ref["acc"], sa[:, None]))
        upd("du", H.rel_err(st["engdot"][ii], ref["eng_dot"], su))
        upd("dt", H.rel_err(st["dt"][ii], ref["dt"]))
        ej = Lg["j_idx"][Lg["j_off"][g]:Lg["j_off"][g + 1]]
        es = Lg["s_idx"][Lg["s_off"][g]:Lg["s_off"][g + 1]]
        spj = np.zeros(len(es), dtype=oracle.SPJ_MONO)
        spj["mass"] = nodes[es, 0]
        spj["pos"] = nodes[es, 1:4]
        a = _epi("grav", st, ii)
        ref = oracle.calc_grav(a, _epj("grav", st, ej), spj)
        src = np.concatenate([st["pos"][ej], spj["pos"]])
        msrc = np.concatenate([st["mass"][ej], spj["mass"]])
        d2 = ((st["pos"][ii][:, None, :] - src[None]) ** 2).sum(-1) + a["eps2"][:, None]
        sg = (G_ * msrc[None] / d2).sum(1)
        upd("grav", H.rel_err(st["grav"][ii], ref["acc"], sg[:, None]))
        upd("pot", H.rel_err(st["pot"][ii], ref["pot"]))
    print("   max rel err:", {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    for k, v in worst.items():
        assert v <= TOL, (k, v)
