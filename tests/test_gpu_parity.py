"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same seeded inputs.  Tolerance for the plain FP64 path: 1e-12 relative
(cancellation-prone vector sums are normalised by the size of the summed
terms, see helpers.rel_err); integer outputs (neighbour counts, lists) exact."""
import numpy as np
import pytest

from fdps_sph_amd import Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from . import helpers as H

pytestmark = pytest.mark.gpu
TOL = 1.0e-12


def _print(*a):
    print(*a, flush=True)


# ------------------------------------------------------------------ functors
@pytest.mark.parametrize("seed", [1, 2])
def test_density_functor(ctx64, oracle, seed):
    _print(f"[dens functor seed={seed}]")
    c = H.cloud(700, seed)
    walks = H.walks_from(c, [1, 63, 64, 65, 130, 7], seed, j_frac=0.6)
    wi = [H.epi_dens(c, ii) for ii, _ in walks]
    wj = [H.epj_dens(c, jj) for _, jj in walks]
    wj[5] = wj[5][:0]  # an empty j-list
    res = ctx64.calc_density(wi, wj)
    for a, b, r in zip(wi, wj, res):
        ref = oracle.calc_density(a, b)
        assert H.rel_err(r["dens"], ref["dens"]) <= TOL
        assert H.rel_err(r["smth"], ref["smth"]) <= TOL
    assert np.all(res[5]["dens"] == 5.0)  # FLAG_GI floor with no neighbours


def test_derivative_functor(ctx64, oracle):
    _print("[drvt functor]")
    c = H.cloud(600, 3)
    walks = H.walks_from(c, [64, 100, 3], 3, j_frac=0.7)
    wi = [H.epi_drvt(c, ii) for ii, _ in walks]
    wj = [H.epj_drvt(c, jj) for _, jj in walks]
    res = ctx64.calc_derivative(wi, wj)
    for a, b, r in zip(wi, wj, res):
        ref = oracle.calc_derivative(a, b)
        sd, sr = oracle.drvt_term_scale(a, b)
        assert H.rel_err(r["div_v"], ref["div_v"], sd) <= TOL
        assert H.rel_err(r["rot_v"], ref["rot_v"], sr[:, None]) <= TOL
        assert H.rel_err(r["grad_smth"], ref["grad_smth"]) <= TOL


def test_hydro_functor(ctx64, oracle):
    _print("[hydro functor]")
    c = H.cloud(600, 4)
    walks = H.walks_from(c, [64, 90, 5], 4, j_frac=0.7)
    wi = [H.epi_hydro(c, ii) for ii, _ in walks]
    wj = [H.epj_hydro(c, jj) for _, jj in walks]
    res = ctx64.calc_hydro(wi, wj)
    for a, b, r in zip(wi, wj, res):
        ref = oracle.calc_hydro(a, b)
        sa, su = oracle.hydro_term_scale(a, b)
        assert H.rel_err(r["acc"], ref["acc"], sa[:, None]) <= TOL
        assert H.rel_err(r["eng_dot"], ref["eng_dot"], su) <= TOL
        assert H.rel_err(r["dt"], ref["dt"]) <= TOL


def test_gravity_functor(ctx64, oracle):
    _print("[grav functor]")
    c = H.cloud(500, 5)
    walks = H.walks_from(c, [64, 70, 1], 5, j_frac=0.8)
    rng = np.random.default_rng(5)
    wi = [H.epi_grav(c, ii) for ii, _ in walks]
    wj = [H.epj_grav(c, jj) for _, jj in walks]
    ws = []
    for _ in walks:
        s = np.zeros(40, dtype=oracle.SPJ_MONO)
        s["mass"] = rng.random(40) * 1e22
        s["pos"] = rng.normal(0, 3e7, (40, 3))
        ws.append(s)
    res = ctx64.calc_gravity(wi, wj, ws)
    for a, b, s, r in zip(wi, wj, ws, res):
        ref = oracle.calc_grav(a, b, s)
        sa = np.abs(ref["acc"]).max()
        assert H.rel_err(r["acc"], ref["acc"], sa * 1e-3) <= TOL
        assert H.rel_err(r["pot"], ref["pot"]) <= TOL
        assert np.all(r["dt"] == 1.0e30)
    res2 = ctx64.calc_gravity(wi, wj)  # P2P only
    for a, b, r in zip(wi, wj, res2):
        ref = oracle.calc_grav(a, b)
        assert H.rel_err(r["pot"], ref["pot"]) <= TOL


# ------------------------------------------------------------------ EoS
@pytest.mark.parametrize("name,mat", [("granite", 0), ("iron", 1), ("dunite", 2)])
def test_eos_bilinear(ctx64, oracle, name, mat):
    _print(f"[eos {name}]")
    tab = oracle.Table(E.PATHS[name])
    cells = tab.cells()
    rax, uax = cells[:, 0, 0], cells[0, :, 1]
    rng = np.random.default_rng(mat)
    rho = np.concatenate([np.exp(rng.uniform(np.log(rax[0]) - 1, np.log(rax[-1]) + 1, 3000)), rax, rax[[0, -1, -2]]])
    u = np.concatenate([np.exp(rng.uniform(np.log(uax[0]) - 1, np.log(uax[-1]) + 1, 3000)),
                        rng.choice(uax, len(rax)), uax[[0, -1, -2]]])
    m = np.full(len(rho), mat, dtype=np.int32)
    p, cs = ctx64.pressure(m, rho, u)
    ref_p = np.array([tab.pressure(a, b) for a, b in zip(rho, u)])
    ref_c = np.array([tab.sound_speed(a, b) for a, b in zip(rho, u)])
    assert np.array_equal(p, ref_p)
    assert np.array_equal(cs, ref_c)
    for prop in (2, 5):
        out = ctx64.property(m, rho, u, prop)
        ref = np.array([tab.bilinear(a, b, prop) for a, b in zip(rho, u)])
        assert np.array_equal(out, ref)


def test_eos_internal_energy(ctx64, oracle):
    _print("[eos internal energy]")
    tab = oracle.Table(E.PATHS["granite"])
    rng = np.random.default_rng(9)
    rho = np.exp(rng.uniform(np.log(10.0), np.log(20000.0), 500))
    S = rng.uniform(-2000, 6000, 500)
    m = np.zeros(500, dtype=np.int32)
    for abs_int in (True, False):
        out = ctx64.internal_energy(m, rho, S, abs_int)
        ref = np.array([tab.internal_energy(a, b, abs_int) for a, b in zip(rho, S)])
        assert np.array_equal(np.isnan(out), np.isnan(ref))
        ok = ~np.isnan(ref)
        assert np.array_equal(out[ok], ref[ok])
    one = ctx64.internal_energy(np.zeros(1, np.int32), [2680.0], [3000.0])
    assert abs(one[0] - 1654186.594026882) <= 1e-15 * 1654186.594026882


def test_eos_unknown_material(ctx64):
    from fdps_sph_amd import SphGpuError
    with pytest.raises(SphGpuError) as e:
        ctx64.pressure(np.array([7], np.int32), [1.0], [1.0])
    assert e.value.code == -5


# ------------------------------------------------------------------ lists and tree
def _oracle_state(oracle, st):
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


def _ic_state(oracle, ic):
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
    return s


@pytest.fixture(scope="module")
def small_impact():
    return IC.granite_impact(1800, 200, seed=1000)


def test_lists_match_bruteforce(ctx64, oracle, small_impact):
    _print("[lists gather/symmetric/gravity n=2000]")
    sim = Simulation(ctx64, small_impact)
    for kind in (0, 1):
        L = sim.lists(kind)
        st = sim.state()
        os_ = _oracle_state(oracle, st)
        t = oracle.Tree(os_, oracle.params())
        assert np.array_equal(t.group_start(), L["group_start"])
        if kind == 1:
            t.update_maxH()
        G = t.ngroups
        for g in range(G):
            ref = t.list_sph(g, kind, brute=True)
            got = L["j_idx"][L["j_off"][g]:L["j_off"][g + 1]]
            assert np.array_equal(got, ref), (kind, g)
            assert np.array_equal(t.list_sph(g, kind, brute=False), ref)
    L = sim.lists(2)
    st = sim.state()
    os_ = _oracle_state(oracle, st)
    t = oracle.Tree(os_, oracle.params())
    mom = sim.nodes()
    assert mom.shape[0] == t.nnodes
    for k in range(t.nnodes):
        _, _, m = t.node(k)
        assert np.array_equal(mom[k], m[:10]), k
    for g in range(t.ngroups):
        ej, es = t.list_grav(g, 0.5)
        assert np.array_equal(L["j_idx"][L["j_off"][g]:L["j_off"][g + 1]], ej), g
        assert np.array_equal(L["s_idx"][L["s_off"][g]:L["s_off"][g + 1]], es), g


# ------------------------------------------------------------------ full step
def _pres_scale(tables, rho, u, mat, rel=1e-7):
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


def _compare(st, os_, tables, tol=TOL, prm=None):
    """Per-particle relative errors; sums whose terms cancel (acc, du/dt,
    gravity, div v, rot v) are measured against the size of their summed terms
    (the oracle's per-particle sum of |term|) where that exceeds the value."""
    from oracle import oracle as O
    assert np.array_equal(st["id"], os_.id)
    assert np.array_equal(st["ngb"], os_.ngb)
    sc = O.state_term_scales(os_, prm if prm is not None else O.params(nsys=int(os_.sys.max()) + 1))
    errs = {}
    errs["dens"] = H.rel_err(st["dens"], os_.rho)
    errs["smth"] = H.rel_err(st["smth"], os_.h)
    acc = np.stack([os_.ax, os_.ay, os_.az], 1)
    errs["acc"] = H.rel_err(st["acc"], acc, sc["acc"][:, None])
    grav = np.stack([os_.gx, os_.gy, os_.gz], 1)
    errs["grav"] = H.rel_err(st["grav"], grav, sc["grav"][:, None])
    errs["du"] = H.rel_err(st["engdot"], os_.du, sc["du"])
    errs["divv"] = H.rel_err(st["divv"], os_.divv, sc["div"])
    errs["rotv"] = H.rel_err(st["rotv"], np.stack([os_.rotx, os_.roty, os_.rotz], 1), sc["rot"][:, None])
    errs["gradh"] = H.rel_err(st["gradh"], os_.gs)
    errs["pot"] = H.rel_err(st["pot"], os_.pot)
    errs["dt"] = H.rel_err(st["dt"], os_.dti)
    errs["pos"] = H.rel_err(st["pos"], np.stack([os_.x, os_.y, os_.z], 1), 1e6)
    # P is a table lookup (bit-exact, see test_eos_bilinear) of rho and u, which
    # carry their own rounding-level differences: P is compared on the scale of
    # the bulk modulus rho c^2 and of the table's own sensitivity to its inputs
    # (the bilinear interpolant can be far steeper than c^2 between grid points)
    psc = np.maximum(os_.rho * os_.cs ** 2, _pres_scale(tables, os_.rho, os_.u, os_.mat))
    errs["pres"] = H.rel_err(st["pres"], os_.p, psc)
    _print("   max rel err:", {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert v <= tol, (k, v)


def test_full_step_reference_snapshot(ctx64, oracle):
    """2,000 particles cut from the reference's own input/tar.dat (evolved,
    half-vaporised core and mantle, granite and iron tables): init + 1 step."""
    import os
    from fdps_sph_amd import snapshot as SN
    _print("[full step, reference snapshot subset]")
    ic = SN.to_ic(SN.read_ascii(os.path.join(os.path.dirname(__file__), "golden", "tar_subset_2000.dat")))
    sim = Simulation(ctx64, ic)
    os_ = _ic_state(oracle, ic)
    os_.id[:] = ic["id"]
    tabs = [oracle.Table(E.PATHS[k]) for k in ("granite", "iron", "dunite")]
    osim = oracle.Sim(os_, tabs, oracle.params())
    osim.init()
    _compare(sim.state(), os_, tabs)
    sim.step(1)
    osim.step()
    _compare(sim.state(), os_, tabs)


def test_full_step_granite_10k(ctx64, oracle):
    """10,000-particle granite impact (9,000 + 1,000): init + 2 steps vs oracle."""
    _print("[full step granite 10k]")
    ic = IC.granite_impact(9000, 1000, seed=1000)
    sim = Simulation(ctx64, ic)
    os_ = _ic_state(oracle, ic)
    tabs = [oracle.Table(E.PATHS["granite"]), oracle.Table(E.PATHS["iron"]), oracle.Table(E.PATHS["dunite"])]
    osim = oracle.Sim(os_, tabs, oracle.params(fast_lists=1))
    osim.init()
    _compare(sim.state(), os_, tabs)
    for step in range(2):
        sim.step(1)
        osim.step()
        dt, t = sim.time()
        _print(f"   step {step + 1}: dt gpu {dt[0]!r} oracle {osim.dt[0]!r}")
        assert abs(dt[0] - osim.dt[0]) <= TOL * osim.dt[0]
        _compare(sim.state(), os_, tabs)


def test_gather_list_reuse(ctx64):
    """The density and derivative passes of a force evaluation share one
    gather list while it covers every support; the full-step parity tests
    above run through the same path."""
    ic = IC.granite_impact(9000, 1000, seed=1000)
    sim = Simulation(ctx64, ic)
    sim.step(3)
    builds, evals = sim.list_builds()
    _print(f"[gather lists] {builds} builds for {evals} force evaluations")
    assert evals == 4
    assert evals <= builds <= 2 * evals


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_list_overflow_redo(ctx, request):
    """List capacities far too small for the lists: every force evaluation
    overflows, grows its capacities and runs again from the saved h.  The
    results equal those of a run whose capacities fit from the start, bit for
    bit, and the gather-list constructions are counted once."""
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(9000, 1000, seed=1001)
    big = Simulation(c, ic)
    small = Simulation(c, ic, list_cap_sph=16, list_cap_epj=16, list_cap_spj=16)
    big.step(2)
    small.step(2)
    _print(f"[list overflow redo] builds {big.list_builds()} vs {small.list_builds()}")
    assert big.list_builds() == small.list_builds()
    for f in ("POS", "VEL", "DENS", "SMTH", "ENG", "ACC", "GRAV", "ENGDOT", "DT", "ID"):
        assert np.array_equal(big.get(f), small.get(f)), f


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_frontier_spill_and_overflow(ctx, request):
    """Tree walks keep their frontier in LDS and continue it in global memory.
    With one LDS entry every walk runs through the global part; with a global
    part of 4 entries as well every force evaluation overflows the frontier,
    grows it and runs again.  Both give the default run's results bit for
    bit, in either precision form."""
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(9000, 1000, seed=1002)
    ref = Simulation(c, ic)
    spill = Simulation(c, ic, list_front_lds=1)
    tiny = Simulation(c, ic, list_front_lds=1, list_cap_front=4)
    for sim in (ref, spill, tiny):
        sim.step(2)
    _print(f"[frontier spill] builds {ref.list_builds()} {spill.list_builds()} {tiny.list_builds()}")
    assert ref.list_builds() == spill.list_builds() == tiny.list_builds()
    for f in ("POS", "VEL", "DENS", "SMTH", "ENG", "ACC", "GRAV", "ENGDOT", "DT", "ID", "NGB"):
        a = ref.get(f)
        assert np.array_equal(a, spill.get(f)), f
        assert np.array_equal(a, tiny.get(f)), f


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_second_stream_matches_serial(ctx, request):
    """The gravity and hydro list walks run on a second stream beside the
    density passes and the gravity kernel; with kernel statistics on, the same
    work runs in the same order on one stream.  Both give the same results, bit
    for bit, over steps whose force evaluations rebuild and reuse lists."""
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(20000, 2000, seed=1003)
    over = Simulation(c, ic)
    serial = Simulation(c, ic)
    serial.set_stats(True)
    for sim in (over, serial):
        sim.step(3)
    _print(f"[second stream] builds {over.list_builds()} vs {serial.list_builds()}")
    assert over.list_builds() == serial.list_builds()
    for f in ("POS", "VEL", "DENS", "SMTH", "ENG", "ACC", "GRAV", "ENGDOT", "DT", "ID", "NGB", "PRES", "DIVV"):
        assert np.array_equal(over.get(f), serial.get(f)), f
