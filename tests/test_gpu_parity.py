"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same inputs.  Tolerance for the plain FP64 path: 1e-12 relative (sums whose
terms cancel are normalised by the size of the summed terms); integer outputs
(neighbour counts, lists) exact.  Full steps are compared one at a time, each
from the GPU's own state loaded into the oracle (tests/parity.py)."""
import numpy as np
import pytest

from fdps_sph_amd import Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from . import helpers as H
from .parity import TOL, compare, fp64_check, ic_state, oracle_state, run_per_step

pytestmark = pytest.mark.gpu


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


@pytest.fixture(scope="module")
def ctx_till():
    from fdps_sph_amd import Context
    c = Context(0, "fp64")
    c.set_tillotson(0, E.GRANITE_TILLOTSON)
    c.load_table(1, E.PATHS["iron"])
    yield c
    c.close()


def test_eos_tillotson(ctx_till, oracle):
    """Analytic Tillotson granite (src/EoS.h:91-113, 446) in every regime:
    P within 1e-12 of the size of its terms, c_s within 1e-12."""
    _print("[eos tillotson]")
    rho0, u0, uiv, ucv, A, B, a, b, alpha, beta = E.GRANITE_TILLOTSON
    t = oracle.Table(tillotson=E.GRANITE_TILLOTSON)
    rng = np.random.default_rng(17)
    rho = rho0 * np.exp(rng.uniform(np.log(0.05), np.log(3.0), 4000))
    u = np.exp(rng.uniform(np.log(1e3), np.log(1e8), 4000))
    rho = np.concatenate([rho, [rho0, rho0, 0.8 * rho0, 0.8 * rho0]])
    u = np.concatenate([u, [1e6, 0.0, uiv, ucv]])
    m = np.zeros(len(rho), np.int32)
    p, cs = ctx_till.pressure(m, rho, u)
    ref = np.array([t.pressure(r, e) for r, e in zip(rho, u)])
    ref_c = np.array([t.sound_speed(r, e) for r, e in zip(rho, u)])
    eta = rho / rho0
    mu = eta - 1
    scale = a * rho * u + b * rho * u / (u / u0 / eta ** 2 + 1) + A * np.abs(mu) + B * mu * mu
    assert H.rel_err(p, ref, scale) <= TOL
    assert H.rel_err(cs, ref_c, 1e-12 * ref_c.max()) <= TOL
    assert np.all(np.isnan(ctx_till.property(m, rho, u, 2))) and np.all(np.isnan(ctx_till.property(m, rho, u, 5)))
    assert np.array_equal(ctx_till.property(m, rho, u, 3), p)
    from fdps_sph_amd import SphGpuError
    with pytest.raises(SphGpuError):
        ctx_till.internal_energy(m[:3], rho[:3], u[:3])


def test_full_step_tillotson_granite(ctx_till, oracle):
    """A granite impact (4,500 + 500) on the analytic Tillotson material:
    init + 2 steps, each from the GPU's state."""
    _print("[full step, analytic Tillotson granite 5k]")
    ic = IC.granite_impact(4500, 500, seed=1007)
    tabs = [oracle.Table(tillotson=E.GRANITE_TILLOTSON), oracle.Table(E.PATHS["iron"])]
    prm = oracle.params()
    run_per_step(Simulation(ctx_till, ic), oracle, ic, tabs, prm, 2, fp64_check(tabs, prm))


def test_eos_unknown_material(ctx64):
    from fdps_sph_amd import SphGpuError
    with pytest.raises(SphGpuError) as e:
        ctx64.pressure(np.array([7], np.int32), [1.0], [1.0])
    assert e.value.code == -5


# ------------------------------------------------------------------ lists and tree
@pytest.fixture(scope="module")
def small_impact():
    return IC.granite_impact(1800, 200, seed=1000)


def test_lists_match_bruteforce(ctx64, oracle, small_impact):
    _print("[lists gather/symmetric/gravity n=2000]")
    sim = Simulation(ctx64, small_impact)
    for kind in (0, 1):
        L = sim.lists(kind)
        st = sim.state()
        os_ = oracle_state(oracle, st)
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
    os_ = oracle_state(oracle, st)
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


def test_compacted_gather_lists(ctx64, small_impact):
    """The gather lists a force evaluation uses (export kind 3): built with the
    8 % margin on h, then compacted on the matrix cores to the entries within
    R_g = 1.08 max_i 2.5 h_i of some particle of the group.  Each group's list
    must stay in ascending index order, keep every particle inside some
    member's margin sphere (1.08 * 2.5 h_i, FP64 distances), and hold nothing
    beyond R_g plus the FP32 test's room or outside the search box."""
    _print("[compacted gather lists n=2000]")
    sim = Simulation(ctx64, small_impact)
    L = sim.lists(3)
    st = sim.state()
    x, h = st["pos"], st["smth"]
    gs = L["group_start"]
    kept_total = box_total = 0
    for g in range(len(gs) - 1):
        i = np.arange(gs[g], gs[g + 1])
        got = L["j_idx"][L["j_off"][g]:L["j_off"][g + 1]]
        assert np.all(np.diff(got) > 0), g
        d = np.sqrt(((x[None, :, :] - x[i][:, None, :]) ** 2).sum(axis=2))  # (n_i, n)
        Hm = (2.5 * h[i]) * 1.08  # the margin radius, rounded as the list construction rounds it
        need = np.nonzero(np.any(d <= (Hm * (1 - 1e-9))[:, None], axis=0))[0]
        assert np.all(np.isin(need, got)), g
        R = 1.08 * np.max(2.5 * h[i])
        assert np.all(d[:, got].min(axis=0) <= R * (1 + 1e-4)), g
        lo = (x[i] - Hm[:, None]).min(axis=0)
        hi = (x[i] + Hm[:, None]).max(axis=0)
        inbox = np.all((x >= lo) & (x <= hi), axis=1)
        assert np.all(inbox[got]), g
        kept_total += len(got)
        box_total += int(inbox.sum())
    _print(f"  kept {kept_total} of {box_total} box entries")
    assert kept_total < box_total


def test_compacted_symmetric_lists(ctx64, small_impact):
    """The symmetric lists as the hydro force terms read them (export kind 4):
    each group's list is an ordered subsequence of its symmetric list (kind 1)
    that keeps every particle inside either support of some member
    (|x_i - x_j| < 2.5 max(h_i, h_j), FP64 distances) and nothing beyond
    max(max_i 2.5 h_i, 2.5 h_j) plus the FP32 test's room."""
    _print("[compacted symmetric lists n=2000]")
    sim = Simulation(ctx64, small_impact)
    full, comp = sim.lists(1), sim.lists(4)
    st = sim.state()
    x, h = st["pos"], st["smth"]
    gs = comp["group_start"]
    assert np.array_equal(gs, full["group_start"])
    kept_total = full_total = 0
    for g in range(len(gs) - 1):
        i = np.arange(gs[g], gs[g + 1])
        lst = full["j_idx"][full["j_off"][g]:full["j_off"][g + 1]]
        got = comp["j_idx"][comp["j_off"][g]:comp["j_off"][g + 1]]
        pos = np.searchsorted(lst, got) if np.all(np.diff(lst) > 0) else None
        if pos is not None:  # an ordered subsequence of an ascending list
            assert np.all(np.diff(got) > 0) and np.all(lst[np.minimum(pos, len(lst) - 1)] == got), g
        else:
            it = iter(lst.tolist())
            assert all(any(v == w for w in it) for v in got.tolist()), g
        d = np.sqrt(((x[lst][None, :, :] - x[i][:, None, :]) ** 2).sum(axis=2))  # (n_i, n_list)
        H = 2.5 * np.maximum(h[i][:, None], h[lst][None, :])
        need = lst[np.any(d < H * (1 - 1e-9), axis=0)]
        assert np.all(np.isin(need, got)), g
        reach = np.maximum(2.5 * h[i].max(), 2.5 * h[lst]) * (1 + 2e-4)
        assert np.all(np.isin(got, lst[d.min(axis=0) <= reach])), g
        kept_total += len(got)
        full_total += len(lst)
    _print(f"  kept {kept_total} of {full_total} symmetric-list entries")
    assert kept_total < full_total


# ------------------------------------------------------------------ multiwalk boundary
def test_two_tags_in_flight_and_empty_lists(ctx64, oracle):
    """FDPS double-buffers multiwalk batches (tag_max > 1): batches of
    different kinds are dispatched on tags 0, 1 and 2 before any retrieve,
    then retrieved out of order.  Walks with empty j-lists give the reference
    functors' values: dt = 0.6 h / 0 = inf for hydro, zero force and
    potential for gravity."""
    _print("[two tags in flight, empty j-lists]")
    c = H.cloud(500, 31)
    wh = H.walks_from(c, [64, 30, 5], 31, j_frac=0.5)
    wg = H.walks_from(c, [20, 64, 64], 32, j_frac=0.7)
    wd = H.walks_from(c, [40, 64], 33, j_frac=0.6)
    hi = [H.epi_hydro(c, i) for i, _ in wh]
    hj = [H.epj_hydro(c, j) for _, j in wh]
    hj[2] = hj[2][:0]
    gi = [H.epi_grav(c, i) for i, _ in wg]
    gj = [H.epj_grav(c, j) for _, j in wg]
    gj[0] = gj[0][:0]
    spj = [np.zeros(0, dtype=oracle.SPJ_MONO) for _ in wg]
    di = [H.epi_dens(c, i) for i, _ in wd]
    dj = [H.epj_dens(c, j) for _, j in wd]
    n0 = ctx64.dispatch("hydro", hi, hj, tag=0)
    n1 = ctx64.dispatch("grav", gi, gj, spj, tag=1)
    n2 = ctx64.dispatch("dens", di, dj, tag=2)
    # a retrieve with a null n_epi array is an argument error and leaves the batch pending
    from fdps_sph_amd import _lib
    import ctypes as C
    pr = (C.c_void_p * 3)()
    assert ctx64._L.sphgpu_grav_retrieve(ctx64.handle, 1, 3, None, pr) == _lib.ERR_ARG
    rg = ctx64.retrieve("grav", n1, tag=1)
    rd = ctx64.retrieve("dens", n2, tag=2)
    rh = ctx64.retrieve("hydro", n0, tag=0)
    for a, b, r in zip(hi, hj, rh):
        ref = oracle.calc_hydro(a, b)
        if len(b) == 0:
            assert np.all(np.isinf(r["dt"])) and np.all(np.isinf(ref["dt"]))
            assert np.all(r["acc"] == 0) and np.all(r["eng_dot"] == 0)
            continue
        sa, su = oracle.hydro_term_scale(a, b)
        assert H.rel_err(r["acc"], ref["acc"], sa[:, None]) <= TOL
        assert H.rel_err(r["eng_dot"], ref["eng_dot"], su) <= TOL
        assert H.rel_err(r["dt"], ref["dt"]) <= TOL
    for a, b, r in zip(gi, gj, rg):
        ref = oracle.calc_grav(a, b)
        if len(b) == 0:
            assert np.all(r["acc"] == 0) and np.all(r["pot"] == 0)
            continue
        assert H.rel_err(r["pot"], ref["pot"]) <= TOL
        assert H.rel_err(r["acc"], ref["acc"], np.abs(ref["acc"]).max() * 1e-3) <= TOL
    for a, b, r in zip(di, dj, rd):
        ref = oracle.calc_density(a, b)
        assert H.rel_err(r["dens"], ref["dens"]) <= TOL
    from fdps_sph_amd import SphGpuError
    with pytest.raises(SphGpuError):  # each tag was retrieved once
        ctx64.retrieve("hydro", n0, tag=0)


# ------------------------------------------------------------------ full step
def _tables(oracle, grid=None):
    return [oracle.Table(cells=E.table(k, grid)) for k in ("granite", "iron", "dunite")]


def test_full_step_reference_snapshot(ctx64, oracle):
    """2,000 particles cut from the reference's own input/tar.dat (evolved,
    half-vaporised core and mantle, granite and iron tables): init + 2 steps."""
    import os
    from fdps_sph_amd import snapshot as SN
    _print("[full step, reference snapshot subset]")
    ic = SN.to_ic(SN.read_ascii(os.path.join(os.path.dirname(__file__), "golden", "tar_subset_2000.dat")))
    tabs = _tables(oracle)
    prm = oracle.params()
    run_per_step(Simulation(ctx64, ic), oracle, ic, tabs, prm, 2, fp64_check(tabs, prm))


def test_full_step_granite_10k(ctx64, oracle):
    """The 10,000-particle granite impact (9,000 + 1,000, 9 km/s, 45 deg):
    init + 2 steps."""
    _print("[full step granite 10k]")
    ic = IC.granite_impact(9000, 1000, seed=1000)
    tabs = _tables(oracle)
    prm = oracle.params()
    run_per_step(Simulation(ctx64, ic), oracle, ic, tabs, prm, 2, fp64_check(tabs, prm))


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


_FIELDS = ("POS", "VEL", "DENS", "SMTH", "ENG", "ACC", "GRAV", "ENGDOT", "DT", "ID", "NGB", "PRES", "DIVV")


def _run(ctx, ic, nsteps, stats=False, **params):
    """Run a simulation to completion and keep its outputs (a context holds
    one simulation at a time)."""
    sim = Simulation(ctx, ic, **params)
    if stats:
        sim.set_stats(True)
    sim.step(nsteps)
    return {f: sim.get(f) for f in _FIELDS}, sim.list_builds()


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_list_overflow_redo(ctx, request):
    """List capacities far too small for the lists: every force evaluation
    overflows, grows its capacities and runs again from the saved h.  The
    results equal those of a run whose capacities fit from the start, bit for
    bit, and the gather-list constructions are counted once."""
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(9000, 1000, seed=1001)
    big, bb = _run(c, ic, 2)
    small, sb = _run(c, ic, 2, list_cap_sph=16, list_cap_epj=16, list_cap_spj=16)
    _print(f"[list overflow redo] builds {bb} vs {sb}")
    assert bb == sb
    for f in _FIELDS:
        assert np.array_equal(big[f], small[f]), f


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_frontier_spill_and_overflow(ctx, request):
    """Tree walks keep their frontier in LDS and continue it in global memory.
    With one LDS entry every walk runs through the global part; with a global
    part of 4 entries as well every force evaluation overflows the frontier,
    grows it and runs again.  Both give the default run's results bit for
    bit, in either precision form."""
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(9000, 1000, seed=1002)
    ref, rb = _run(c, ic, 2)
    spill, sb = _run(c, ic, 2, list_front_lds=1)
    tiny, tb = _run(c, ic, 2, list_front_lds=1, list_cap_front=4)
    _print(f"[frontier spill] builds {rb} {sb} {tb}")
    assert rb == sb == tb
    for f in _FIELDS:
        assert np.array_equal(ref[f], spill[f]), f
        assert np.array_equal(ref[f], tiny[f]), f


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_second_stream_matches_serial(ctx, request):
    """The gravity and hydro list walks run on a second stream beside the
    density passes and the gravity kernel; with kernel statistics on, the same
    work runs in the same order on one stream.  Both give the same results, bit
    for bit, over steps whose force evaluations rebuild and reuse lists."""
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(20000, 2000, seed=1003)
    over, ob = _run(c, ic, 3)
    serial, sb = _run(c, ic, 3, stats=True)
    _print(f"[second stream] builds {ob} vs {sb}")
    assert ob == sb
    for f in _FIELDS:
        assert np.array_equal(over[f], serial[f]), f


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_persistent_launches_match(ctx, request):
    """The interaction kernels as persistent launches over per-XCD group
    counters give the one-workgroup-per-group launches' results bit for bit,
    on a batch of several impacts (groups of every system in one launch)."""
    c = request.getfixturevalue(ctx)
    ic = IC.batch_of_impacts(n_systems=6, n_min=600, n_max=4000, seed=9)
    ref, rb = _run(c, ic, 2, nsys=6)
    per, pb = _run(c, ic, 2, nsys=6, persistent=15)
    _print(f"[persistent launches] builds {rb} vs {pb}")
    assert rb == pb
    for f in _FIELDS:
        assert np.array_equal(ref[f], per[f]), f


def test_set_precision_matches_dedicated_context(ctxmix):
    """A context switched to the precision form runs it exactly as a context
    created in that form, and switched back runs FP64 again."""
    from fdps_sph_amd import Context
    ic = IC.granite_impact(9000, 1000, seed=1005)
    ref, _ = _run(ctxmix, ic, 2)
    c = Context(0, "fp64")
    try:
        for m, k in enumerate(("granite", "iron", "dunite")):
            c.load_table(m, E.PATHS[k])
        fp, _ = _run(c, ic, 2)
        c.set_precision("mixed")
        sw, _ = _run(c, ic, 2)
        c.set_precision("fp64")
        back, _ = _run(c, ic, 2)
    finally:
        c.close()
    for f in _FIELDS:
        assert np.array_equal(ref[f], sw[f]), f
        assert np.array_equal(fp[f], back[f]), f
    assert not np.array_equal(fp["ACC"], sw["ACC"])


def test_replaced_simulation_raises(ctx64):
    ic = IC.granite_impact(450, 50, seed=5)
    a = Simulation(ctx64, ic)
    Simulation(ctx64, ic)
    with pytest.raises(RuntimeError):
        a.step(1)


def test_list_export_leaves_the_state(ctx64):
    """sphgpu_sim_lists rebuilds the tree and lists of the current state; a
    single size-only call (no arrays) between steps changes nothing that
    follows, bit for bit."""
    import ctypes as C
    ic = IC.granite_impact(9000, 1000, seed=1004)
    ref, _ = _run(ctx64, ic, 2)
    sim = Simulation(ctx64, ic)
    sim.step(1)
    ng, nj, ns = C.c_int64(), C.c_int64(), C.c_int64()
    from fdps_sph_amd._lib import check
    check(sim._L.sphgpu_sim_lists(sim.ctx.handle, 1, C.byref(ng), C.byref(nj), C.byref(ns), None, None, None, None,
                                  None))
    assert ng.value > 0 and nj.value > 0
    sim.step(1)
    for f in _FIELDS:
        assert np.array_equal(ref[f], sim.get(f)), f
