"""GPU parity with i-groups of more than one 64-lane wave (FDPS's n_group,
`src/main.cpp:88-90`, up to 1024 here): a group of n particles runs as
ceil(n / 64) one-wave work items that share its interaction lists, and the
list walks and the gather-list compaction take the box and radius of the
whole group.  Checked on the iron/dunite impact (20k particles, 200 x 200
tables): FP64 against the oracle built with the same n_group (init + 2 steps,
each step from the GPU's own state), the precision form against the FP64
oracle, and persistent launches and recorded steps against direct ones bit
for bit."""
import numpy as np
import pytest

from fdps_sph_amd import Context, Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from .parity import fp64_check, mixed_check, run_per_step

pytestmark = pytest.mark.gpu
NAMES = ("granite", "iron", "dunite")


def _ctx(precision):
    c = Context(0, precision)
    for m, k in enumerate(NAMES):
        c.set_table(m, E.table(k, 200))
    return c


@pytest.fixture(scope="module")
def ctxg():
    c = _ctx("fp64")
    yield c
    c.close()


@pytest.fixture(scope="module")
def tabsg(oracle):
    return [oracle.Table(cells=E.table(k, 200)) for k in NAMES]


@pytest.fixture(scope="module")
def impact():
    return IC.iron_dunite_impact(18000, 2000, seed=1201)


@pytest.mark.parametrize("n_group", [128, 256, 1024])
def test_multiwave_groups_fp64(ctxg, oracle, tabsg, impact, n_group):
    print(f"[iron/dunite 20k fp64 n_group={n_group}]", flush=True)
    ctxg.set_precision("fp64")
    prm = oracle.params(n_group=n_group)
    sim = Simulation(ctxg, impact, n_group=n_group)
    sizes = np.diff(sim.lists(0)["group_start"])
    print(f"   {len(sizes)} groups, largest {sizes.max()}", flush=True)
    assert sizes.max() > 64
    run_per_step(sim, oracle, impact, tabsg, prm, 2, fp64_check(tabsg, prm))


def test_multiwave_groups_precision_form(ctxg, oracle, tabsg, impact):
    print("[iron/dunite 20k precision form n_group=256]", flush=True)
    ctxg.set_precision("mixed")
    try:
        prm = oracle.params(n_group=256)
        sim = Simulation(ctxg, impact, n_group=256)
        run_per_step(sim, oracle, impact, tabsg, prm, 2, mixed_check(tabsg, prm), dt_tol=1e-5)
    finally:
        ctxg.set_precision("fp64")


_FIELDS = ("POS", "VEL", "DENS", "SMTH", "ENG", "ACC", "GRAV", "ENGDOT", "DT", "ID", "NGB")


def _run(ctx, ic, steps, **params):
    sim = Simulation(ctx, ic, **params)
    sim.step(steps)
    return {f: sim.get(f) for f in _FIELDS}, sim.time()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multiwave_groups_launch_forms(ctxg, impact, precision):
    """n_group 256: every interaction kernel in its persistent form, and steps
    replayed from recorded graphs, give the direct run's results bit for bit."""
    ctxg.set_precision(precision)
    try:
        ref, tref = _run(ctxg, impact, 4, n_group=256, graph=0)
        for params in ({"persistent": 15, "graph": 0}, {"graph": 1}):
            got, tgot = _run(ctxg, impact, 4, n_group=256, **params)
            print(f"[n_group 256 {precision} {params}]", flush=True)
            for f in _FIELDS:
                assert np.array_equal(got[f], ref[f]), (params, f)
            assert np.array_equal(tgot[0], tref[0]) and np.array_equal(tgot[1], tref[1])
    finally:
        ctxg.set_precision("fp64")
