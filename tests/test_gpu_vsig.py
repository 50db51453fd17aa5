"""The hydro pass's dt as a kernel of its own (`sphgpu_sim_params.vsig_split`).

CalcHydroForce (`src/force.h:112-147`) forms dt_i = 0.3 * 2 h_i / v_sig_max,
v_sig_max over every j of the group's list (`:122-124, 145`).  A simulation
step can form that maximum in a kernel of its own on its second stream,
beside the gravity kernel, and leave the hydro kernel the force terms.
Checked on the iron/dunite impact (20k particles, 200 x 200 tables):

* FP64: the split step gives the one-kernel step's results bit for bit, run
  directly and replayed from recorded graphs, and matches the oracle per step
  (every field within 1e-12, dt included);
* the precision form with the split (its dt then formed in FP64): per-step
  parity against the FP64 oracle at the precision form's bar.

With the split, the hydro kernel can read the symmetric lists compacted to
the entries within reach of either support (`hydro_compact`, the default)
while dt runs over the whole lists: the force sums then take the same
nonzero terms in another order, so the step is held to the oracle per step
(both forms), gives the uncompacted step's dt bit for bit at the initial
force evaluation, and replays from recorded graphs bit for bit.
"""
import numpy as np
import pytest

from fdps_sph_amd import Context, Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from .parity import fp64_check, mixed_check, run_per_step

pytestmark = pytest.mark.gpu
NAMES = ("granite", "iron", "dunite")
FIELDS = ("POS", "VEL", "DENS", "SMTH", "ENG", "ACC", "GRAV", "ENGDOT", "DT", "ID", "NGB", "PRES", "SNDS")


@pytest.fixture(scope="module")
def ctxv():
    c = Context(0, "fp64")
    for m, k in enumerate(NAMES):
        c.set_table(m, E.table(k, 200))
    yield c
    c.close()


@pytest.fixture(scope="module")
def tabsv(oracle):
    return [oracle.Table(cells=E.table(k, 200)) for k in NAMES]


@pytest.fixture(scope="module")
def impact():
    return IC.iron_dunite_impact(18000, 2000, seed=1301)


def _run(ctx, ic, steps, **params):
    sim = Simulation(ctx, ic, **params)
    sim.step(steps)
    return {f: sim.get(f) for f in FIELDS}, sim.time()


def test_split_fp64_bit_identical(ctxv, impact):
    ctxv.set_precision("fp64")
    ref, tref = _run(ctxv, impact, 4, vsig_split=0, graph=0)
    for params in ({"vsig_split": 1, "graph": 0}, {"vsig_split": 1, "graph": 1}, {"vsig_split": 3, "graph": 0}):
        params["hydro_compact"] = 0
        print(f"[fp64 {params}]", flush=True)
        got, tgot = _run(ctxv, impact, 4, **params)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f]), (params, f)
        assert np.array_equal(tgot[0], tref[0]) and np.array_equal(tgot[1], tref[1])


@pytest.mark.parametrize("compact", [0, 1])
def test_split_fp64_oracle(ctxv, oracle, tabsv, impact, compact):
    print(f"[iron/dunite 20k fp64 vsig_split=1 hydro_compact={compact}]", flush=True)
    ctxv.set_precision("fp64")
    prm = oracle.params()
    sim = Simulation(ctxv, impact, vsig_split=1, hydro_compact=compact)
    run_per_step(sim, oracle, impact, tabsv, prm, 2, fp64_check(tabsv, prm))


def test_compact_fp64_dt_and_replay(ctxv, impact):
    """Compacted lists: the initial evaluation's dt equals the whole lists' bit
    for bit (dt reads the whole lists either way), and recorded steps replay
    the direct steps bit for bit."""
    ctxv.set_precision("fp64")
    dts = []
    for compact in (0, 1):
        print(f"[fp64 initial evaluation hydro_compact={compact}]", flush=True)
        sim = Simulation(ctxv, impact, vsig_split=1, hydro_compact=compact)
        dts.append((sim.get("DT"), sim.get("ID"), sim.time()))
    assert np.array_equal(dts[0][1], dts[1][1])
    assert np.array_equal(dts[0][0], dts[1][0])
    assert np.array_equal(dts[0][2][0], dts[1][2][0])
    ref, tref = _run(ctxv, impact, 4, vsig_split=1, hydro_compact=1, graph=0)
    print("[fp64 hydro_compact=1 recorded]", flush=True)
    got, tgot = _run(ctxv, impact, 4, vsig_split=1, hydro_compact=1, graph=1)
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f
    assert np.array_equal(tgot[0], tref[0]) and np.array_equal(tgot[1], tref[1])


@pytest.mark.parametrize("compact", [0, 1])
def test_split_precision_form(ctxv, oracle, tabsv, impact, compact):
    print(f"[iron/dunite 20k precision form vsig_split=2 hydro_compact={compact}]", flush=True)
    ctxv.set_precision("mixed")
    try:
        prm = oracle.params()
        sim = Simulation(ctxv, impact, vsig_split=2, hydro_compact=compact)
        run_per_step(sim, oracle, impact, tabsv, prm, 2, mixed_check(tabsv, prm), dt_tol=1e-5)
    finally:
        ctxv.set_precision("fp64")


def test_split_params_checked(ctxv, impact):
    with pytest.raises(Exception):
        Simulation(ctxv, impact, vsig_split=4)
    with pytest.raises(Exception):
        Simulation(ctxv, impact, hydro_compact=2)
