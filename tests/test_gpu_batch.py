"""Batching form: independent impacts of uneven particle counts in one
device-resident simulation (one launch per stage over the groups of every
system), against the oracle's multi-system restatement, in both precision
forms.  Lists never cross systems and each system keeps its own dt and time.
Comparisons run step by step from the GPU's own state (tests/parity.py)."""
import numpy as np
import pytest

from fdps_sph_amd import Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from .parity import fp64_check, mixed_check, run_per_step

pytestmark = pytest.mark.gpu
NSYS = 8


@pytest.fixture(scope="module")
def batch_ic():
    # log-uniform counts in [600, 4000], speeds U[5, 15] km/s, angles U[0, 75] deg
    return IC.batch_of_impacts(n_systems=NSYS, n_min=600, n_max=4000, seed=3)


def test_batch_lists_stay_in_system(ctx64, batch_ic):
    print("[batch lists]", flush=True)
    sim = Simulation(ctx64, batch_ic, nsys=NSYS)
    for kind in (0, 1, 2):
        L = sim.lists(kind)
        sys_ = sim.get("SYS")
        gs = L["group_start"]
        for g in range(len(gs) - 1):
            s = sys_[gs[g]]
            assert np.all(sys_[gs[g]:gs[g + 1]] == s)
            assert np.all(sys_[L["j_idx"][L["j_off"][g]:L["j_off"][g + 1]]] == s)


def test_batch_matches_oracle(ctx64, oracle, batch_ic):
    print(f"[batch vs oracle, {NSYS} systems, {len(batch_ic['mass'])} particles, counts "
          f"{batch_ic['counts'].tolist()}]", flush=True)
    sim = Simulation(ctx64, batch_ic, nsys=NSYS)
    tabs = [oracle.Table(E.PATHS[k]) for k in ("granite", "iron", "dunite")]
    prm = oracle.params(nsys=NSYS)
    run_per_step(sim, oracle, batch_ic, tabs, prm, 2, fp64_check(tabs, prm))
    dt, _ = sim.time()
    assert len(set(dt.tolist())) == NSYS  # independent time steps


def test_batch_precision_form_matches_oracle(ctxmix, oracle, batch_ic):
    """The batching form in the precision form (the 64-impact workload's
    arithmetic) against the FP64 oracle on the same inputs, each step."""
    print(f"[mixed batch vs oracle, {NSYS} systems]", flush=True)
    sim = Simulation(ctxmix, batch_ic, nsys=NSYS)
    tabs = [oracle.Table(E.PATHS[k]) for k in ("granite", "iron", "dunite")]
    prm = oracle.params(nsys=NSYS)
    run_per_step(sim, oracle, batch_ic, tabs, prm, 2, mixed_check(tabs, prm), dt_tol=1e-5)


def test_autotune_picks_a_timed_candidate(ctx64, batch_ic):
    """The autotuner times every (n_group, n_leaf) pair it is given -- groups of
    one wave and of several -- then the launch forms and walk waves at the best
    pair, and returns the fastest row of its table; that configuration runs."""
    from fdps_sph_amd import autotune
    res = autotune.tune(ctx64, batch_ic, n_groups=(64, 256), n_leaves=(8, 16), persistents=(0, 8),
                        front_lds=(512,), wave_counts=(2048, 4096), vsig_splits=(0, 1), warmup=1, steps=1,
                        nsys=NSYS)
    print(f"[autotune] best {res.params()} {res.ms_per_step:.3f} ms; {len(res.table)} candidates", flush=True)
    pairs = {(r["n_group"], r["n_leaf"]) for r in res.table}
    assert pairs == {(64, 8), (64, 16), (256, 8), (256, 16)}
    assert {r["persistent"] for r in res.table} == {0, 8}
    assert {r["walk_waves"] for r in res.table} == {2048, 4096}
    assert {r["vsig_split"] for r in res.table} == {0, 1}
    assert res.ms_per_step == min(r["ms_per_step"] for r in res.table) > 0
    sim = Simulation(ctx64, batch_ic, nsys=NSYS, **res.params())
    sim.step(2)
    assert np.isfinite(sim.get("ACC")).all() and np.all(sim.time()[1] > 0)
