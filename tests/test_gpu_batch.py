"""Batching form: several independent impacts in one device-resident
simulation (one launch per stage over the groups of every system), against the
oracle's multi-system restatement; lists never cross systems and each system
keeps its own dt and time."""
import numpy as np
import pytest

from fdps_sph_amd import Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from . import helpers as H
from .test_gpu_parity import TOL, _compare, _ic_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_ic():
    return IC.batch_of_impacts(n_systems=3, n_min=600, n_max=2500, seed=3)


def test_batch_lists_stay_in_system(ctx64, batch_ic):
    print("[batch lists]", flush=True)
    sim = Simulation(ctx64, batch_ic, nsys=3)
    for kind in (0, 1, 2):
        L = sim.lists(kind)
        sys_ = sim.get("SYS")
        gs = L["group_start"]
        for g in range(len(gs) - 1):
            s = sys_[gs[g]]
            assert np.all(sys_[gs[g]:gs[g + 1]] == s)
            assert np.all(sys_[L["j_idx"][L["j_off"][g]:L["j_off"][g + 1]]] == s)


def test_batch_matches_oracle(ctx64, oracle, batch_ic):
    print("[batch vs oracle, 3 systems]", flush=True)
    sim = Simulation(ctx64, batch_ic, nsys=3)
    os_ = _ic_state(oracle, batch_ic)
    tabs = [oracle.Table(E.PATHS[k]) for k in ("granite", "iron", "dunite")]
    osim = oracle.Sim(os_, tabs, oracle.params(nsys=3))
    osim.init()
    _compare(sim.state(), os_, tabs)
    sim.step(1)
    osim.step()
    dt, t = sim.time()
    print("   dt per system", dt, "oracle", osim.dt, flush=True)
    assert len(set(dt.tolist())) == 3  # independent time steps
    assert H.rel_err(dt, osim.dt) <= TOL
    assert H.rel_err(t, osim.time) <= TOL
    _compare(sim.state(), os_, tabs)
