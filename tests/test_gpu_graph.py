"""Recorded steps (sphgpu_sim_params.graph): each step replays a HIP graph of
the whole step -- kick/drift, the sort and tree build with device-side
counts, the lists, every force kernel, the final kick and the time step --
recorded once per parity of the particle buffer sets.  A replayed run must
equal the run launched step by step, bit for bit, in both precision forms,
over a batch of impacts, with persistent launches, and when a replayed
evaluation overflows its list capacities (its final kick and time step then
leave the state untouched and the evaluation runs again directly)."""
import numpy as np
import pytest

from fdps_sph_amd import Simulation
from fdps_sph_amd import ic as IC

pytestmark = pytest.mark.gpu

_FIELDS = ("POS", "VEL", "DENS", "SMTH", "ENG", "ACC", "GRAV", "ENGDOT", "DT", "ID", "NGB", "PRES", "DIVV",
           "VELHALF", "ENGHALF", "BAL", "GRADH", "POT")


def _print(*a):
    print(*a, flush=True)


def _run(ctx, ic, nsteps, shrink_after=None, **params):
    sim = Simulation(ctx, ic, **params)
    if shrink_after is None:
        sim.step(nsteps)
    else:
        sim.step(shrink_after)
        sim.set_capacities(16, 16, 16)
        sim.step(nsteps - shrink_after)
    out = {f: sim.get(f) for f in _FIELDS}
    out["time"] = sim.time()
    return out, sim.list_builds(), sim.graph_stats()


def _same(a, b):
    for f in _FIELDS:
        assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(a["time"][0], b["time"][0]) and np.array_equal(a["time"][1], b["time"][1])


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_graph_steps_match_direct(ctx, request):
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(9000, 1000, seed=1010)
    _print(f"[graph vs direct, {ctx}, 10k, 6 steps]")
    ref, rb, rg = _run(c, ic, 6)
    got, gb, gg = _run(c, ic, 6, graph=1)
    _print(f"  direct {rg}, recorded {gg} (replayed steps, captures); builds {rb} {gb}")
    assert rg == (0, 0)
    assert gg[0] >= 4 and gg[1] >= 2
    assert rb == gb
    _same(ref, got)


def test_graph_batch_persistent(ctx64):
    """Eight impacts of uneven sizes in one simulation, the interaction
    kernels as persistent launches, steps replayed from graphs."""
    ic = IC.batch_of_impacts(n_systems=8, n_min=600, n_max=4000, seed=11)
    _print("[graph batch, 8 systems, persistent launches]")
    ref, rb, _ = _run(ctx64, ic, 4, nsys=8, persistent=15)
    got, gb, gg = _run(ctx64, ic, 4, nsys=8, persistent=15, graph=1)
    _print(f"  recorded {gg}")
    assert gg[0] >= 3
    assert rb == gb
    _same(ref, got)


@pytest.mark.parametrize("ctx", ["ctx64", "ctxmix"])
def test_graph_replay_overflow_redo(ctx, request):
    """Capacities cut to 16 entries between steps: the next step is recorded
    with them and replayed, its evaluation overflows every list, the replay's
    final kick and time step leave the state as the evaluation found it, and
    the evaluation runs again directly with doubled capacities.  The run
    equals the direct run with the same capacity cut, bit for bit."""
    c = request.getfixturevalue(ctx)
    ic = IC.granite_impact(9000, 1000, seed=1011)
    _print(f"[graph replay overflow, {ctx}]")
    ref, rb, _ = _run(c, ic, 5, shrink_after=2)
    got, gb, gg = _run(c, ic, 5, shrink_after=2, graph=1)
    _print(f"  recorded {gg}; builds {rb} {gb}")
    assert gg[0] >= 3 and gg[1] >= 3
    assert rb == gb
    _same(ref, got)
