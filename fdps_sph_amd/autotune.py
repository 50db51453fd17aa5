"""Pick the tree/group parameters of a device-resident simulation by timing.

FDPS exposes the leaf size (n_leaf, `src/main.cpp:86-87`) and the i-group
size (n_group, `src/main.cpp:88-90`) as tuning knobs and leaves the choice to
the user.  Here the same two knobs are timed on the workload itself: each
candidate builds a Simulation from the same initial state, runs `warmup`
steps untimed and `steps` timed steps, and the fastest mean step wins.

On this device the i-group is one 64-lane wavefront (one lane per
i-particle), so n_group ranges over divisors of the wave width (larger FDPS
groups only amortise host-side list construction, which is not on this path);
the j-tile staged through LDS is one wave of j-particles per tile by
construction.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from .sph import Context, Simulation

N_GROUP_CHOICES = (32, 48, 64)
N_LEAF_CHOICES = (8, 16, 32)


@dataclass
class TuneResult:
    n_group: int
    n_leaf: int
    ms_per_step: float
    table: list = field(default_factory=list)  # (n_group, n_leaf, ms_per_step) per candidate


def tune(ctx: Context, ic: dict, *, n_groups=N_GROUP_CHOICES, n_leaves=N_LEAF_CHOICES, warmup: int = 1,
         steps: int = 3, log=None, **params) -> TuneResult:
    """Time every (n_group, n_leaf) pair on `ic` and return the fastest.

    `params` are passed to every Simulation (theta, nsys, ...).  Each candidate
    starts from the same initial state, so all of them time the same steps.
    """
    table = []
    for ng in n_groups:
        for nl in n_leaves:
            sim = Simulation(ctx, ic, n_group=ng, n_leaf=nl, **params)
            if warmup:
                sim.step(warmup)
            t0 = time.perf_counter()
            sim.step(steps)  # returns after the context's stream has drained
            ms = 1e3 * (time.perf_counter() - t0) / steps
            table.append((ng, nl, ms))
            if log:
                log(f"[autotune] n_group={ng:3d} n_leaf={nl:3d}: {ms:8.3f} ms/step")
            del sim
    best = min(table, key=lambda r: r[2])
    return TuneResult(best[0], best[1], best[2], table)
