"""Pick the tree/group/launch parameters of a device-resident simulation by timing.

FDPS exposes the leaf size (n_leaf, `src/main.cpp:86-87`) and the i-group
size (n_group, `src/main.cpp:88-90`) as tuning knobs and leaves the choice to
the user.  Here they are timed on the workload itself, together with two
knobs of this implementation:

* persistent: which interaction kernels run as one persistent launch taking
  the groups of every system from per-XCD counters (a bitmask: 1 density,
  2 derivative, 4 hydro, 8 gravity), the others one workgroup per group;
* list_front_lds: the entries of a tree walk's frontier kept in LDS before it
  continues in global memory -- the LDS tile of the list construction.  The
  interaction kernels' j-tile is one wave of j-particles (64 entries) by
  construction: each lane stages one entry;
* walk_waves: the persistent waves of each list walk, which runs beside the
  force kernels and shares the GPU with them;
* vsig_split: whether the hydro pass's dt (v_sig_max over each list) runs as
  a kernel of its own beside the gravity kernel (0: inside the hydro kernel,
  3: split in either precision form).

A 64-lane wavefront holds up to 64 i-particles, one per lane; a larger
i-group (n_group 128 .. 1024, FDPS's own range) runs as one work item per
wave of its particles, the waves sharing the group's lists: fewer, longer
lists and fewer walks, against more pairs outside the supports.  The search
is staged: every (n_group, n_leaf) pair first, then the launch form and the
LDS frontier at the best pair, then the walks' waves, then the dt split.  Each candidate builds
a Simulation from the same initial state, runs `warmup` steps untimed and
`steps` timed steps, and the fastest mean step wins.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from .sph import Context, Simulation

N_GROUP_CHOICES = (32, 64, 128, 256, 512, 1024)
N_LEAF_CHOICES = (8, 16, 32)
# persistent-launch bitmask: none, gravity only, every interaction kernel
PERSISTENT_CHOICES = (0, 8, 15)
FRONT_LDS_CHOICES = (128, 512)
WALK_WAVES_CHOICES = (2048, 4096, 8192)
VSIG_SPLIT_CHOICES = (0, 3)


@dataclass
class TuneResult:
    n_group: int
    n_leaf: int
    ms_per_step: float
    persistent: int = 0
    list_front_lds: int = 512
    walk_waves: int = 4096
    vsig_split: int = 1
    table: list = field(default_factory=list)  # dict per candidate: the six knobs and ms_per_step

    def params(self) -> dict:
        return {"n_group": self.n_group, "n_leaf": self.n_leaf, "persistent": self.persistent,
                "list_front_lds": self.list_front_lds, "walk_waves": self.walk_waves,
                "vsig_split": self.vsig_split}


def _time(ctx, ic, warmup, steps, **params) -> float:
    sim = Simulation(ctx, ic, **params)
    if warmup:
        sim.step(warmup)
    t0 = time.perf_counter()
    sim.step(steps)  # returns after the context's stream has drained
    return 1e3 * (time.perf_counter() - t0) / steps


def tune(ctx: Context, ic: dict, *, n_groups=N_GROUP_CHOICES, n_leaves=N_LEAF_CHOICES,
         persistents=PERSISTENT_CHOICES, front_lds=FRONT_LDS_CHOICES, wave_counts=WALK_WAVES_CHOICES,
         vsig_splits=VSIG_SPLIT_CHOICES, warmup: int = 1, steps: int = 3, log=None, **params) -> TuneResult:
    """Time the candidates on `ic` and return the fastest.

    `params` are passed to every Simulation (theta, nsys, ...).  Each candidate
    starts from the same initial state, so all of them time the same steps.
    """
    table = []
    knobs = ("n_group", "n_leaf", "persistent", "list_front_lds", "walk_waves", "vsig_split")
    params = {k: v for k, v in params.items() if k not in knobs}

    def run(**knobs):
        ms = _time(ctx, ic, warmup, steps, **knobs, **params)
        row = dict(knobs, ms_per_step=ms)
        table.append(row)
        if log:
            log("[autotune] " + " ".join(f"{k}={v}" for k, v in knobs.items()) + f": {ms:8.3f} ms/step")
        return row

    ww0 = 4096 if 4096 in wave_counts else wave_counts[0]
    vs0 = vsig_splits[0]
    base = {"persistent": persistents[0], "list_front_lds": front_lds[-1], "walk_waves": ww0, "vsig_split": vs0}
    for ng in n_groups:
        for nl in n_leaves:
            run(n_group=ng, n_leaf=nl, **base)
    best = min(table, key=lambda r: r["ms_per_step"])
    for p in persistents:
        for fl in front_lds:
            if p == best["persistent"] and fl == best["list_front_lds"]:
                continue
            run(n_group=best["n_group"], n_leaf=best["n_leaf"], persistent=p, list_front_lds=fl, walk_waves=ww0,
                vsig_split=vs0)
    for knob, choices in (("walk_waves", wave_counts), ("vsig_split", vsig_splits)):
        best = min(table, key=lambda r: r["ms_per_step"])
        for v in choices:
            if v != best[knob]:
                run(**{k: best[k] for k in knobs if k != knob}, **{knob: v})
    best = min(table, key=lambda r: r["ms_per_step"])
    return TuneResult(best["n_group"], best["n_leaf"], best["ms_per_step"], best["persistent"],
                      best["list_front_lds"], best["walk_waves"], best["vsig_split"], table)
