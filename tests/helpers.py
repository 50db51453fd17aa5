"""Input generators and comparison helpers shared by the tests."""
from __future__ import annotations

import numpy as np

from fdps_sph_amd import structs as S


def cloud(n: int, seed: int, dx: float = 1.0e5, hjit: float = 0.3, vel: float = 1.0e3):
    """A jittered lattice of n particles with varied h, mass, density and state."""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(n ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)[:n]
    pos = (g + 0.3 * (rng.random((n, 3)) - 0.5)) * dx + np.array([1.0e6, -2.0e6, 3.0e5])
    return {
        "pos": pos,
        "vel": rng.normal(0.0, vel, (n, 3)),
        "mass": 2680.0 * dx ** 3 * (1.0 + 0.1 * (rng.random(n) - 0.5)),
        "smth": 1.2 * dx * (1.0 + hjit * (rng.random(n) - 0.5)),
        "dens": 2680.0 * (1.0 + 0.2 * (rng.random(n) - 0.5)),
        "pres": 1.0e9 * (1.0 + 0.5 * rng.random(n)),
        "snds": 3.0e3 * (1.0 + 0.5 * rng.random(n)),
        "grad_smth": 1.0 + 0.2 * (rng.random(n) - 0.5),
        "Bal": rng.random(n),
    }


def walks_from(c: dict, sizes_i, seed: int, j_frac: float = 1.0):
    """Walks (i-range, j-index list) over the cloud: i-groups are contiguous
    chunks of the given sizes, j-lists are random subsets (in random order)."""
    rng = np.random.default_rng(seed)
    n = len(c["mass"])
    walks = []
    start = 0
    for ni in sizes_i:
        ii = np.arange(start, start + ni) % n
        start += ni
        nj = int(round(j_frac * n))
        jj = rng.permutation(n)[:nj]
        walks.append((ii, jj))
    return walks


def epi_dens(c, ii):
    a = np.zeros(len(ii), S.EPI_DENS)
    a["pos"] = c["pos"][ii]; a["mass"] = c["mass"][ii]; a["smth"] = c["smth"][ii]; a["id"] = ii
    return a


def epj_dens(c, jj):
    a = np.zeros(len(jj), S.EPJ_DENS)
    a["mass"] = c["mass"][jj]; a["pos"] = c["pos"][jj]; a["smth"] = c["smth"][jj]
    return a


def epi_drvt(c, ii):
    a = np.zeros(len(ii), S.EPI_DRVT)
    a["pos"] = c["pos"][ii]; a["vel"] = c["vel"][ii]; a["smth"] = c["smth"][ii]; a["dens"] = c["dens"][ii]
    return a


def epj_drvt(c, jj):
    a = np.zeros(len(jj), S.EPJ_DRVT)
    a["mass"] = c["mass"][jj]; a["pos"] = c["pos"][jj]; a["vel"] = c["vel"][jj]; a["smth"] = c["smth"][jj]
    return a


def epi_hydro(c, ii):
    a = np.zeros(len(ii), S.EPI_HYDRO)
    for f in ("pos", "vel", "smth", "dens", "pres", "snds", "grad_smth", "Bal"):
        a[f] = c[f][ii]
    a["AVa"] = 1.0
    a["id"] = ii
    return a


def epj_hydro(c, jj):
    a = np.zeros(len(jj), S.EPJ_HYDRO)
    for f in ("pos", "vel", "dens", "mass", "smth", "pres", "grad_smth", "snds", "Bal"):
        a[f] = c[f][jj]
    a["AVa"] = 1.0
    a["id"] = jj
    return a


def epi_grav(c, ii):
    a = np.zeros(len(ii), S.EPI_GRAV)
    a["pos"] = c["pos"][ii]; a["eps2"] = c["smth"][ii] ** 2; a["id"] = ii
    return a


def epj_grav(c, jj):
    a = np.zeros(len(jj), S.EPJ_GRAV)
    a["pos"] = c["pos"][jj]; a["mass"] = c["mass"][jj]; a["id"] = jj
    return a


def rel_err(x, ref, scale=None):
    """max |x - ref| / max(|ref|, scale): scale bounds the error of sums whose
    terms cancel (the size of the summed terms)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.abs(ref)
    if scale is not None:
        den = np.maximum(den, np.asarray(scale))
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.abs(x - ref) / den)) if x.size else 0.0
