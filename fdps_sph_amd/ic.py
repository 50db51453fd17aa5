"""Seeded initial conditions for the giant-impact workloads.

The reference builds its bodies in GI_universal::setupIC (src/GI.h:35-445)
from a lattice cut to a sphere, with an unseeded rand() particle removal
(src/GI.h:10-24).  The generator here follows the same idea with a seeded
SplitMix64 stream so that every run (CPU oracle and GPU) starts from the same
particles:

* each body is a cubic lattice cut to a sphere (or to a core + mantle pair of
  shells, one lattice spacing per material so that particles have equal
  mass), trimmed to exactly the requested count by seeded removal;
* positions get seeded uniform jitter of +-5 % of the spacing per axis;
* the impactor centre is at (sqrt((2 R_t)^2 - b^2), b, 0) with impact
  parameter b = (R_t + R_i) sin(angle), moving along -x.

Material ids index the EoS tables loaded into the context (see MATERIALS).
"""
from __future__ import annotations

import math

import numpy as np

G = 6.67e-11  # src/force.h:172

# material ids used by the workloads
GRANITE, IRON, DUNITE = 0, 1, 2
RHO0 = {GRANITE: 2680.0,  # Tillotson granite rho0 (src/EoS.h:446)
        IRON: 7800.0,      # iron rho0 (src/EoS.h:448)
        DUNITE: 3320.0}    # dunite reference density (no value in the reference)


class SplitMix64:
    """SplitMix64 generator (Steele, Lea, Flood 2014) -> uniform doubles in [0, 1)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            gamma = np.uint64(0x9E3779B97F4A7C15)
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + gamma * idx
            self.state = self.state + gamma * np.uint64(n)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _lattice_r(dx: float, rmax: float, rmin: float = 0.0):
    """Cubic lattice points (cell centres, spacing dx) with rmin < r <= rmax,
    in lattice (x, y, z index) order, and their radii."""
    k = int(math.ceil(rmax / dx)) + 1
    a = (np.arange(-k, k + 1) + 0.5) * dx
    X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    r = np.sqrt((p * p).sum(1))
    sel = (r <= rmax) & (r > rmin)
    return p[sel], r[sel]


def _lattice(dx: float, rmax: float, rmin: float = 0.0) -> np.ndarray:
    return _lattice_r(dx, rmax, rmin)[0]


def _trim(p: np.ndarray, n: int, rng: SplitMix64) -> np.ndarray:
    if len(p) < n:
        raise ValueError(f"lattice holds {len(p)} < {n} particles")
    if len(p) == n:
        return p
    # seeded removal: keep the n entries with the smallest random keys, in lattice order
    keys = rng.uniform(len(p))
    keep = np.sort(np.argsort(keys, kind="stable")[:n])
    return p[keep]


def _shell(n: int, rho: float, m: float, rmin: float, rng: SplitMix64, jitter=0.05):
    """n particles of mass m at density rho filling the shell outside rmin."""
    vol = n * m / rho
    rmax = (rmin ** 3 + 3.0 * vol / (4.0 * math.pi)) ** (1.0 / 3.0)
    dx = (m / rho) ** (1.0 / 3.0)
    # boundary effects: widen the outer radius in steps of 0.2 % until the shell
    # holds n points.  One lattice wide enough for every radius tried is cut
    # once; a point's coordinates and its lattice order do not depend on the
    # extent of the lattice it was generated in.
    wide = 1.25
    pw, rw = _lattice_r(dx, rmax * wide, rmin)
    rs = np.sort(rw)
    grow = 1.0
    while np.searchsorted(rs, rmax * grow, side="right") < n:
        grow += 0.002
        if rmax * grow > rmax * wide:
            wide *= 1.25
            pw, rw = _lattice_r(dx, rmax * wide, rmin)
            rs = np.sort(rw)
    p = pw[rw <= rmax * grow]
    p = _trim(p, n, rng)
    p = p + (rng.uniform(3 * n).reshape(n, 3) - 0.5) * (2.0 * jitter * dx)
    return p, rmax, dx


def body(layers, m: float, rng: SplitMix64):
    """layers: [(n, material)] from the centre outwards; equal particle mass m."""
    pos, mat = [], []
    rmin = 0.0
    for n, mt in layers:
        if n <= 0:
            continue
        p, rmin, _ = _shell(n, RHO0[mt], m, rmin, rng)
        pos.append(p)
        mat.append(np.full(n, mt, dtype=np.int32))
    return np.concatenate(pos), np.concatenate(mat), rmin


def impact(target_layers, impactor_layers, m_target: float, m_impactor: float, speed: float, angle_deg: float,
           seed: int, eng: float = 1.0e6, speed_in_vesc: bool = False):
    """Two-body impact; returns the IC dict consumed by Simulation."""
    rng = SplitMix64(seed)
    pt, mt, Rt = body(target_layers, m_target, rng)
    pi, mi, Ri = body(impactor_layers, m_impactor, rng)
    Mt = m_target * len(pt)
    Mi = m_impactor * len(pi)
    if speed_in_vesc:
        vesc = math.sqrt(2.0 * G * (Mt + Mi) / (Rt + Ri))  # as src/GI.h:204
        speed = speed * vesc
    b = (Rt + Ri) * math.sin(math.radians(angle_deg))
    cx = math.sqrt(max((2.0 * Rt) ** 2 - b * b, 0.0))
    pi = pi + np.array([cx, b, 0.0])
    pos = np.concatenate([pt, pi])
    vel = np.zeros_like(pos)
    vel[len(pt):, 0] = -speed
    mass = np.concatenate([np.full(len(pt), m_target), np.full(len(pi), m_impactor)])
    mat = np.concatenate([mt, mi])
    dens = np.array([RHO0[int(k)] for k in range(3)])[mat]
    return {"pos": pos, "vel": vel, "mass": mass, "dens": dens, "eng": np.full(len(mass), eng), "mat": mat,
            "tag": np.concatenate([np.zeros(len(pt), np.int32), np.ones(len(pi), np.int32)]),
            "R_target": Rt, "R_impactor": Ri}


def granite_impact(n_target=9000, n_impactor=1000, R_target=6371e3, R_impactor=3390e3, speed=9000.0, angle=45.0,
                   seed=1000):
    """Uniform granite target + impactor with equal mass within each body (the
    10k-particle CPU-runnable workload: 9,000 + 1,000 particles, 9 km/s, 45 deg)."""
    rho = RHO0[GRANITE]
    mt = rho * 4.0 / 3.0 * math.pi * R_target ** 3 / n_target
    mi = rho * 4.0 / 3.0 * math.pi * R_impactor ** 3 / n_impactor
    return impact([(n_target, GRANITE)], [(n_impactor, GRANITE)], mt, mi, speed, angle, seed)


def iron_dunite_impact(n_target=900_000, n_impactor=100_000, M_target=5.97e24, core_mass_frac=0.3, vesc_factor=1.1,
                       angle=45.0, seed=1001):
    """Iron core (30 % of the mass) + dunite mantle target and a 10x lighter
    impactor of the same structure, equal particle mass, 1.1 v_esc at 45 deg."""
    m = M_target / n_target
    nc_t = int(round(core_mass_frac * n_target))
    nc_i = int(round(core_mass_frac * n_impactor))
    return impact([(nc_t, IRON), (n_target - nc_t, DUNITE)], [(nc_i, IRON), (n_impactor - nc_i, DUNITE)], m, m,
                  vesc_factor, angle, seed, speed_in_vesc=True)


def three_material_impact(n_target=12_000_000, n_impactor=4_000_000, M_target=5.97e24, core_mass_frac=0.3,
                          vesc_factor=1.1, angle=45.0, seed=1004):
    """Iron core + dunite mantle target and a granite impactor, equal particle
    mass (the impactor mass follows from its particle count), 1.1 v_esc at
    45 deg: the 16M-particle three-material workload."""
    m = M_target / n_target
    nc_t = int(round(core_mass_frac * n_target))
    return impact([(nc_t, IRON), (n_target - nc_t, DUNITE)], [(n_impactor, GRANITE)], m, m, vesc_factor, angle, seed,
                  speed_in_vesc=True)


def batch_of_impacts(n_systems=64, n_min=5_000, n_max=200_000, seed=1003):
    """Independent granite impacts with log-uniform particle counts, speeds
    U[5,15] km/s and angles U[0,75] deg; 90/10 target/impactor split."""
    rng = SplitMix64(seed)
    u = rng.uniform(3 * n_systems).reshape(n_systems, 3)
    ns = np.round(np.exp(np.log(n_min) + u[:, 0] * (np.log(n_max) - np.log(n_min)))).astype(np.int64)
    parts = []
    for s in range(n_systems):
        nt = int(round(0.9 * ns[s]))
        ni = int(ns[s]) - nt
        ic = granite_impact(nt, ni, speed=5000.0 + 10000.0 * u[s, 1], angle=75.0 * u[s, 2], seed=seed + 1 + s)
        ic["sys"] = np.full(len(ic["mass"]), s, dtype=np.int32)
        parts.append(ic)
    out = {k: np.concatenate([p[k] for p in parts]) for k in ("pos", "vel", "mass", "dens", "eng", "mat", "sys")}
    out["counts"] = ns
    return out
