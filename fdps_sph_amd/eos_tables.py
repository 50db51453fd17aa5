"""EoS tables: the reference's rho-u tables and their regridding.

The bundled tables are the reference's eos/*.rho_u.txt data (120 x 120 cells,
fields rho, u, T, P, c_s, S; read at src/EoS.h:444-450).  ``regrid`` resamples
a table onto an N x M grid whose density and energy axes are both
log-spaced between the source table's first and last nodes (the energy axis
is log-spaced in the source already, eos/ANEOS_conversion.py:83-96), taking
every property from the source by the reference's bilinear rule
(src/interpolation.h:60-102).
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_DIR = os.path.join(HERE, "eos")
PATHS = {
    "granite": os.path.join(TABLE_DIR, "granite.rho_u.txt"),
    "iron": os.path.join(TABLE_DIR, "iron___.rho_u.txt"),
    "dunite": os.path.join(TABLE_DIR, "dunite_.rho_u.txt"),
}


def read(path: str) -> np.ndarray:
    """cells[nrho][nu][6] of a reference-format table (src/EoS.h:290-385)."""
    with open(path) as f:
        head = f.readline().split()
        nr, nu = int(head[1]), int(head[2])
        rows = [ln for ln in f if ln.strip() and not ln.startswith("#")]
    data = np.array([[float(v) for v in ln.split()[:6]] for ln in rows[: nr * nu]])
    return data.reshape(nr, nu, 6)


def write(path: str, cells: np.ndarray) -> None:
    nr, nu, _ = cells.shape
    with open(path, "w") as f:
        f.write(f"#  {nr} {nu} :Grid numbers for density and internal energy\n")
        f.write("# Density, Internal energy, Temperature (K), Pressure (Pa), Sound speed (m/s), Entropy (J/K/kg)\n")
        for r in range(nr):
            for c in range(nu):
                f.write(" " + " ".join(f"{v: .15E}" for v in cells[r, c]) + "\n")


def bilinear(cells: np.ndarray, rho, u, prop: int) -> np.ndarray:
    """Vectorised BilinearInterpolation::interpolate of one property."""
    rax = cells[:, 0, 0]
    uax = cells[0, :, 1]
    nr, nu = len(rax), len(uax)
    rho = np.asarray(rho, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    li = np.minimum(np.searchsorted(rax, rho, side="left"), nr - 1)
    ci = np.minimum(np.searchsorted(uax, u, side="left"), nu - 1)
    T = cells[:, :, prop]
    inside = (li < nr - 1) & (ci < nu - 1) & (li > 0) & (ci > 0)
    out = T[li, ci].copy()
    l0 = np.where(inside, li - 1, 0)
    c0 = np.where(inside, ci - 1, 0)
    l1 = np.where(inside, li, 1)
    c1 = np.where(inside, ci, 1)
    xi = (rho - rax[l0]) / (rax[l1] - rax[l0])
    eta = (u - uax[c0]) / (uax[c1] - uax[c0])
    v = (1 - xi) * (1 - eta) * T[l0, c0] + xi * (1 - eta) * T[l1, c0] + (1 - xi) * eta * T[l0, c1] + xi * eta * T[l1, c1]
    out[inside] = v[inside]
    return out


def regrid(cells: np.ndarray, nrho: int = 200, nu: int = 200) -> np.ndarray:
    rax = cells[:, 0, 0]
    uax = cells[0, :, 1]
    new_r = np.exp(np.linspace(np.log(rax[0]), np.log(rax[-1]), nrho))
    new_u = np.exp(np.linspace(np.log(uax[0]), np.log(uax[-1]), nu))
    new_r[0], new_r[-1] = rax[0], rax[-1]
    new_u[0], new_u[-1] = uax[0], uax[-1]
    R, U = np.meshgrid(new_r, new_u, indexing="ij")
    out = np.empty((nrho, nu, 6))
    out[:, :, 0] = R
    out[:, :, 1] = U
    for prop in (2, 3, 4, 5):
        out[:, :, prop] = bilinear(cells, R.ravel(), U.ravel(), prop).reshape(nrho, nu)
    return out


_CACHE: dict = {}


def table(name: str, grid: int | None = None) -> np.ndarray:
    key = (name, grid)
    if key not in _CACHE:
        c = read(PATHS[name])
        _CACHE[key] = c if grid is None else regrid(c, grid, grid)
    return _CACHE[key]
