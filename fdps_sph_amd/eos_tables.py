"""EoS tables: the reference's rho-u tables and the conversion that makes them.

The bundled ``*.rho_u.txt`` tables are the reference's eos/*.rho_u.txt data
(120 x 120 cells, fields rho, u, T, P, c_s, S; read at src/EoS.h:444-450).
They were made by eos/ANEOS_conversion.py from the iSALE-format ANEOS tables
``*.table.txt`` (rho, T, u, P, c_s, S on a 120-density x 61-temperature grid),
which are bundled here as well.  ``convert`` restates that conversion
(eos/ANEOS_conversion.py:38-128):

* u_min = 1.01 min(u) and u_max = max(u) over the whole source table; the new
  energy axis is u_min (u_max/u_min)^(m/(nu-1)), m = 0..nu-1;
* per density row, T(u) by linear interpolation of the row's (u, T) pairs with
  linear extrapolation, then P, c_s and S as linear functions of T (again
  extrapolated) at those temperatures;
* P clipped at 0.

The reference script loops its per-row conversion over ``range(nu)`` while
indexing density rows, so it only runs for nu equal to the number of source
densities (120).  ``convert`` takes the density axis as an argument instead:
the source rows (nrho=None) or nrho log-spaced densities between the first
and last source density, at which the source's u, P, c_s and S are
interpolated linearly in density at each source temperature before the
per-row conversion (the same linear-in-density rule as the run-time lookup,
src/interpolation.h:60-102).  This is how the 200 x 200 tables are built.

The shipped granite and iron rho-u tables were made by an earlier form of the
conversion: energy axis from u_min = min(u) (no 1.01 factor) and no P >= 0
clip (they hold 359 and 368 negative pressures); the dunite table follows the
script as it stands.
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
SOURCE_PATHS = {
    "granite": os.path.join(TABLE_DIR, "granite.table.txt"),
    "iron": os.path.join(TABLE_DIR, "iron___.table.txt"),
    "dunite": os.path.join(TABLE_DIR, "dunite_.table.txt"),
}
# Tillotson parameters of granite (rho0, u0, u_iv, u_cv, A, B, a, b, alpha,
# beta), src/EoS.h:446
GRANITE_TILLOTSON = (2680.0, 16.0e6, 3.5e6, 18.0e6, 18.0e9, 18.0e9, 0.5, 1.3, 5.0, 5.0)
# tables whose shipped rho-u form was written before the P >= 0 clip
UNCLIPPED = ("granite", "iron")


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


def _number(tok: str) -> float:
    """A number of the iSALE table: 'm.mmmE+ee', or 'm.mmm-eee' with the 'E'
    dropped once the exponent has three digits (eos/ANEOS_conversion.py:20-34
    reads both as mantissa * 10**exponent)."""
    if "E" in tok:
        m, e = tok.split("E")
    else:
        k = max(tok.rfind("-"), tok.rfind("+"))
        if k <= 0:
            return float(tok)
        m, e = tok[:k], tok[k:]
    return float(m) * 10 ** float(e)


def read_source(path: str) -> dict:
    """An iSALE-format ANEOS table: a header line, then one block of
    (rho, T, u, P, c_s, S) lines per density, blocks separated by blank lines,
    every block over the same temperatures."""
    with open(path) as f:
        lines = f.read().split("\n")[1:]
    blocks, cur = [], []
    for ln in lines:
        if ln.strip():
            cur.append([_number(t) for t in ln.split()[:6]])
        elif cur:
            blocks.append(cur)
            cur = []
    if cur:
        blocks.append(cur)
    nt = len(blocks[0])
    if any(len(b) != nt for b in blocks):
        raise ValueError(f"{path}: density blocks of unequal length")
    a = np.array(blocks)  # [nr][nt][6]
    return {"density": a[:, 0, 0].copy(), "temperature": a[0, :, 1].copy(), "energy": a[:, :, 2].copy(),
            "pressure": a[:, :, 3].copy(), "soundspeed": a[:, :, 4].copy(), "entropy": a[:, :, 5].copy()}


def _interp_linear(x, y, xn):
    """Piecewise-linear interpolation with linear extrapolation beyond both
    ends: the points sorted by x (stably), each new point placed by a
    left-sided binary search clipped to the first and last intervals, and
    evaluated on the chord of its interval."""
    o = np.argsort(x, kind="mergesort")
    x = np.asarray(x)[o]
    y = np.asarray(y)[o]
    k = np.clip(np.searchsorted(x, xn), 1, len(x) - 1)
    lo, hi = k - 1, k
    slope = (y[hi] - y[lo]) / (x[hi] - x[lo])
    return slope * (xn - x[lo]) + y[lo]


def convert(src: dict, nu: int = 120, nrho: int | None = None, clip: bool = True,
            umin_factor: float = 1.01) -> np.ndarray:
    """cells[nrho][nu][6] = (rho, u, T, P, c_s, S) from a source table (see
    the module docstring; eos/ANEOS_conversion.py:83-128).  clip=False and
    umin_factor=1 give the earlier form of the conversion that the shipped
    granite and iron tables were made with."""
    T = src["temperature"]
    umin = np.min(src["energy"]) * umin_factor
    umax = np.max(src["energy"])
    delta = (umax / umin) ** (1.0 / (nu - 1))
    new_u = np.array([umin * delta ** m for m in range(nu)])
    dens = src["density"]
    fields = ("energy", "pressure", "soundspeed", "entropy")
    if nrho is None:
        rows = dens
        rowf = {f: src[f] for f in fields}
    else:
        rows = np.exp(np.linspace(np.log(dens[0]), np.log(dens[-1]), nrho))
        rows[0], rows[-1] = dens[0], dens[-1]
        k = np.clip(np.searchsorted(dens, rows), 1, len(dens) - 1)
        w = ((rows - dens[k - 1]) / (dens[k] - dens[k - 1]))[:, None]
        rowf = {f: (1.0 - w) * src[f][k - 1] + w * src[f][k] for f in fields}
    out = np.empty((len(rows), nu, 6))
    for m in range(len(rows)):
        t = _interp_linear(rowf["energy"][m], T, new_u)
        out[m, :, 0] = rows[m]
        out[m, :, 1] = new_u
        out[m, :, 2] = t
        out[m, :, 3] = _interp_linear(T, rowf["pressure"][m], t)
        out[m, :, 4] = _interp_linear(T, rowf["soundspeed"][m], t)
        out[m, :, 5] = _interp_linear(T, rowf["entropy"][m], t)
    if clip:
        out[:, :, 3] = np.maximum(out[:, :, 3], 0.0)
    return out


def regrid(name: str, nrho: int = 200, nu: int = 200) -> np.ndarray:
    """The nrho x nu table of a material converted from its ANEOS source."""
    return convert(read_source(SOURCE_PATHS[name]), nu=nu, nrho=nrho)


_CACHE: dict = {}


def table(name: str, grid: int | None = None) -> np.ndarray:
    """The shipped 120 x 120 table (grid None) or a grid x grid conversion of
    the material's ANEOS source."""
    key = (name, grid)
    if key not in _CACHE:
        _CACHE[key] = read(PATHS[name]) if grid is None else regrid(name, grid, grid)
    return _CACHE[key]
