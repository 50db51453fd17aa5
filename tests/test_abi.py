"""The C-ABI library loads and exports every entry point of include/fdps_sph_gpu.h
(no GPU needed: no compute call is made)."""
import ctypes
import os
import re

import numpy as np
import pytest

from fdps_sph_amd import _lib
from fdps_sph_amd import structs as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fdps_sph_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sphgpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("sphgpu_dens_dispatch", "sphgpu_dens_retrieve", "sphgpu_drvt_dispatch", "sphgpu_drvt_retrieve",
                 "sphgpu_hydro_dispatch", "sphgpu_hydro_retrieve", "sphgpu_grav_dispatch", "sphgpu_grav_retrieve",
                 "sphgpu_eos_pressure", "sphgpu_sim_init", "sphgpu_sim_step"):
        assert must in names


def test_library_exports_every_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_reference_sizes():
    for name, size in S.SIZES.items():
        assert getattr(S, name).itemsize == size, name


def test_errors_without_context():
    L = _lib.lib()
    assert L.sphgpu_dens_retrieve(None, 0, 0, None, None) == _lib.ERR_ARG
    assert L.sphgpu_destroy(None) == _lib.OK
    assert b"context" in L.sphgpu_last_error()
    p = _lib.SimParams()
    L.sphgpu_sim_default_params(ctypes.byref(p))
    assert p.n_group == 64 and p.n_leaf == 16 and p.theta == 0.5
    assert p.list_cap_front == 4096 and p.list_front_lds == 512


@pytest.mark.parametrize("name,kw", [("granite", dict(clip=False, umin_factor=1.0)),
                                     ("iron", dict(clip=False, umin_factor=1.0)),
                                     ("dunite", {})])
def test_conversion_reproduces_shipped_tables(name, kw):
    """eos/ANEOS_conversion.py turned the iSALE ANEOS tables (*.table.txt) into
    the shipped rho-u tables, printed to 9 significant digits; converting the
    same source at nu = 120 gives every cell to that precision (granite and
    iron by the script's earlier form: u_min = min u, no P clip)."""
    from fdps_sph_amd import eos_tables as E
    src = E.read_source(E.SOURCE_PATHS[name])
    ship = E.read(E.PATHS[name])
    c = E.convert(src, nu=120, **kw)
    assert c.shape == ship.shape == (120, 120, 6)
    zero = ship == 0  # clipped pressures: exactly 0 on both sides
    assert np.array_equal(c[zero], ship[zero])
    rel = np.abs(c[~zero] - ship[~zero]) / np.abs(ship[~zero])
    assert rel.max() <= 6e-9, (name, rel.max())
    if name == "dunite":  # its source has negative pressures; the clip removes them
        raw = E.convert(src, nu=120, clip=False)
        assert (raw[:, :, 3] < 0).sum() > 0 and np.array_equal(c[:, :, 3], np.maximum(raw[:, :, 3], 0))


def test_regrid_200():
    """200 log-spaced densities and energies: the first and last rows are the
    source's first and last densities, converted exactly as at the source
    rows; every other row lies between its two source rows."""
    from fdps_sph_amd import eos_tables as E
    src = E.read_source(E.SOURCE_PATHS["iron"])
    r = E.regrid("iron", 200, 200)
    rows = E.convert(src, nu=200)
    assert r.shape == (200, 200, 6)
    lr = np.log(r[:, 0, 0])
    assert r[0, 0, 0] == src["density"][0] and r[-1, 0, 0] == src["density"][-1]
    assert np.allclose(np.diff(lr), (lr[-1] - lr[0]) / 199, rtol=1e-9)
    assert np.array_equal(r[:, :, 1], rows[:1, :, 1].repeat(200, 0))  # one energy axis for every row
    assert np.array_equal(r[0], rows[0]) and np.array_equal(r[-1], rows[-1])
    assert (r[:, :, 3] >= 0).all() and np.isfinite(r).all()
    # a new row between source rows k-1 and k: at every source temperature its
    # energy lies between theirs, so its T at a given u does as well
    k = np.searchsorted(src["density"], r[1:-1, 0, 0])
    lo, hi = rows[k - 1, :, 2], rows[k, :, 2]
    T = r[1:-1, :, 2]
    assert np.all((T >= np.minimum(lo, hi) * (1 - 1e-12)) & (T <= np.maximum(lo, hi) * (1 + 1e-12)))
