"""The entry point that allocates nothing (sphgpu_eos_pressure_noalloc),
called through the C-ABI: its results equal sphgpu_eos_pressure's bit for
bit on two calls in a row, the second smaller than the first (the shared
scratch keeps the first call's contents), and it refuses sizes beyond its
fixed scratch with a capacity error."""
import numpy as np
import pytest

from fdps_sph_amd import _lib
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd.sph import Context

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    mat = rng.integers(0, 3, n).astype(np.int32)
    rho = np.exp(rng.uniform(np.log(1.0), np.log(3e4), n))
    u = np.exp(rng.uniform(np.log(1e3), np.log(1e9), n))
    return mat, rho, u


def _noalloc(ctx, mat, rho, u):
    L = _lib.lib()
    n = len(rho)
    p = np.full(n, np.nan)
    c = np.full(n, np.nan)
    rc = L.sphgpu_eos_pressure_noalloc(ctx.handle, n, _lib.ptr(mat), _lib.ptr(rho), _lib.ptr(u), _lib.ptr(p),
                                       _lib.ptr(c))
    return rc, p, c


def test_noalloc_pressure_matches_allocating_entry():
    ctx = Context(0, "fp64")
    try:
        for m, k in enumerate(("granite", "iron", "dunite")):
            ctx.load_table(m, E.PATHS[k])
        for n, seed in ((300_000, 1), (4_321, 2)):  # the second call smaller than the first
            print(f"[noalloc eos n={n}]", flush=True)
            mat, rho, u = _inputs(n, seed)
            rc, p, c = _noalloc(ctx, mat, rho, u)
            assert rc == _lib.OK, _lib.lib().sphgpu_last_error()
            ref_p, ref_c = ctx.pressure(mat, rho, u)
            assert np.array_equal(p, ref_p)
            assert np.array_equal(c, ref_c)
        # sizes beyond the fixed scratch are refused before any work
        n = _lib.NOALLOC_MAX_N + 1
        rc = _lib.lib().sphgpu_eos_pressure_noalloc(ctx.handle, n, None, None, None, None, None)
        assert rc == _lib.ERR_CAPACITY
        assert b"SPHGPU_NOALLOC_MAX_N" in _lib.lib().sphgpu_last_error()
        # an unknown material is an EoS error, as for the allocating entry
        mat, rho, u = _inputs(10, 3)
        mat[4] = 9
        rc, _, _ = _noalloc(ctx, mat, rho, u)
        assert rc == _lib.ERR_EOS
        assert _lib.lib().sphgpu_eos_pressure_noalloc(ctx.handle, 0, None, None, None, None, None) == _lib.OK
    finally:
        ctx.close()
