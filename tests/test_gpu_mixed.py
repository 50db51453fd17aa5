"""Precision form (FP32 group-relative pair arithmetic, compensated sums)
against the FP64 oracle on the same inputs.  Tolerances: 1e-6 relative on
density and smoothing length, 1e-5 on accelerations and du/dt (relative to
each particle's own magnitude, or to 1e-3 of the largest one for vectors whose
terms cancel to near zero), 1e-5 on the gravitational potential."""
import numpy as np
import pytest

from fdps_sph_amd import Context, Simulation
from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

from . import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx32():
    c = Context(0, "mixed")
    c.load_table(0, E.PATHS["granite"])
    c.load_table(1, E.PATHS["iron"])
    c.load_table(2, E.PATHS["dunite"])
    yield c
    c.close()


def test_mixed_functors(ctx32, oracle):
    """Errors are measured against the physical scale of each output: sigma_v / h
    for velocity gradients, P / (rho h) for the pressure acceleration and
    (P / rho) (sigma_v / h) for du/dt -- these sums cancel strongly for the
    random velocities used here, so their own value is no scale."""
    print("[mixed functors]", flush=True)
    c = H.cloud(700, 21)
    walks = H.walks_from(c, [64, 100, 3], 21, j_frac=0.7)
    sig_v = float(np.std(c["vel"]))
    res = ctx32.calc_density([H.epi_dens(c, i) for i, _ in walks], [H.epj_dens(c, j) for _, j in walks])
    for (i, j), r in zip(walks, res):
        ref = oracle.calc_density(H.epi_dens(c, i), H.epj_dens(c, j))
        assert H.rel_err(r["dens"], ref["dens"]) < 1e-6
        assert H.rel_err(r["smth"], ref["smth"]) < 1e-6
    res = ctx32.calc_derivative([H.epi_drvt(c, i) for i, _ in walks], [H.epj_drvt(c, j) for _, j in walks])
    for (i, j), r in zip(walks, res):
        ref = oracle.calc_derivative(H.epi_drvt(c, i), H.epj_drvt(c, j))
        sc = sig_v / c["smth"][i]
        assert H.rel_err(r["div_v"], ref["div_v"], sc) < 1e-5
        assert H.rel_err(r["rot_v"], ref["rot_v"], sc[:, None]) < 1e-5
        assert H.rel_err(r["grad_smth"], ref["grad_smth"]) < 1e-5
    res = ctx32.calc_hydro([H.epi_hydro(c, i) for i, _ in walks], [H.epj_hydro(c, j) for _, j in walks])
    for (i, j), r in zip(walks, res):
        ref = oracle.calc_hydro(H.epi_hydro(c, i), H.epj_hydro(c, j))
        pa = c["pres"][i] / (c["dens"][i] * c["smth"][i])
        assert H.rel_err(r["acc"], ref["acc"], pa[:, None]) < 1e-5
        assert H.rel_err(r["eng_dot"], ref["eng_dot"], c["pres"][i] / c["dens"][i] * sig_v / c["smth"][i]) < 1e-5
        assert H.rel_err(r["dt"], ref["dt"]) < 1e-6
    res = ctx32.calc_gravity([H.epi_grav(c, i) for i, _ in walks], [H.epj_grav(c, j) for _, j in walks])
    for (i, j), r in zip(walks, res):
        ref = oracle.calc_grav(H.epi_grav(c, i), H.epj_grav(c, j))
        sa = np.abs(ref["acc"]).max(1, keepdims=True)
        assert H.rel_err(r["acc"], ref["acc"], sa) < 1e-5
        assert H.rel_err(r["pot"], ref["pot"]) < 1e-5


def test_mixed_simulation_vs_fp64(ctx64, ctx32):
    """Same initial state, one force evaluation in each precision."""
    print("[mixed sim vs fp64, 10k granite]", flush=True)
    ic = IC.granite_impact(9000, 1000, seed=1000)
    a = Simulation(ctx64, ic)
    b = Simulation(ctx32, ic)
    sa, sb = a.state(), b.state()
    ia, ib = np.argsort(sa["id"]), np.argsort(sb["id"])
    def err(f, scale_frac=None):
        x, y = sb[f][ib], sa[f][ia]
        if y.ndim == 2:
            num = np.linalg.norm(x - y, axis=1)
            den = np.linalg.norm(y, axis=1)
        else:
            num, den = np.abs(x - y), np.abs(y)
        if scale_frac is not None:
            den = np.maximum(den, scale_frac * den.max())
        return float(np.max(num / den))
    e = {"dens": err("dens"), "smth": err("smth"), "grav": err("grav"), "pot": err("pot"),
         "acc": err("acc", 1e-3), "engdot": err("engdot", 1e-3)}
    print("   max rel err:", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert e["dens"] < 1e-6 and e["smth"] < 1e-6
    assert e["grav"] < 1e-5 and e["pot"] < 1e-5
    assert e["acc"] < 1e-4 and e["engdot"] < 1e-4
