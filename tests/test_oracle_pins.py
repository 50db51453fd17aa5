"""Closed-form and identity pins of the CPU oracle's hydro, derivative,
integrator and inverse-EoS functions (CPU only).

Each check follows from the mathematics of the scheme, not from a second copy
of the oracle's code, and is chosen so that a plausible slip fails it:

* energy identity: at uniform h and f = 1 the pair terms of CalcHydroForce
  (src/force.h:118-143) exchange kinetic and internal energy exactly,
  sum_i m_i (du_i + v_i . a_i) = 0, for any P, rho, c_s, Bal and velocities;
  it fixes du/dt's share of each acceleration term (pressure: A_i, viscosity:
  AV/2);
* continuum limits: under homologous compression v = -eps x the viscous heating
  of a lattice particle tends to (2 c eps^2 M1 + 15/4 eps^3 M2) B/rho_bar with
  the kernel moments M1 = int r W, M2 = int r^2 W (integration by parts of
  the pair sum); this fixes the viscosity's absolute scale, its signal
  velocity, its density and Balsara averages;
* the pressure force of a static cloud is minus the gradient of the kernel
  sums sum_j m_j (A_i W(x - x_j, h_i) + A_j W(x - x_j, h_j)), and du/dt is
  A_i/2 times the Lagrangian rate of f_i sum_j m_j W(., h_i) + sum_j m_j f_j
  W(., h_j) (central differences of the pinned kernel W): these fix which of
  A, h and f belong to i and to j;
* head-on pair time step dt = 0.6 h / (c_i + c_j + 3 |w|) and its maximum over
  the whole list;
* grad-h factor f = 1 / (1 + h/(3 rho) d rho/d h) with d rho/d h from central
  differences of the pinned CalcDensity (src/force.h:99-102);
* the KDK integrator with its predictor (src/class.h:201-229) and the damping
  force (src/GI.h:458-465): free drift, constant acceleration, the damped
  velocity factor inside the 5000 s window and none after it, and the
  predicted energy that the step's EoS pass reads;
* InternalEnergy (src/interpolation.h:107-335): exact inverse on a table whose
  entropy is linear in (rho, u), and the piecewise-linear inverse inside the
  right bracket on a curved table;
* analytic Tillotson (src/EoS.h:91-113): its limits at rho0, under cold
  compression and for a hot dilute gas, continuity of the regime blend, and
  c^2 against central differences of P.
"""
import math

import numpy as np
import pytest

from fdps_sph_amd import eos_tables as E
from fdps_sph_amd import ic as IC

G = 6.67e-11


def _lattice(k, dx=1.0):
    a = np.arange(k) - (k - 1) / 2
    return np.stack(np.meshgrid(a, a, a, indexing="ij"), -1).reshape(-1, 3) * dx


def _hydro(oracle, pos, vel, h, rho, pres, cs, gs, bal, m):
    n = len(pos)
    epi = np.zeros(n, dtype=oracle.EPI_HYDRO)
    epj = np.zeros(n, dtype=oracle.EPJ_HYDRO)
    for a in (epi, epj):
        a["pos"] = pos
        a["vel"] = vel
        a["smth"] = h
        a["dens"] = rho
        a["pres"] = pres
        a["snds"] = cs
        a["grad_smth"] = gs
        a["Bal"] = bal
        a["AVa"] = 1.0
    epj["mass"] = m
    return epi, epj


def _moment(oracle, h, p):
    """int r^p W(r, h) d^3r by Gauss-Legendre quadrature of the pinned W."""
    H = 2.5 * h
    x, w = np.polynomial.legendre.leggauss(200)
    r = 0.5 * H * (x + 1)
    vals = np.array([oracle.W([ri, 0, 0], h) for ri in r])
    return 4 * math.pi * 0.5 * H * np.sum(w * vals * r ** (2 + p))


# ------------------------------------------------------------------ hydro: energy identity
@pytest.mark.parametrize("seed", [1, 2])
def test_hydro_energy_identity(oracle, seed):
    rng = np.random.default_rng(seed)
    pos = _lattice(5) + 0.3 * (rng.random((125, 3)) - 0.5)
    n = len(pos)
    m = 1 + 0.2 * rng.random(n)
    vel = rng.normal(size=(n, 3))
    epi, epj = _hydro(oracle, pos, vel, np.full(n, 1.1), 1 + 0.3 * rng.random(n), 1 + rng.random(n),
                      1 + rng.random(n), np.ones(n), rng.random(n), m)
    r = oracle.calc_hydro(epi, epj)
    kin = m * np.einsum("ij,ij->i", vel, r["acc"])
    heat = m * r["eng_dot"]
    # both exchanges are non-trivial, and they cancel to rounding
    assert np.abs(heat).sum() > 0.1 * np.abs(kin).sum()
    assert abs(kin.sum() + heat.sum()) <= 1e-13 * (np.abs(kin).sum() + np.abs(heat).sum())
    # viscosity alone (P = 0) conserves energy as well, at any h and f: its
    # symmetrised gradient enters the acceleration and du/dt alike
    h = 1.0 + 0.4 * rng.random(n)
    f = 0.7 + 0.6 * rng.random(n)
    epi, epj = _hydro(oracle, pos, vel, h, 1 + 0.3 * rng.random(n), np.zeros(n), 1 + rng.random(n), f,
                      0.2 + rng.random(n), m)
    r = oracle.calc_hydro(epi, epj)
    kin = m * np.einsum("ij,ij->i", vel, r["acc"])
    heat = m * r["eng_dot"]
    assert np.all(heat >= -1e-15 * np.abs(heat).max())  # viscosity only heats
    assert abs(kin.sum() + heat.sum()) <= 1e-13 * (np.abs(kin).sum() + np.abs(heat).sum())


# ------------------------------------------------------------------ hydro: viscosity scale
def test_hydro_viscous_heating_continuum_limit(oracle):
    """Homologous compression v = -eps x of a unit lattice (m = 1, number
    density 1) around a centre particle with its own density and Balsara
    value: du/dt -> (2 c eps^2 M1 + 15/4 eps^3 M2) (B_i + B_j)/2 / rho_bar,
    rho_bar = (rho_i + rho_j)/2."""
    k = 17
    pos = _lattice(k)
    n = len(pos)
    c = n // 2
    # h = 2.2 lattice spacings: the lattice sums reach the integrals to ~1e-4
    eps, cs, h = 0.05, 1.0, 2.2
    rho = np.full(n, 1.0)
    rho[c] = 2.0
    bal = np.full(n, 0.8)
    bal[c] = 0.4
    epi, epj = _hydro(oracle, pos, -eps * pos, np.full(n, h), rho, np.zeros(n), np.full(n, cs), np.ones(n), bal,
                      np.ones(n))
    r = oracle.calc_hydro(epi[c:c + 1], epj)[0]
    M1, M2 = _moment(oracle, h, 1), _moment(oracle, h, 2)
    expect = (2 * cs * eps ** 2 * M1 + 3.75 * eps ** 3 * M2) * 0.5 * (0.4 + 0.8) / (0.5 * (2.0 + 1.0))
    assert abs(r["eng_dot"] - expect) <= 2e-3 * expect, (r["eng_dot"], expect)
    assert np.abs(r["acc"]).max() <= 1e-12 * expect  # no net force at the centre of symmetry


# ------------------------------------------------------------------ hydro: i/j roles
def _cloud(rng, n=40):
    pos = rng.normal(size=(n, 3)) * 1.2
    return {
        "pos": pos, "h": 1.0 + 0.8 * rng.random(n), "rho": 1 + rng.random(n), "pres": 1 + rng.random(n),
        "cs": 1 + rng.random(n), "f": 0.6 + 0.8 * rng.random(n), "m": 1 + rng.random(n),
    }


def _kernel_sum(oracle, x, pos, h, w):
    """sum_j w_j W(x - x_j, h_j) with the pinned kernel."""
    return sum(w[j] * oracle.W(x - pos[j], h[j]) for j in range(len(pos)))


def test_hydro_pressure_force_is_kernel_gradient(oracle):
    rng = np.random.default_rng(11)
    c = _cloud(rng)
    n = len(c["m"])
    A = c["f"] * c["pres"] / c["rho"] ** 2
    epi, epj = _hydro(oracle, c["pos"], np.zeros((n, 3)), c["h"], c["rho"], c["pres"], c["cs"], c["f"],
                      np.ones(n), c["m"])
    res = oracle.calc_hydro(epi, epj)
    checked = 0
    for i in range(0, n, 3):
        x = c["pos"][i]
        wi = c["m"] * A[i]
        wj = c["m"] * A

        def phi(y):
            return (sum(wi[j] * oracle.W(y - c["pos"][j], c["h"][i]) for j in range(n))
                    + _kernel_sum(oracle, y, c["pos"], c["h"], wj))
        d = 1e-4
        grad = np.array([(phi(x + d * e) - phi(x - d * e)) / (2 * d) for e in np.eye(3)])
        scale = np.abs(res["acc"][i]).max()
        if scale < 1e-3:
            continue
        checked += 1
        assert np.allclose(res["acc"][i], -grad, rtol=0, atol=2e-6 * scale + 1e-9), i
    assert checked >= 8


def test_hydro_energy_rate_is_lagrangian_derivative(oracle):
    rng = np.random.default_rng(12)
    c = _cloud(rng)
    n = len(c["m"])
    vel = rng.normal(size=(n, 3))
    A = c["f"] * c["pres"] / c["rho"] ** 2
    # Bal = 0: no viscosity
    epi, epj = _hydro(oracle, c["pos"], vel, c["h"], c["rho"], c["pres"], c["cs"], c["f"], np.zeros(n), c["m"])
    res = oracle.calc_hydro(epi, epj)
    for i in range(0, n, 4):
        def psi(t):
            p = c["pos"] + t * vel
            y = p[i]
            own = sum(c["m"][j] * oracle.W(y - p[j], c["h"][i]) for j in range(n))
            other = _kernel_sum(oracle, y, p, c["h"], c["m"] * c["f"])
            return c["f"][i] * own + other
        d = 1e-4
        rate = (psi(d) - psi(-d)) / (2 * d)
        expect = 0.5 * A[i] * rate
        assert abs(res["eng_dot"][i] - expect) <= 2e-6 * abs(A[i]) * np.abs(c["m"]).sum() + 1e-12, i


# ------------------------------------------------------------------ hydro: time step
def test_hydro_head_on_pair_time_step(oracle):
    d, V, h = 1.5, 0.7, 1.2
    ci, cj = 1.3, 0.8
    pos = np.array([[0.0, 0, 0], [d, 0, 0]])
    for sign, expect in ((1.0, 0.6 * h / (ci + cj + 3 * 2 * V)),  # approaching: w = -2V
                         (-1.0, 0.6 * h / max(2 * ci, ci + cj))):  # receding: w = 0
        vel = np.array([[sign * V, 0, 0], [-sign * V, 0, 0]])
        epi, epj = _hydro(oracle, pos, vel, np.full(2, h), np.ones(2), np.ones(2), np.array([ci, cj]),
                          np.ones(2), np.ones(2), np.ones(2))
        r = oracle.calc_hydro(epi[:1], epj)[0]
        assert abs(r["dt"] - expect) <= 1e-15 * expect, (sign, r["dt"], expect)
    # oblique approach: only the component along the separation counts
    vel = np.array([[V, 2.0, 0], [-V, -1.0, 0.5]])
    epi, epj = _hydro(oracle, pos, vel, np.full(2, h), np.ones(2), np.ones(2), np.array([ci, cj]),
                      np.ones(2), np.ones(2), np.ones(2))
    r = oracle.calc_hydro(epi[:1], epj)[0]
    assert abs(r["dt"] - 0.6 * h / (ci + cj + 6 * V)) <= 1e-15 * r["dt"]


# ------------------------------------------------------------------ derivative: grad-h factor
def test_grad_smth_matches_density_h_derivative(oracle):
    rng = np.random.default_rng(13)
    pos = _lattice(6) + 0.2 * (rng.random((216, 3)) - 0.5)
    m = 100.0 * (1 + 0.3 * rng.random(len(pos)))  # densities far above the floor of 5
    epj_d = np.zeros(len(pos), dtype=oracle.EPJ_DENS)
    epj_d["pos"] = pos
    epj_d["mass"] = m
    epj_v = np.zeros(len(pos), dtype=oracle.EPJ_DRVT)
    epj_v["pos"] = pos
    epj_v["mass"] = m

    def dens(i, h):
        e = np.zeros(1, dtype=oracle.EPI_DENS)
        e["pos"] = pos[i]
        e["mass"] = m[i]
        e["smth"] = h
        return oracle.calc_density(e, epj_d)["dens"][0]
    seen = []
    for i in (0, 5, 40, 107, 130):  # corner, edge and interior particles
        for h in (0.9, 1.3):
            rho = dens(i, h)
            assert rho > 5.0
            dh = 1e-5 * h
            drho = (dens(i, h + dh) - dens(i, h - dh)) / (2 * dh)
            expect = 1.0 / (1.0 + h * drho / (3.0 * rho))
            e = np.zeros(1, dtype=oracle.EPI_DRVT)
            e["pos"] = pos[i]
            e["smth"] = h
            e["dens"] = rho
            f = oracle.calc_derivative(e, epj_v)["grad_smth"][0]
            assert abs(f - expect) <= 1e-6 * abs(expect), (i, h, f, expect)
            seen.append(f)
    assert max(abs(np.array(seen) - 1)) > 0.1  # cases far from f = 1


# ------------------------------------------------------------------ integrator
def _isolated_state(oracle, pos, vel, m, rho0=5.0, u=1.0e6):
    """rho0 = 5: an isolated light particle's density is the floor of 5
    (src/force.h:18-20), so its h = 1.2 (m/5)^(1/3) is a fixed point of the
    density passes from the start."""
    n = len(m)
    s = oracle.State(n)
    s.x[:], s.y[:], s.z[:] = np.asarray(pos, float).T
    s.vx[:], s.vy[:], s.vz[:] = np.asarray(vel, float).T
    s.m[:] = m
    s.rho[:] = rho0
    s.u[:] = u
    return s


def _flat(oracle):
    """A table with P = 0 and c_s = 1000 m/s everywhere: isolated particles
    then feel no pressure and take dt_i = 0.6 h_i / 2000 s."""
    R, U = np.meshgrid(np.geomspace(1e-3, 1e5, 16), np.geomspace(1.0, 1e9, 16), indexing="ij")
    cells = np.zeros((16, 16, 6))
    cells[:, :, 0], cells[:, :, 1], cells[:, :, 4] = R, U, 1000.0
    return [oracle.Table(cells=cells)]


def _by_id(s):
    o = np.argsort(s.id)
    return {k: getattr(s, k)[o] for k in ("x", "y", "z", "vx", "vy", "vz", "u", "p", "cs", "rho", "du", "dti")}


def test_kdk_free_drift(oracle):
    """Light particles far apart: no SPH neighbour, gravity below 1e-20 of
    the velocities; one step is x += dt v, v and u unchanged."""
    pos = [[0.0, 0, 0], [3e6, 0, 0], [0, 5e6, 1e6]]
    vel = [[1.0, -2.0, 0.5], [0.0, 3.0, -1.0], [-4.0, 0.0, 2.0]]
    s = _isolated_state(oracle, pos, vel, [1.0, 1.0, 1.0])
    sim = oracle.Sim(s, _flat(oracle), oracle.params(ext_force_until=0.0))
    sim.init()
    b = _by_id(s)
    dt = sim.dt[0]
    assert 0 < dt < 1e29
    sim.step()
    a = _by_id(s)
    for k, q in (("x", "vx"), ("y", "vy"), ("z", "vz")):
        assert np.allclose(a[k], b[k] + dt * b[q], rtol=0, atol=1e-9 * dt)
        assert np.allclose(a[q], b[q], rtol=0, atol=1e-15)
    assert np.array_equal(a["u"], b["u"])
    assert abs(sim.time[0] - dt) <= 1e-15 * dt


def test_kdk_constant_acceleration(oracle):
    """A light particle at rest 1e9 m from a heavy one: a = G M / d^2 is
    constant over a step to 1e-12, so KDK gives v = a dt, x = a dt^2 / 2."""
    M, d = 1.0e24, 1.0e9
    s = _isolated_state(oracle, [[0.0, 0, 0], [d, 0, 0]], [[0.0, 0, 0], [0.0, 0, 0]], [1.0, M])
    sim = oracle.Sim(s, _flat(oracle), oracle.params(ext_force_until=0.0))
    sim.init()
    dt = sim.dt[0]
    a = G * M / d ** 2
    sim.step()
    o = np.argsort(s.id)
    assert abs(s.vx[o][0] - a * dt) <= 1e-9 * a * dt
    assert abs(s.x[o][0] - 0.5 * a * dt * dt) <= 1e-9 * a * dt * dt
    assert s.vy[o][0] == 0.0 and s.vz[o][0] == 0.0


@pytest.mark.parametrize("t0,damped", [(0.0, True), (4000.0, True), (5000.0, False)])
def test_kdk_damping_window(oracle, t0, damped):
    """Isolated particles with the damping force a = -0.05 v / dt_i inside
    the window (src/GI.h:458-465): the force of the initial evaluation (t = 0) damps the first
    half kick, v_half = 0.975 v0, x1 = x0 + dt v_half; the step's force acts
    on the predicted velocity 0.95 v0 if t0 + dt < 5000 s, giving
    v1 = (0.975 - 0.025 * 0.95) v0, else v1 = v_half."""
    # one particle: its list holds only itself, so dt_i = 0.6 h / (2 c) stays
    # fixed (a second particle would enter v_sig_max through its group's list)
    s = _isolated_state(oracle, [[1.0e3, -2.0e3, 5.0e2]], [[10.0, -20.0, 5.0]], [1.0])
    sim = oracle.Sim(s, _flat(oracle), oracle.params(ext_force_until=5000.0))
    sim.init()
    b = _by_id(s)
    dt = sim.dt[0]
    assert abs(dt - 0.6 * s.h[0] / 2000.0) <= 1e-15 * dt
    sim.time[:] = t0
    sim.step()
    a = _by_id(s)
    vh = 0.975
    f = vh - 0.025 * 0.95 if damped else vh
    for k, q in (("x", "vx"), ("y", "vy"), ("z", "vz")):
        assert np.allclose(a[q], f * b[q], rtol=1e-12, atol=0)
        assert np.allclose(a[k], b[k] + dt * vh * b[q], rtol=0, atol=1e-9 * dt * np.abs(b[q]).max())


def test_step_eos_reads_predicted_energy(oracle):
    """The step's EoS pass evaluates P and c_s at the predicted energy
    u0 + dt du0 (src/class.h:215-219, src/main.cpp:141), while the state ends
    with the final-kick energy."""
    ic = IC.granite_impact(450, 50, seed=21)
    s = _isolated_state(oracle, ic["pos"], ic["vel"], ic["mass"], rho0=2680.0)
    s.rho[:] = ic["dens"]
    s.mat[:] = 0
    tab = oracle.Table(E.PATHS["granite"])
    sim = oracle.Sim(s, [tab], oracle.params())
    sim.init()
    b = _by_id(s)
    dt = sim.dt[0]
    assert np.abs(b["du"]).max() > 0
    sim.step()
    a = _by_id(s)
    u_pred = b["u"] + dt * b["du"]
    p = np.array([tab.pressure(r, u) for r, u in zip(a["rho"], u_pred)])
    c = np.array([tab.sound_speed(r, u) for r, u in zip(a["rho"], u_pred)])
    assert np.array_equal(a["p"], p)
    assert np.array_equal(a["cs"], c)
    assert not np.array_equal(a["u"], u_pred)


# ------------------------------------------------------------------ InternalEnergy
def _table_cells(rho_axis, u_axis, S):
    nr, nu = len(rho_axis), len(u_axis)
    R, U = np.meshgrid(rho_axis, u_axis, indexing="ij")
    cells = np.zeros((nr, nu, 6))
    cells[:, :, 0] = R
    cells[:, :, 1] = U
    cells[:, :, 2] = 1.0
    cells[:, :, 3] = 1.0
    cells[:, :, 4] = 1.0
    cells[:, :, 5] = S(R, U)
    return cells


def test_internal_energy_linear_table_is_exact(oracle):
    """S = a + b rho + c u on a 120 x 120 table: the two row-wise linear
    inversions and the density interpolation reproduce u = (S - a - b rho)/c."""
    rax = np.exp(np.linspace(np.log(100.0), np.log(20000.0), 120))
    uax = np.exp(np.linspace(np.log(1e4), np.log(1e8), 120))
    a_, b_, c_ = 100.0, 0.05, 1e-3
    t = oracle.Table(cells=_table_cells(rax, uax, lambda R, U: a_ + b_ * R + c_ * U))
    rng = np.random.default_rng(14)
    for _ in range(200):
        r = rng.integers(2, 116)
        rho = rax[r] + rng.uniform(0.05, 0.95) * (rax[r + 1] - rax[r])
        u = math.exp(rng.uniform(math.log(uax[3]), math.log(uax[-4])))
        S = a_ + b_ * rho + c_ * u
        for abs_int in (True, False):
            got = t.internal_energy(rho, S, abs_int)
            assert abs(got - u) <= 1e-9 * u, (rho, u, got, abs_int)


def test_internal_energy_brackets_on_curved_table(oracle):
    """S = sqrt(u) scaled, the same in every row: the result is the chord
    inverse inside the bracket S_k <= S <= S_k+1 of the row's entropies."""
    rax = np.exp(np.linspace(np.log(100.0), np.log(20000.0), 120))
    uax = np.exp(np.linspace(np.log(1e4), np.log(1e8), 120))
    t = oracle.Table(cells=_table_cells(rax, uax, lambda R, U: 10.0 * np.sqrt(U)))
    Srow = 10.0 * np.sqrt(uax)
    rng = np.random.default_rng(15)
    for _ in range(200):
        r = rng.integers(2, 116)
        rho = rax[r] + rng.uniform(0.05, 0.95) * (rax[r + 1] - rax[r])
        S = rng.uniform(Srow[3], Srow[-4])
        k = int(np.searchsorted(Srow, S)) - 1
        expect = uax[k] + (S - Srow[k]) / (Srow[k + 1] - Srow[k]) * (uax[k + 1] - uax[k])
        for abs_int in (True, False):
            got = t.internal_energy(rho, S, abs_int)
            assert abs(got - expect) <= 1e-9 * expect, (rho, S, got, expect, abs_int)


# ------------------------------------------------------------------ analytic Tillotson
GRANITE = E.GRANITE_TILLOTSON  # rho0, u0, u_iv, u_cv, A, B, a, b, alpha, beta (src/EoS.h:446)


def test_tillotson_closed_forms(oracle):
    """Limits the Tillotson form fixes: at rho0 P = (a + b/(u/u0 + 1)) rho0 u;
    cold compression (u = 0) P = A mu + B mu^2; a hot dilute gas P = a rho u
    (both exponentials vanish); the blend between u_iv and u_cv meets each
    branch at its end."""
    rho0, u0, uiv, ucv, A, B, a, b, alpha, beta = GRANITE
    t = oracle.Table(tillotson=GRANITE)
    for u in (1e3, 1e6, 3e7):
        expect = (a + b / (u / u0 + 1.0)) * rho0 * u
        assert abs(t.pressure(rho0, u) - expect) <= 1e-15 * expect
    for eta in (1.05, 1.3, 2.0):
        mu = eta - 1.0
        assert abs(t.pressure(eta * rho0, 0.0) - (A * mu + B * mu * mu)) <= 1e-15 * (A * mu + B * mu * mu)
        # compressed states keep that cold part at any u, hot ones included:
        # the thermal part is (a + b / (u / (u0 eta^2) + 1)) rho u
        for u in (1e6, 3e7, 1e8):
            th = (a + b / (u / u0 / eta ** 2 + 1.0)) * eta * rho0 * u
            got = t.pressure(eta * rho0, u) - (A * mu + B * mu * mu)
            assert abs(got - th) <= 1e-14 * (th + A * mu), (eta, u, got, th)
    rho, u = 1e-3 * rho0, 2.0 * ucv
    assert t.pressure(rho, u) == a * rho * u
    # in expansion, P_co below u_iv and P_ex above u_cv, continuous across both
    rho = 0.8 * rho0
    for edge in (uiv, ucv):
        lo, hi = t.pressure(rho, edge * (1 - 1e-12)), t.pressure(rho, edge * (1 + 1e-12))
        assert abs(hi - lo) <= 1e-9 * abs(lo), (edge, lo, hi)


@pytest.mark.parametrize("rho_f,u", [(1.3, 1e6), (1.0, 5e6), (0.8, 1e6), (0.8, 1e7), (0.8, 4e7), (0.5, 2.5e7),
                                     (0.2, 6e7)])
def test_tillotson_sound_speed_is_thermodynamic(oracle, rho_f, u):
    """c^2 = dP/drho|_u + P/rho^2 dP/du|_rho against central differences of
    the pressure itself, in every regime."""
    rho0 = GRANITE[0]
    t = oracle.Table(tillotson=GRANITE)
    rho = rho_f * rho0
    P = t.pressure(rho, u)
    dr = (t.pressure(rho * (1 + 1e-6), u) - t.pressure(rho * (1 - 1e-6), u)) / (2e-6 * rho)
    du = (t.pressure(rho, u * (1 + 1e-6)) - t.pressure(rho, u * (1 - 1e-6))) / (2e-6 * u)
    c2 = dr + P / rho ** 2 * du
    assert c2 > 0
    assert abs(t.sound_speed(rho, u) ** 2 - c2) <= 1e-6 * c2, (t.sound_speed(rho, u) ** 2, c2)
    assert math.isnan(t.temperature(rho, u)) and math.isnan(t.internal_energy(rho, 100.0))
