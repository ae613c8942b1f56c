"""Checks that pin the oracle independently of the reference build: closed
forms of the limiter, the EoS and the primitive recovery, translational
invariance, and 0+1D Bjorken flow against an independent ODE solution."""
import math
import struct

import numpy as np
import pytest

import vhtest as V


def _L():
    return V.oracle_lib()


def test_minmod_truth_table():
    """sign(0) = +1 and minmod = (sx + sy) min(|x|,|y|) / 2 (FluxLimiter.cpp:14-32)."""
    L = _L()
    cases = {(1.0, 2.0): 1.0, (-1.0, -2.0): -1.0, (1.0, -2.0): 0.0, (-3.0, 2.0): 0.0, (0.0, 5.0): 0.0,
             (0.0, -5.0): 0.0, (2.5, 2.5): 2.5, (-0.0, 3.0): 0.0}
    for (x, y), want in cases.items():
        got = L.orc_minmod(x, y)
        assert got == want and math.copysign(1, got) == math.copysign(1, want + 0.0), (x, y, got)


def test_limited_slope_closed_forms():
    L = _L()
    h = 0.37
    # linear data: the centred slope (theta = 1.1 > 1 never limits it)
    assert L.orc_approx_deriv(1.0, 1.0 + h, 1.0 + 2 * h) == pytest.approx(h, rel=1e-15)
    # extremum: zero
    assert L.orc_approx_deriv(0.0, 1.0, 0.5) == 0.0
    # steep right side: limited by theta * left difference
    assert L.orc_approx_deriv(0.0, 1.0, 4.0) == 1.1
    # steep left side: limited by theta * right difference
    assert L.orc_approx_deriv(0.0, 3.0, 3.5) == pytest.approx(0.55, rel=1e-15)


def test_conformal_eos_closed_forms():
    L = _L()
    E = L.orc_eos_create(None)
    rng = np.random.default_rng(0)
    for e in rng.uniform(1e-6, 500, 200):
        assert L.orc_eos_pressure(E, e) == e / 3.0
        assert L.orc_eos_cs2(E, e) == 1.0 / 3.0
        T = L.orc_eos_temperature(E, e)
        assert T == pytest.approx((e / 13.8997) ** 0.25, rel=2e-7)
    for T in (0.3, 0.76, 1.5):
        assert L.orc_eos_energy_density(E, T) == pytest.approx(13.8997 * T ** 4, rel=5e-7)
    L.orc_eos_destroy(E)


def test_table_eos_nodes_and_constant_sound_speed():
    """Interpolation returns the table at its nodes; the reference's table c_s^2
    is a one-segment difference quotient clamped to [0.01, 1/3], which is 1/3
    for every e on this table (EquationOfState.cpp:141-153)."""
    L = _L()
    tab = V.eos_table()
    raw = tab.reshape(-1, 4)
    E = L.orc_eos_create(tab.ctypes.data)
    e_nodes = raw[:, 0] / V.HBARC
    p_nodes = raw[:, 1] / V.HBARC
    T_nodes = raw[:, 3] / V.HBARC
    for i in (0, 1, 17, 5000, 54321, 99998):
        assert L.orc_eos_pressure(E, e_nodes[i]) == pytest.approx(max(1e-15, p_nodes[i]), rel=1e-12)
        assert L.orc_eos_temperature(E, e_nodes[i]) == pytest.approx(max(1e-15, T_nodes[i]), rel=1e-12)
    for e in np.geomspace(1e-6, 2 * e_nodes[-1], 300):
        assert L.orc_eos_cs2(E, e) == 1.0 / 3.0
    # T(e(T)) = T within the bisection tolerance (temperatures inside the table)
    for T in (0.12 / V.HBARC, 0.15 / V.HBARC, 2.0):
        e = L.orc_eos_energy_density(E, T)
        assert L.orc_eos_temperature(E, e) == pytest.approx(T, rel=1e-7)
    L.orc_eos_destroy(E)


def _conserved(e, p, ut, ux, uy, un, pi):
    """T^{tau mu} + pi^{tau mu} from primitives (EnergyMomentumTensor.cpp:292-306)."""
    q = np.zeros(15)
    q[0] = (e + p) * ut * ut - p + pi[0]
    q[1] = (e + p) * ut * ux + pi[1]
    q[2] = (e + p) * ut * uy + pi[2]
    q[3] = (e + p) * ut * un + pi[3]
    q[4:14] = pi
    return q


@pytest.mark.parametrize("t", [0.6, 3.0])
def test_recovery_round_trip_conformal(t):
    """(e, u) -> T^{tau mu} -> recovery returns (e, u) to single precision:
    the closed-form root passes through sqrtf (EnergyMomentumTensor.cpp:113)."""
    L = _L()
    o = V.Oracle(3, 3, 1, 0.1, 0.1, 0.1, 0.02, t, 0.2)
    rng = np.random.default_rng(4)
    out = np.zeros(6)
    for _ in range(200):
        e = rng.uniform(0.2, 80)
        ux, uy, un = rng.normal(0, 1.5), rng.normal(0, 1.5), rng.normal(0, 0.5 / t)
        ut = math.sqrt(1 + ux * ux + uy * uy + t * t * un * un)
        q = _conserved(e, e / 3, ut, ux, uy, un, np.zeros(10))
        L.orc_inferred(o.h, t, q.ctypes.data, e, out.ctypes.data)
        assert out[0] == pytest.approx(e, rel=3e-6)
        assert out[2] == pytest.approx(ut, rel=3e-6)
        assert out[3] == pytest.approx(ux, rel=1e-5, abs=1e-6)
        assert out[5] == pytest.approx(un, rel=1e-5, abs=1e-6)


def test_recovery_round_trip_table():
    """Table EoS: Newton-Raphson from a nearby previous e converges to the
    state's e within the solver's 1e-3 stopping rule (EnergyMomentumTensor.cpp:20,44)."""
    L = _L()
    tab = V.eos_table()
    o = V.Oracle(3, 3, 1, 0.1, 0.1, 0.1, 0.02, 1.0, 0.2, table=tab)
    E = L.orc_eos_create(tab.ctypes.data)
    rng = np.random.default_rng(5)
    out = np.zeros(6)
    for _ in range(100):
        e = rng.uniform(0.5, 60)
        p = L.orc_eos_pressure(E, e)
        ux, uy = rng.normal(0, 1.0), rng.normal(0, 1.0)
        ut = math.sqrt(1 + ux * ux + uy * uy)
        q = _conserved(e, p, ut, ux, uy, 0.0, np.zeros(10))
        L.orc_inferred(o.h, 1.0, q.ctypes.data, e * rng.uniform(0.9, 1.1), out.ctypes.data)
        assert out[0] == pytest.approx(e, rel=2e-3)
    L.orc_eos_destroy(E)


def test_uniform_state_stays_uniform():
    """Every cell of a uniform lattice sees the same stencil (clamped edges are
    copies), so the state stays uniform bit for bit."""
    nx, ny, nz, t0, dt = 7, 6, 5, 0.8, 0.02
    o = V.Oracle(nx, ny, nz, 0.1, 0.1, 0.1, dt, t0, 0.2)
    N = nx * ny * nz
    pi = np.zeros((10, N))
    pi[4] = pi[7] = 0.05  # pixx = piyy
    pi[9] = -0.1 / t0 ** 2  # pinn
    o.set_ic({"e": np.full(N, 20.0), "p": np.full(N, 20.0 / 3), "ut": np.ones(N), "ux": np.zeros(N),
              "uy": np.zeros(N), "un": np.zeros(N), "pi": pi.ravel()})
    o.init_conserved(t0)
    for n in range(1, 6):
        o.step(t0 + (n - 1) * dt)
    for f in ("e", "pixx", "pinn", "ttt", "ut"):
        a = o.get(f)
        assert np.all(a == a[0]), f


def test_bjorken_flow_against_ode():
    """0+1D Bjorken flow: uniform lattice, u = (1,0,0,0).  With
    Phi = -tau^2 pi^{eta eta} and pi^xx = pi^yy = Phi/2 the reference's
    equations reduce to de/dtau = -(e + p - Phi)/tau and
    dPhi/dtau = 4 beta_pi/(3 tau) - Phi/tau_pi - (delta_pipi + tau_pipi/3) Phi/tau
    (beta_pi = (e+p)/5, tau_pi = 5 eta/s / T; the notebook's Bjorken system,
    mathematica/gubserFlowTest.nb).  The lattice evolves it with Heun's method,
    so it must agree with an accurate ODE solution to O(dt^2)."""
    from scipy.integrate import solve_ivp

    etas, t0, dt, nsteps = 0.2, 0.6, 0.005, 200
    e0, phi0 = 30.0, 1.5
    d_pp, t_pp = 1.33333, 1.42857

    def rhs(tau, y):
        e, phi = y
        p = e / 3
        T = (e / 13.8997) ** 0.25
        return [-(e + p - phi) / tau,
                4 * (e + p) / 5 / (3 * tau) - phi * T / (5 * etas) - (d_pp + t_pp / 3) * phi / tau]

    sol = solve_ivp(rhs, (t0, t0 + nsteps * dt), [e0, phi0], rtol=1e-11, atol=1e-12, dense_output=True)
    nx = ny = nz = 3
    N = nx * ny * nz
    o = V.Oracle(nx, ny, nz, 0.1, 0.1, 0.1, dt, t0, etas)
    pi = np.zeros((10, N))
    pi[4] = pi[7] = phi0 / 2
    pi[9] = -phi0 / t0 ** 2
    o.set_ic({"e": np.full(N, e0), "p": np.full(N, e0 / 3), "ut": np.ones(N), "ux": np.zeros(N),
              "uy": np.zeros(N), "un": np.zeros(N), "pi": pi.ravel()})
    o.init_conserved(t0)
    for n in range(1, nsteps + 1):
        o.step(t0 + (n - 1) * dt)
    tau = t0 + nsteps * dt
    e_want, phi_want = sol.sol(tau)
    e_got = o.get("e")[0]
    phi_got = -tau ** 2 * o.get("pinn")[0]
    assert e_got == pytest.approx(e_want, rel=2e-5)
    assert phi_got == pytest.approx(phi_want, rel=2e-4)
    # pi^xx = -tau^2 pi^{eta eta} / 2 holds to the scheme's O(dt^2)
    assert o.get("pixx")[0] == pytest.approx(phi_got / 2, rel=5e-4)
    # a plausible slip (the tau_pipi/3 term dropped) is far outside that bar
    sol2 = solve_ivp(lambda tau, y: [rhs(tau, y)[0], rhs(tau, y)[1] + t_pp / 3 * y[1] / tau], (t0, tau),
                     [e0, phi0], rtol=1e-11, atol=1e-12)
    assert abs(sol2.y[1, -1] - phi_got) > 50 * abs(phi_want - phi_got)


ETAS_MIN, ETAS_TK, SLOPE_LO, SLOPE_HI = 0.08, 0.154, 0.5, 1.0


def _etas_oracle():
    return V.Oracle(3, 3, 3, 0.1, 0.1, 0.1, 0.005, 0.6, ETAS_MIN, etas_Tk_GeV=ETAS_TK, etas_slope_lo=SLOPE_LO,
                    etas_slope_hi=SLOPE_HI)


def test_eta_over_s_of_T_minimum_in_GeV():
    """eta/s(T) (the extension the drop-in configuration names: minimum 0.08
    at T = 154 MeV) scanned over T = 50...400 MeV with T handed over in
    fm^-1, as the source terms do: the minimum sits at 154 MeV to the scan's
    resolution and equals 0.08, and the two sides rise with the configured
    slopes per GeV.  A fm^-1/GeV slip would put the minimum at 30 MeV or
    30 GeV."""
    o = _etas_oracle()
    TGeV = np.linspace(0.05, 0.4, 3501)  # 0.1 MeV steps
    vals = np.array([o.etas(t / V.HBARC) for t in TGeV])
    k = int(np.argmin(vals))
    assert abs(TGeV[k] - ETAS_TK) <= 1e-4
    assert vals[k] == pytest.approx(ETAS_MIN, abs=1e-5)
    assert o.etas(ETAS_TK / V.HBARC) == pytest.approx(ETAS_MIN, abs=1e-12)
    lo, hi = TGeV < ETAS_TK - 1e-3, TGeV > ETAS_TK + 1e-3
    assert np.polyfit(TGeV[lo], vals[lo], 1)[0] == pytest.approx(-SLOPE_LO, rel=1e-9)
    assert np.polyfit(TGeV[hi], vals[hi], 1)[0] == pytest.approx(SLOPE_HI, rel=1e-9)
    # zero slopes: the reference's constant, bit for bit
    c = V.Oracle(3, 3, 3, 0.1, 0.1, 0.1, 0.005, 0.6, 0.2)
    assert all(c.etas(t / V.HBARC) == 0.2 for t in (0.05, 0.154, 0.3))


def _bjorken_ode(etas_of_TGeV, e0, phi0, t0, t1):
    from scipy.integrate import solve_ivp
    d_pp, t_pp = 1.33333, 1.42857

    def rhs(tau, y):
        e, phi = y
        p = e / 3
        T = (e / 13.8997) ** 0.25  # fm^-1
        return [-(e + p - phi) / tau,
                4 * (e + p) / 5 / (3 * tau) - phi * T / (5 * etas_of_TGeV(T * 0.197326938))
                - (d_pp + t_pp / 3) * phi / tau]

    return solve_ivp(rhs, (t0, t1), [e0, phi0], rtol=1e-11, atol=1e-12, dense_output=True)


def test_bjorken_flow_with_eta_over_s_of_T():
    """0+1D Bjorken flow (as test_bjorken_flow_against_ode) with eta/s(T):
    the fluid cools from 240 to 135 MeV, through the minimum at 154 MeV, and
    the lattice must follow an ODE solved independently with
    eta/s(T) = 0.08 + 0.5 (0.154 - T) below and 0.08 + 1.0 (T - 0.154) above
    154 MeV (T in GeV).  The same ODE with the minimum placed at 0.154 fm^-1
    (a unit slip) or with constant eta/s = 0.08 misses the lattice by far
    more than the scheme's error."""
    t0, dt, nsteps = 0.6, 0.005, 600
    e0, phi0 = 30.0, 1.0

    def etas_T(TGeV):
        return ETAS_MIN + (SLOPE_HI * (TGeV - ETAS_TK) if TGeV > ETAS_TK else SLOPE_LO * (ETAS_TK - TGeV))

    tau = t0 + nsteps * dt
    sol = _bjorken_ode(etas_T, e0, phi0, t0, tau)
    T_end = (sol.sol(tau)[0] / 13.8997) ** 0.25 * V.HBARC
    assert (e0 / 13.8997) ** 0.25 * V.HBARC > 0.2 and T_end < 0.14  # the run crosses the minimum
    N = 27
    o = _etas_oracle()
    pi = np.zeros((10, N))
    pi[4] = pi[7] = phi0 / 2
    pi[9] = -phi0 / t0 ** 2
    o.set_ic({"e": np.full(N, e0), "p": np.full(N, e0 / 3), "ut": np.ones(N), "ux": np.zeros(N),
              "uy": np.zeros(N), "un": np.zeros(N), "pi": pi.ravel()})
    o.init_conserved(t0)
    for n in range(1, nsteps + 1):
        assert o.step(t0 + (n - 1) * dt) == 0
    e_want, phi_want = sol.sol(tau)
    e_got, phi_got = o.get("e")[0], -tau ** 2 * o.get("pinn")[0]
    assert e_got == pytest.approx(e_want, rel=2e-5)
    assert phi_got == pytest.approx(phi_want, rel=5e-4)
    err = abs(phi_got - phi_want)
    slip = _bjorken_ode(lambda TGeV: etas_T(TGeV / 0.197326938), e0, phi0, t0, tau).sol(tau)[1]
    const = _bjorken_ode(lambda TGeV: ETAS_MIN, e0, phi0, t0, tau).sol(tau)[1]
    assert abs(slip - phi_got) > 20 * err
    assert abs(const - phi_got) > 20 * err
