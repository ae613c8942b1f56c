"""Checks that pin the oracle's bulk-viscosity extension: zeta/s(T) peaked at
a chosen temperature with a chosen height, the table's own sound speed
(cs2_mode 1, which makes the bulk coefficients (1/3 - c_s^2) non-zero), the
fdlibm-method exp both the oracle and the device use for zeta/s, and the
bound on the bulk relaxation rate.  The dynamics are pinned by 0+1D Bjorken
flow against an ODE solved independently from the reference's equations
(rhic/src/SourceTerms.cpp:55-313) reduced to u = (1,0,0,0)."""
import math
import struct

import numpy as np
import pytest

import vhtest as V

HB = V.HBARC


def _L():
    return V.oracle_lib()


def _ulps(a, b):
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    return abs(ia - ib)


def test_fdlibm_exp_within_one_ulp_of_exp():
    """The oracle's exp (fdlibm method) against the host C library's exp
    over the finite range, the interval zeta/s feeds it, tiny arguments and
    the reduction's branch edges: never more than one ulp apart, and exact
    at the special values."""
    L = _L()
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(-745, 709, 40000), rng.uniform(-400, 10, 40000),
                         rng.uniform(-1e-6, 1e-6, 5000), rng.uniform(-3, 3, 20000)])
    edges = [0.5 * math.log(2), 1.5 * math.log(2), 2.0 ** -28, 709.78, -745.13]
    for x0 in edges:
        xs = np.concatenate([xs, np.nextafter(x0, np.inf) + np.arange(-50, 50) * math.ulp(x0),
                             -(x0 + np.arange(-50, 50) * math.ulp(x0))])
    worst = 0
    for x in xs:
        x = float(x)
        want, got = math.exp(x) if x < 709.78 else math.exp(min(x, 709.78)), L.orc_exp_fdlibm(x)
        if x >= 709.78:
            continue
        worst = max(worst, _ulps(want, got))
    assert worst <= 1
    assert L.orc_exp_fdlibm(0.0) == 1.0
    assert L.orc_exp_fdlibm(1.0) == pytest.approx(math.e, rel=2.3e-16)
    assert L.orc_exp_fdlibm(-math.inf) == 0.0
    assert L.orc_exp_fdlibm(math.inf) == math.inf
    assert L.orc_exp_fdlibm(800.0) == math.inf
    assert L.orc_exp_fdlibm(-800.0) == 0.0
    assert math.isnan(L.orc_exp_fdlibm(math.nan))


def _zeta_ref_shape(x):
    """The reference's bulkViscosityToEntropyDensity of its argument x
    (rhic/src/SourceTerms.cpp:40-48), written out with math.exp."""
    if x > 1.05:
        return 0.9 * math.exp(-(x - 1) / 0.025) + 0.25 * math.exp(-(x - 1) / 0.13) + 0.001
    if x < 0.995:
        return 0.9 * math.exp((x - 1) / 0.0025) + 0.22 * math.exp((x - 1) / 0.022) + 0.03
    return -13.77 * x * x + 27.55 * x - 13.45


def test_zeta_over_s_peak_in_GeV():
    """zeta/s(T) with T_peak = 160 MeV and height 0.1, T handed over in
    fm^-1 as the source terms do: the maximum of a 0.1 MeV scan sits at
    160 MeV and equals 0.1 (the reference's quadratic peaks at its argument
    27.55/27.54 = 1.00036 with 0.3300018: 160.06 MeV, 0.1000005); the tails follow the reference's shape scaled by
    0.1/0.33.  Without the extension the function is the reference's (to an
    ulp of exp).  A fm^-1/GeV slip would put the peak at 32 MeV or 810 MeV."""
    L = _L()
    TGeV = np.linspace(0.05, 0.5, 4501)
    vals = np.array([L.orc_zeta_over_s(t / HB, 0.16, 0.1) for t in TGeV])
    k = int(np.argmax(vals))
    assert abs(TGeV[k] - 0.16 * 27.55 / 27.54) <= 0.6e-4
    assert vals[k] == pytest.approx(0.1, rel=1e-5)
    assert L.orc_zeta_over_s(0.16 / HB, 0.16, 0.1) == pytest.approx(0.1, rel=1e-12)
    for t in (0.06, 0.12, 0.155, 0.165, 0.2, 0.45):
        assert L.orc_zeta_over_s(t / HB, 0.16, 0.1) == pytest.approx(0.1 / 0.33 * _zeta_ref_shape(t / 0.16),
                                                                   rel=1e-13)
        x = (t / HB) / 1.01355
        assert L.orc_zeta_over_s(t / HB, 0.0, 0.0) == pytest.approx(_zeta_ref_shape(x), rel=1e-14)
    # peak placed at the reference's own 1.01355 fm^-1 with height 0.33: the reference's function
    for T in (0.5, 0.9, 1.0, 1.01355, 1.2, 2.0):
        assert L.orc_zeta_over_s(T, 1.01355 * HB, 0.33) == pytest.approx(L.orc_zeta_over_s(T, 0.0, 0.0), rel=1e-13)
    assert L.orc_zeta_over_s(0.8, 0.16, 0.0) == 0.0


def test_table_sound_speed_mode():
    """cs2_mode 1 is the table's sound speed dp/de over one table segment
    (at a node: (p[i+1] - p[i]) / (e[i+1] - e[i]) from the raw rows),
    clamped to [0.01, 1/3].  Its minimum over T = 100...400 MeV lies in the
    crossover (lattice QCD's softest point, 140-170 MeV), and it tends to
    ~0.3 at high T.  Mode 0 stays the reference's constant 1/3."""
    L = _L()
    tab = V.eos_table()
    raw = tab.reshape(-1, 4)
    E = L.orc_eos_create(tab.ctypes.data)
    e_nodes, p_nodes = raw[:, 0] / HB, raw[:, 1] / HB
    L.orc_eos_set_cs2_mode(E, 1)
    de = e_nodes[1] - e_nodes[0]
    for i in (3, 100, 5000, 50000, 99000):
        want = min(1 / 3, max(0.01, (p_nodes[i + 1] - p_nodes[i]) / de))
        assert L.orc_eos_cs2(E, e_nodes[i]) == pytest.approx(want, rel=1e-6)
    es = np.geomspace(0.05, 400.0, 3000)
    cs2 = np.array([L.orc_eos_cs2(E, e) for e in es])
    T = np.array([L.orc_eos_temperature(E, e) * HB for e in es])
    assert np.all((cs2 >= 0.01) & (cs2 <= 1 / 3))
    window = (T > 0.1) & (T < 0.4)
    kmin = int(np.argmin(np.where(window, cs2, 1.0)))
    assert 0.14 <= T[kmin] <= 0.17
    assert 0.12 <= cs2[kmin] <= 0.17
    assert 0.28 <= cs2[-1] < 1 / 3
    L.orc_eos_set_cs2_mode(E, 0)
    assert all(L.orc_eos_cs2(E, e) == 1 / 3 for e in es[::50])
    L.orc_eos_destroy(E)


def _bulk_ode(E, L, zeta_of_TGeV, e0, phi0, Pi0, t0, t1, drop=None, method="RK45"):
    """The reference's equations for 0+1D Bjorken flow with bulk, in
    Phi = -tau^2 pi^{eta eta}, pi^xx = pi^yy = Phi/2 and Pi:
      de/dtau   = -(e + p + Pi - Phi)/tau
      dPhi/dtau = 4 beta_pi/(3 tau) - Phi/tau_pi - (delta_pipi + tau_pipi/3) Phi/tau
                  + (2/3) lambda_piPi Pi/tau
      dPi/dtau  = -beta_Pi/tau - Pi/tau_Pi - delta_PiPi Pi/tau + lambda_Pipi Phi/tau
    (theta = 1/tau, pi:sigma = Phi/tau, sigma^{eta eta} = -2/(3 tau^3) in
    the reference's shear tensor), beta_pi = (e+p)/5, tau_pi = 5 eta/s / T,
    beta_Pi = 15 a^2 (e+p), 1/tau_Pi = 15 a^2 T/(zeta/s), lambda_Pipi = 8a/5,
    a = 1/3 - c_s^2, with the EoS functions of the table."""
    from scipy.integrate import solve_ivp
    d_pp, t_pp, d_PP, l_pP = 1.33333, 1.42857, 0.666667, 1.2
    etas = 0.08

    def rhs(tau, y):
        e, phi, Pi = y
        p, T, cs2 = L.orc_eos_pressure(E, e), L.orc_eos_temperature(E, e), L.orc_eos_cs2(E, e)
        a = 1 / 3 - cs2
        bPi = 0.0 if drop == "beta_Pi" else 15 * a * a * (e + p)
        lPp = 8 * a / 5 * (-1 if drop == "lambda_sign" else 1)
        de = -(e + p + Pi - phi) / tau
        dphi = 4 * (e + p) / 5 / (3 * tau) - phi * T / (5 * etas) - (d_pp + t_pp / 3) * phi / tau \
            + 2 / 3 * l_pP * Pi / tau
        dPi = -bPi / tau - Pi * 15 * a * a * T / zeta_of_TGeV(T * HB) - d_PP * Pi / tau + lPp * phi / tau
        return [de, dphi, dPi]

    return solve_ivp(rhs, (t0, t1), [e0, phi0, Pi0], method=method, rtol=1e-10, atol=1e-12, dense_output=True)


def _bjorken_lattice(dt, t0, nsteps, e0, phi0, Pi0, **kw):
    tab = V.eos_table()
    o = V.Oracle(3, 3, 3, 0.1, 0.1, 0.1, dt, t0, 0.08, bulk=True, table=tab, cs2_mode=1, **kw)
    L = _L()
    E = L.orc_eos_create(tab.ctypes.data)
    N = 27
    pi = np.zeros((10, N))
    pi[4] = pi[7] = phi0 / 2
    pi[9] = -phi0 / t0 ** 2
    o.set_ic({"e": np.full(N, e0), "p": np.full(N, L.orc_eos_pressure(E, e0)), "ut": np.ones(N),
              "ux": np.zeros(N), "uy": np.zeros(N), "un": np.zeros(N), "pi": pi.ravel(), "Pi": np.full(N, Pi0)})
    o.init_conserved(t0)
    for n in range(1, nsteps + 1):
        assert o.step(t0 + (n - 1) * dt) == 0
    L.orc_eos_set_cs2_mode(E, 1)
    tau = t0 + nsteps * dt
    return o, E, L, tau


def test_bjorken_flow_with_bulk_against_ode():
    """0+1D Bjorken flow on the table EoS with its sound speed (cs2_mode 1)
    and zeta/s peaked at 200 MeV (height 0.1): the fluid cools from 255 to
    175 MeV, through the peak, and e, Phi and Pi follow the ODE solution.
    The bar for Phi and Pi (1e-3, 3e-3) is set by the reference scheme
    itself: u^tau comes from a single-precision square root
    (EnergyMomentumTensor.cpp), so it carries 1e-7..1e-6 of rounding that the
    (u - u_prev)/dt terms of theta and sigma amplify by 1/dt (the agreement
    is 1e-5 at dt = 0.005 and 1e-3 at dt = 0.0025).  The ODE with beta_Pi
    dropped, with lambda_Pipi's sign flipped, or with the peak temperature
    read in fm^-1 misses the lattice's Pi by more than ten times that bar."""
    t0, dt, nsteps, Tp = 0.6, 0.005, 680, 0.2
    e0, phi0, Pi0 = 30.0, 1.0, -0.3
    o, E, L, tau = _bjorken_lattice(dt, t0, nsteps, e0, phi0, Pi0, zeta_Tpeak_GeV=Tp, zeta_peak=0.1)

    def zeta(TGeV):
        return 0.1 / 0.33 * _zeta_ref_shape(TGeV / Tp)

    sol = _bulk_ode(E, L, zeta, e0, phi0, Pi0, t0, tau)
    e_want, phi_want, Pi_want = sol.sol(tau)
    T0, T1 = L.orc_eos_temperature(E, e0) * HB, L.orc_eos_temperature(E, e_want) * HB
    assert T0 > 0.25 and T1 < 0.18  # the run crosses the peak
    e_got, phi_got, Pi_got = o.get("e")[0], -tau ** 2 * o.get("pinn")[0], o.get("Pi")[0]
    assert e_got == pytest.approx(e_want, rel=1e-4)
    assert phi_got == pytest.approx(phi_want, rel=1e-3)
    assert Pi_got == pytest.approx(Pi_want, rel=3e-3)
    assert abs(Pi_want) > 1e-3 * e_want  # bulk pressure is not negligible here
    bar = 10 * 3e-3 * abs(Pi_want)
    for drop in ("beta_Pi", "lambda_sign"):
        wrong = _bulk_ode(E, L, zeta, e0, phi0, Pi0, t0, tau, drop=drop).sol(tau)[2]
        assert abs(wrong - Pi_got) > bar, drop
    slip = _bulk_ode(E, L, lambda TGeV: zeta(TGeV / HB), e0, phi0, Pi0, t0, tau).sol(tau)[2]
    assert abs(slip - Pi_got) > bar
    L.orc_eos_destroy(E)


def test_bulk_relaxation_rate_bound():
    """With a tiny zeta/s (height 1e-4) the relaxation rate 15 a^2 T/(zeta/s)
    exceeds 2/dt, where the explicit step diverges; the bound tau_Pi >= dt
    (beta_Pi scaled alike) keeps it stable, and Pi stays within a few percent
    of the stiff ODE's solution, which sits at the Navier-Stokes value.
    zeta/s = 0 gives Pi ~ 0 (only the lambda_Pipi coupling to the shear
    stress feeds it)."""
    t0, dt, nsteps = 0.6, 0.02, 150
    e0, phi0, Pi0 = 30.0, 1.0, 0.0
    tab = V.eos_table()
    L = _L()
    E = L.orc_eos_create(tab.ctypes.data)
    L.orc_eos_set_cs2_mode(E, 1)
    a = 1 / 3 - L.orc_eos_cs2(E, e0)
    rate = 15 * a * a * L.orc_eos_temperature(E, e0) / L.orc_zeta_over_s(L.orc_eos_temperature(E, e0), 0.16, 1e-4)
    assert rate * dt > 2  # unbounded, the explicit step would diverge
    L.orc_eos_destroy(E)
    o, E, L, tau = _bjorken_lattice(dt, t0, nsteps, e0, phi0, Pi0, zeta_Tpeak_GeV=0.16, zeta_peak=1e-4)
    sol = _bulk_ode(E, L, lambda TGeV: 1e-4 / 0.33 * _zeta_ref_shape(TGeV / 0.16), e0, phi0, Pi0, t0, tau,
                    method="Radau")
    e_want, phi_want, Pi_want = sol.sol(tau)
    Pi_got = o.get("Pi")[0]
    assert np.isfinite(Pi_got)
    assert o.get("e")[0] == pytest.approx(e_want, rel=1e-3)
    assert Pi_got == pytest.approx(Pi_want, rel=0.1, abs=1e-4 * e_want)
    L.orc_eos_destroy(E)
    o0, E, L, tau = _bjorken_lattice(dt, t0, nsteps, e0, phi0, Pi0, zeta_Tpeak_GeV=0.16, zeta_peak=0.0)
    assert abs(o0.get("Pi")[0]) < 1e-2 * abs(o.get("Pi")[0]) + 1e-3 * o0.get("p")[0]
    L.orc_eos_destroy(E)


def test_extension_leaves_reference_configurations_unchanged():
    """With the reference's sound speed (cs2_mode 0: c_s^2 = 1/3 on this
    table) the bulk coefficients vanish, so the zeta/s parameters change
    nothing, bit for bit, on a 2+1D state with shear and bulk stress."""
    tab = V.eos_table()
    n = 24
    ic = V.random_ic(n, n, 1, 0.1, 0.6, seed=5, lattice=True, bulk=True)
    outs = []
    for kw in ({}, dict(zeta_Tpeak_GeV=0.16, zeta_peak=0.05)):
        o = V.Oracle(n, n, 1, 0.1, 0.1, 0.1, 0.01, 0.6, 0.1, bulk=True, table=tab, **kw)
        o.set_ic(ic)
        o.init_conserved(0.6)
        for k in range(5):
            assert o.step(0.6 + 0.01 * k) == 0
        outs.append(np.concatenate([o.get(f) for f in ("e", "Pi", "pixx", "ux")]))
    assert np.array_equal(outs[0], outs[1])
