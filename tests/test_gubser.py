"""Viscous Gubser flow (conformal EoS, eta/s = 0.2, q = 1 fm^-1) through the
oracle -- the analytic check of the 2+1D step.  Lattice of the Gubser
configuration: 201 x 201 cells, dx = dy = 0.05 fm, tau0 = 1 fm/c,
dtau = 0.0125 fm/c.  The initial state is the semi-analytic solution at
tau0 (tests/gubser.py).  The regulator runs at T_reg = 100 MeV: with no
regulation the reference algorithm blows up within 40 steps in the grid
corners (r ~ 6.5 fm), where the solution's pibar ~ 0.6; inside r < 1 fm e
stays far above e_reg up to tau = 1.5 fm/c, so the comparison region is
unregulated.  Tolerances are the scheme's discretisation error at this
resolution (the minmod limiter with theta = 1.1 makes the scheme close to
first order at the extremum), and the error shrinks as dx is halved."""
import numpy as np
import pytest

import gubser as Gb
import vhtest as V

ETAS = 0.2
T_REG = 0.1


def _grid(n, dx):
    x = (np.arange(n) - (n - 1) / 2) * dx
    Y, X = np.meshgrid(x, x, indexing="ij")
    return X, Y


def _run(n, dx, nsteps, sol):
    dt = dx / 4
    X, Y = _grid(n, dx)
    e, u, pi = Gb.fields(sol, 1.0, X, Y)
    ic = {"e": e.ravel(), "p": (e / 3).ravel(), "ut": u[0].ravel(), "ux": u[1].ravel(), "uy": u[2].ravel(),
          "un": u[3].ravel(), "pi": np.concatenate([p.ravel() for p in pi])}
    o = V.Oracle(n, n, 1, dx, dx, dx, dt, 1.0, ETAS, T_reg_GeV=T_REG)
    o.set_ic(ic)
    o.init_conserved(1.0)
    for s in range(1, nsteps + 1):
        assert o.step(1.0 + (s - 1) * dt) == 0
    return o, 1.0 + nsteps * dt, X, Y


def _errors(o, tau, X, Y, sol):
    n = X.shape[0]
    e2, u2, pi2 = Gb.fields(sol, tau, X, Y)
    core = X * X + Y * Y < 1.0
    e = o.get("e").reshape(n, n)
    ux = o.get("ux").reshape(n, n)
    c = n // 2
    return {"e": float(np.max(np.abs(e - e2)[core] / e2[core])),
            "ux": float(np.max(np.abs(ux - u2[1])[core])),
            "pinn": abs(o.get("pinn").reshape(n, n)[c, c] / pi2[9][c, c] - 1)}


def test_gubser_core_against_semi_analytic():
    sol = Gb.solve(ETAS)
    o, tau, X, Y = _run(201, 0.05, 40, sol)  # tau = 1.5 fm/c
    err = _errors(o, tau, X, Y, sol)
    assert err["e"] < 2.5e-2 and err["ux"] < 1e-2 and err["pinn"] < 4e-2, err
    # the viscous check distinguishes eta/s: the eta/s = 0.1 solution misses pi^{eta eta}
    wrong = _errors(o, tau, X, Y, Gb.solve(0.1))
    assert wrong["pinn"] > 2.5 * err["pinn"] and wrong["pinn"] > 0.1, (wrong, err)


def test_gubser_error_shrinks_with_resolution():
    sol = Gb.solve(ETAS)
    coarse = _errors(*_run(101, 0.1, 10, sol), sol)   # tau = 1.25
    fine = _errors(*_run(201, 0.05, 20, sol), sol)    # tau = 1.25
    assert fine["e"] < 0.6 * coarse["e"], (fine, coarse)


def test_gubser_solution_satisfies_conservation():
    """The semi-analytic fields conserve T^{mu nu} in Milne coordinates
    (central differences): pins the Gubser-to-Milne mapping of e and u."""
    sol = Gb.solve(ETAS)

    def tmunu(tau, x, y):
        e, u, pi = Gb.fields(sol, tau, np.array([x]), np.array([y]))
        e, p = e[0], e[0] / 3
        ut, ux, uy = u[0][0], u[1][0], u[2][0]
        w = e + p
        T = np.array([[w * ut * ut - p + pi[0][0], w * ut * ux + pi[1][0], w * ut * uy + pi[2][0]],
                      [w * ux * ut + pi[1][0], w * ux * ux + p + pi[4][0], w * ux * uy + pi[5][0]],
                      [w * uy * ut + pi[2][0], w * uy * ux + pi[5][0], w * uy * uy + p + pi[7][0]]])
        return T, p / tau ** 2 + pi[9][0]

    h = 1e-4
    for tau, x, y in ((1.0, 1.0, 0.0), (1.3, 0.4, -0.7), (2.0, 0.0, 1.5)):
        T, tnn = tmunu(tau, x, y)
        dtau = (tmunu(tau + h, x, y)[0] - tmunu(tau - h, x, y)[0]) / (2 * h)
        dx_ = (tmunu(tau, x + h, y)[0] - tmunu(tau, x - h, y)[0]) / (2 * h)
        dy_ = (tmunu(tau, x, y + h)[0] - tmunu(tau, x, y - h)[0]) / (2 * h)
        r0 = dtau[0, 0] + dx_[1, 0] + dy_[2, 0] + (T[0, 0] + tau * tau * tnn) / tau
        r1 = dtau[0, 1] + dx_[1, 1] + dy_[2, 1] + T[0, 1] / tau
        r2 = dtau[0, 2] + dx_[1, 2] + dy_[2, 2] + T[0, 2] / tau
        scale = abs(dtau[0, 0]) + abs(dtau[0, 1]) + abs(dtau[0, 2])
        assert max(abs(r0), abs(r1), abs(r2)) < 1e-5 * scale, (tau, x, y, r0, r1, r2)


def test_ideal_gubser_initial_state_evolves_as_published():
    """The reference's IC type 3 (ideal Gubser flow at tau = 1, built by
    vh_ic_test_problem) through the oracle at eta/s = 0.001: at tau = 1.2 the
    state is the published ideal solution (Gubser, PRD 82 (2010) 085027,
    q = 1): e = e(0, tau = 1) * 4^(4/3) / (tau^(4/3) [1 + 2(tau^2 + r^2) +
    (tau^2 - r^2)^2]^(4/3)) and v_perp = 2 tau r / (1 + tau^2 + r^2), to the
    scheme's error at dx = 0.1 fm inside r < 0.75 fm (0.2% for r < 0.5 fm).
    Further out, around r = 1.2 fm, the minmod limiter clips the extremum of
    the radial momentum density and the error reaches 10% at this resolution
    (6% at dx = 0.05 fm), the same at every eta/s and time step."""
    import cpuvh_amd as C
    n, dx, dt, nsteps = 81, 0.1, 0.0025, 80
    ic = C.analytic_ic(3, n, n, 1, dx, dx, e0=1.0)
    ic["pi"] = np.zeros(10 * n * n)
    o = V.Oracle(n, n, 1, dx, dx, dx, dt, 1.0, 0.001)
    o.set_ic(ic)
    o.init_conserved(1.0)
    for s in range(nsteps):
        assert o.step(1.0 + s * dt) == 0
    tau = 1.0 + nsteps * dt
    X, Y = _grid(n, dx)
    r = np.hypot(X, Y).ravel()
    e_c0 = ic["e"][(n // 2) * n + n // 2]
    e_ana = e_c0 * 4 ** (4 / 3) / (tau ** (4 / 3) * (1 + 2 * (tau * tau + r * r) + (tau * tau - r * r) ** 2) ** (4 / 3))
    v_ana = 2 * tau * r / (1 + tau * tau + r * r)
    e = o.get("e")
    v = np.hypot(o.get("ux"), o.get("uy")) / o.get("ut")
    assert np.max(np.abs(e[r < 0.5] / e_ana[r < 0.5] - 1)) < 3e-3
    assert np.max(np.abs(e[r < 0.75] / e_ana[r < 0.75] - 1)) < 1e-2
    assert np.max(np.abs(v[r < 0.75] - v_ana[r < 0.75])) < 5e-3
    # the expansion is real: the centre has cooled by (1.2)^(-4/3) and more
    assert e[(n // 2) * n + n // 2] < 0.7 * e_c0


def test_israel_stewart_gubser_file_round_trip(tmp_path):
    """IC type 1 (the reference reads gubserIC.dat, which it does not ship):
    the file made from the semi-analytic solution at tau = 1 in the format of
    rhic/src/InitialConditions.cpp:1223-1224 reads back as those fields."""
    import cpuvh_amd as C
    n, dx = 9, 0.5
    path = str(tmp_path / "gubserIC.dat")
    write_gubser_ic(path, Gb.solve(ETAS), n, dx)
    got = C.read_gubser(path, n, n, 2)
    X, Y = _grid(n, dx)
    e, u, pi = Gb.fields(Gb.solve(ETAS), 1.0, X, Y)
    np.testing.assert_allclose(got["e"][: n * n], e.ravel(), rtol=1e-15)
    np.testing.assert_allclose(got["ux"][: n * n], u[1].ravel(), rtol=1e-15, atol=1e-300)
    for c in (0, 1, 2, 4, 5, 7, 9):
        np.testing.assert_allclose(got["pi"][c * 2 * n * n: c * 2 * n * n + n * n], pi[c].ravel(), rtol=1e-15,
                                   atol=1e-300)
    np.testing.assert_array_equal(got["e"][: n * n], got["e"][n * n:])


def write_gubser_ic(path, sol, n, dx):
    """gubserIC.dat for an n x n lattice at tau = 1: x y e ux uy pixx piyy
    pixy pitt pitx pity pinn per line, i (x) outer, j (y) inner."""
    x = (np.arange(n) - (n - 1) / 2) * dx
    with open(path, "w") as f:
        for i in range(n):
            for j in range(n):
                e, u, pi = Gb.fields(sol, 1.0, np.array([x[i]]), np.array([x[j]]))
                vals = [x[i], x[j], e[0], u[1][0], u[2][0], pi[4][0], pi[7][0], pi[5][0], pi[0][0], pi[1][0],
                        pi[2][0], pi[9][0]]
                f.write("\t".join("%.17g" % v for v in vals) + "\n")
