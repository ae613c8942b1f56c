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
