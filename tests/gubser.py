"""Viscous Gubser flow (conformal, q = 1 fm^-1) as the analytic oracle of the
2+1D step (mathematica/gubserFlowTest.nb, "Compare 3+1d code to
Israel-Stewart Gubser": T-hat(rho), pi-bar(rho) from an ODE in de Sitter time
rho, T-hat(0) = 1.2, pi-bar(0) = 0), mapped to Milne coordinates.

Gubser-frame quantities (hats) are those of a fluid at rest in
(rho, theta, phi, eta) with ds^2 = tau^2 (-drho^2 + cosh^2 rho dOmega^2 + deta^2):
  rho(tau, r) = -asinh((1 - tau^2 + r^2) / (2 tau)),
  tanh kappa = 2 tau r / (1 + tau^2 + r^2),  u = (cosh kappa, sinh kappa) in (tau, r),
  T = T-hat(rho) / tau,  e = 13.8997 T^4 (rhic/include/EquationOfState.h:19).
The shear tensor is diagonal in the Gubser frame with (mostly-plus mixed
components) pi^eta_eta = pibar * w (w = e + p; with
d eps / d rho = -(eps + p) theta - pi^{mu nu} sigma_{mu nu} this gives the
ODE's d ln T-hat / d rho = -(2/3) tanh rho + (pibar / 3) tanh rho) and
pi^theta_theta = pi^phi_phi = -pibar * w / 2.  Contravariant components are
the same in either signature, so in the reference's Milne components, with the
unit vector n = (sinh kappa, cosh kappa) orthogonal to u in the (tau, r) plane:
pi^{ab} = pi_theta n^a n^b (a, b in tau, r), pi^phiphi = pi_theta / r^2,
pi^etaeta = pibar w / tau^2, pi_theta = -pibar w / 2.

The pibar equation follows from the reference's relaxation equation
(rhic/src/SourceTerms.cpp:55-313: -pi/tau_pi, 2 beta_pi sigma with
beta_pi = (e+p)/5, -delta_pipi pi theta with delta_pipi = 4/3,
-tau_pipi pi^<mu_lambda sigma^nu>lambda with tau_pipi = 10/7; no acceleration
or vorticity in the Gubser frame):
  d pibar / d rho = -pibar T-hat / (5 eta/s) + (4/15) tanh rho
                    + (10/21) pibar tanh rho - (4/3) pibar^2 tanh rho.
The notebook's system is the same without the tau_pipi term (lam = 0).
"""
import math

import numpy as np
from scipy.integrate import solve_ivp

EOS_FACTOR = 13.8997


TAU_PIPI_TERM = 10.0 / 21.0


def solve(etas, lam=TAU_PIPI_TERM, T0hat=1.2, rho_span=(-12.0, 12.0)):
    """T-hat(rho), pi-bar(rho) on a dense interpolant; lam multiplies the
    pibar tanh(rho) term (10/21 for the reference's equations, 0 for the
    notebook's)."""
    def rhs(rho, y):
        T, pb = y
        th = math.tanh(rho)
        dT = -2.0 / 3.0 * T * th + T * pb * th / 3.0
        dpb = -T * pb / (5.0 * etas) - 4.0 / 3.0 * pb * pb * th + 4.0 / 15.0 * th + lam * pb * th
        return [dT, dpb]

    fw = solve_ivp(rhs, (0.0, rho_span[1]), [T0hat, 0.0], method="Radau", rtol=1e-11, atol=1e-13, dense_output=True)
    bw = solve_ivp(rhs, (0.0, rho_span[0]), [T0hat, 0.0], method="Radau", rtol=1e-11, atol=1e-13, dense_output=True)

    def at(rho):
        rho = np.asarray(rho, dtype=float)
        out = np.empty((2,) + rho.shape)
        m = rho >= 0
        if m.any():
            out[:, m] = fw.sol(rho[m])
        if (~m).any():
            out[:, ~m] = bw.sol(rho[~m])
        return out[0], out[1]

    return at


def fields(sol, tau, x, y):
    """Milne-frame e, u^mu, pi^{mu nu} (conserved-variable order) at (tau, x, y)."""
    r = np.sqrt(x * x + y * y)
    rho = -np.arcsinh((1.0 - tau * tau + r * r) / (2.0 * tau))
    That, pbar = sol(rho)
    T = That / tau
    e = EOS_FACTOR * T ** 4
    w = 4.0 / 3.0 * e
    kappa = np.arctanh(2.0 * tau * r / (1.0 + tau * tau + r * r))
    ch, sh = np.cosh(kappa), np.sinh(kappa)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(r > 0, x / np.where(r > 0, r, 1.0), 1.0)
        s = np.where(r > 0, y / np.where(r > 0, r, 1.0), 0.0)
    ut, ux, uy = ch, sh * c, sh * s
    pth = -pbar * w / 2.0
    pitt = pth * sh * sh
    pitr = pth * sh * ch
    pirr = pth * ch * ch
    r2piphiphi = pth  # r^2 pi^{phi phi}
    pixx = pirr * c * c + r2piphiphi * s * s
    piyy = pirr * s * s + r2piphiphi * c * c
    pixy = (pirr - r2piphiphi) * s * c
    zero = np.zeros_like(e)
    pi = [pitt, pitr * c, pitr * s, zero, pixx, pixy, zero, piyy, zero, pbar * w / (tau * tau)]
    return e, (ut, ux, uy, zero), pi
