"""The two whole runs of the plain FP64 path that the full-run parity tests
replay (tests/test_gpu_fullrun.py) and whose oracle results
tests/golden/make_fullrun.py writes:

- gubser: the Gubser configuration, 201 x 201 x 1, dx = dy = 0.05 fm,
  tau0 = 1.0 fm/c, dtau = 0.0125 fm/c, 400 RK2 steps, conformal EoS,
  eta/s = 0.2, bulk off, the semi-analytic viscous Gubser state (q = 1/fm)
  as initial state.  Regulation below T_reg = 100 MeV: with the regulator off
  the algorithm itself blows up within 40 steps in the grid corners, where the
  solution's inverse Reynolds number is large.
- glauber3d: the 3+1D drop-in configuration, 131 x 131 x 61, dx = dy =
  deta = 0.1, tau0 = 0.6 fm/c, dtau = 0.02 fm/c, 600 steps, seeded MC-Glauber
  Au+Au, table EoS, bulk on, eta/s(T) with its minimum 0.08 at 154 MeV,
  regulation below 100 MeV.

Freezeout at T_F = 150 MeV, flagged every coarse step (factor 2, the default
of the reference's rhic-conf/lattice.properties).
"""
import hashlib
import os
import sys

import numpy as np

import vhtest as V

HERE = os.path.dirname(os.path.abspath(__file__))

FIELDS = ["ttt", "ttx", "tty", "ttn", "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn",
          "pinn", "e", "p", "ut", "ux", "uy", "un", "up_t", "up_x", "up_y", "up_n", "regfac"]
FIELDS_BULK = FIELDS + ["Pi", "regfac_bulk"]

CASES = {
    "gubser": dict(nx=201, ny=201, nz=1, dx=0.05, dt=0.0125, t0=1.0, steps=400, coarse=2, etabar=0.2,
                   T_reg_GeV=0.1, T_F_GeV=0.15, checkpoints=(100, 200, 300, 400)),
    "glauber3d": dict(nx=131, ny=131, nz=61, dx=0.1, dt=0.02, t0=0.6, steps=600, coarse=2, etabar=0.08,
                      T_reg_GeV=0.1, T_F_GeV=0.15, checkpoints=(150, 300, 450, 600)),
}


def make_case(name):
    """(initial state, keyword arguments of Hydro/Oracle, EoS kind)."""
    c = CASES[name]
    if name == "gubser":
        import gubser as Gb
        n, dx = c["nx"], c["dx"]
        x = (np.arange(n) - (n - 1) / 2) * dx
        Y, X = np.meshgrid(x, x, indexing="ij")
        sol = Gb.solve(c["etabar"])
        e, u, pi = Gb.fields(sol, c["t0"], X, Y)
        ic = {"e": e.ravel(), "p": (e / 3).ravel(), "ut": u[0].ravel(), "ux": u[1].ravel(), "uy": u[2].ravel(),
              "un": u[3].ravel(), "pi": np.concatenate([p.ravel() for p in pi])}
        return ic, dict(T_reg_GeV=c["T_reg_GeV"]), "conformal"
    sys.path.insert(0, V.REPO)
    import bench as B
    w = B.WORKLOADS["glauber3d"]
    for k in ("nx", "ny", "nz", "dx", "dt", "t0", "etabar", "T_reg_GeV"):
        assert w[k] == c[k], k
    ic, _ = B.make_ic(w, V.eos_table())
    kw = dict(T_reg_GeV=w["T_reg_GeV"], bulk=True, etas_Tk_GeV=w["etas_Tk_GeV"], etas_slope_lo=w["etas_slope_lo"],
              etas_slope_hi=w["etas_slope_hi"])
    return ic, kw, "table"


def ic_hash(ic):
    h = hashlib.sha256()
    for k in ("e", "p", "ut", "ux", "uy", "un", "pi"):
        h.update(np.ascontiguousarray(ic[k], dtype="<f8").tobytes())
    return np.frombuffer(h.digest(), dtype=np.uint8)


def oracle_eF(case, table):
    L = V.oracle_lib()
    eos = L.orc_eos_create(None if table is None else table.ctypes.data)
    try:
        return L.orc_eos_energy_density(eos, case["T_F_GeV"] / V.HBARC)
    finally:
        L.orc_eos_destroy(eos)


def fixture_path(name):
    return os.path.join(HERE, "golden", f"fullrun_{name}.npz")
