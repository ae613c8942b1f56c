"""The plain FP64 path at the real size of the 3+1D drop-in configuration
(131 x 131 x 61 = 1,046,821 cells, seeded MC-Glauber Au+Au, table EoS, bulk,
eta/s(T), regulation below 100 MeV) against the oracle run in the same
process for the first 10 steps, bit for bit (the whole 600-step run against
the oracle's fixture is tests/test_gpu_fullrun.py)."""
import os
import sys
import time

import numpy as np
import pytest

import vhtest as V
from cpuvh_amd import Hydro

pytestmark = pytest.mark.gpu

FIELDS = ["ttt", "ttx", "tty", "ttn", "pitt", "pixx", "pixy", "pinn", "Pi", "e", "p", "ut", "ux", "uy", "un"]


def test_dropin_configuration_full_size():
    sys.path.insert(0, V.REPO)
    import bench as B

    w = B.WORKLOADS["glauber3d"]
    table = V.eos_table()
    ic, info = B.make_ic(w, table)
    nx, ny, nz, dx, dt, t0 = w["nx"], w["ny"], w["nz"], w["dx"], w["dt"], w["t0"]
    print(f"case full-size drop-in {nx}x{ny}x{nz} glauber b={info['b']:.2f}", flush=True)
    kw = dict(T_reg_GeV=w["T_reg_GeV"], bulk=True, etas_Tk_GeV=w["etas_Tk_GeV"], etas_slope_lo=w["etas_slope_lo"],
              etas_slope_hi=w["etas_slope_hi"])
    g = Hydro(nx, ny, nz, dx, dx, dx, dt, t0, w["etabar"], eos_table=table, **kw)
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, w["etabar"], table=table, **kw)
    g.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic["Pi"])
    o.set_ic(ic)
    g.init_conserved(t0)
    o.init_conserved(t0)
    nsteps = 10
    t_or = time.perf_counter()
    for n in range(1, nsteps + 1):
        t = t0 + (n - 1) * dt
        g.step(t)
        assert o.step(t) == 0
    g.sync()
    print(f"oracle {time.perf_counter() - t_or:.1f} s for {nsteps} steps", flush=True)
    for f in FIELDS:
        a, b = o.get(f), g.get(f)
        assert V.bits_equal(a, b), f"{f}: max rel {V.max_rel(a, b):.3e}"
    g.close()
