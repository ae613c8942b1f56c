"""Thermal vorticity on the device (vh_thermal_vorticity) against the
oracle's restatement (pinned bit for bit to the reference's own
calculateThermalVorticity in tests/test_oracle_vorticity.py): bit for bit,
before the first step (u = previous u) and after each of several steps, 2+1D
and 3+1D, conformal and table EoS."""
import pytest

import vhtest as V
from cpuvh_amd import Hydro

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,nz", [(23, 19, 1), (15, 13, 9)])
@pytest.mark.parametrize("lattice,bulk", [(False, False), (True, True)])
def test_vorticity_bitwise(nx, ny, nz, lattice, bulk):
    print(f"case vorticity nx={nx} ny={ny} nz={nz} lattice={lattice} bulk={bulk}", flush=True)
    dx, dt, t0 = 0.1, 0.02, 0.6
    table = V.eos_table() if lattice else None
    ic = V.random_ic(nx, ny, nz, dx, t0, seed=8, lattice=lattice, bulk=bulk)
    g = Hydro(nx, ny, nz, dx, dx, dx, dt, t0, 0.1, T_reg_GeV=0.1, bulk=bulk, eos_table=table)
    g.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic.get("Pi"))
    g.init_conserved(t0)
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, 0.1, T_reg_GeV=0.1, bulk=bulk, table=table)
    o.set_ic(ic)
    o.init_conserved(t0)
    for n in range(0, 4):
        if n:
            t = t0 + (n - 1) * dt
            g.step(t)
            assert o.step(t) == 0
        got, want = g.thermal_vorticity(), o.vorticity()
        assert V.bits_equal(got, want), f"after {n} steps: max rel {V.max_rel(got, want):.3e}"
