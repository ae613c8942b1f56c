"""The bulk-viscosity extension on the HIP path: the table's own sound speed
(cs2_mode 1, so (1/3 - c_s^2) != 0 and the bulk terms act) and zeta/s(T)
peaked at a chosen temperature with a per-event height, including the
bounded relaxation rate (tiny zeta/s) and zeta/s = 0.  The FP64 paths are
compared with the oracle bit for bit, the mixed form with the FP64 batch
within the precision form's tolerances."""
import numpy as np
import pytest

import vhtest as V
from cpuvh_amd import Batch, Hydro, PREC_MIXED

pytestmark = pytest.mark.gpu

FIELDS = ["ttt", "ttx", "tty", "ttn", "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn",
          "pinn", "Pi", "e", "p", "ut", "ux", "uy", "un"]
ZETA_CASES = [(0.16, 0.1), (0.0, 0.0), (0.16, 1e-4), (0.16, 0.0)]


def _oracle_run(nx, ny, nz, dx, dt, t0, etabar, ic, nsteps, T_reg, cs2_mode, zt, zp):
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=T_reg, bulk=True, table=V.eos_table(),
                 etas_slope_lo=0.5, etas_slope_hi=1.0, cs2_mode=cs2_mode, zeta_Tpeak_GeV=zt, zeta_peak=zp)
    o.set_ic(ic)
    o.init_conserved(t0)
    for n in range(1, nsteps + 1):
        assert o.step(t0 + (n - 1) * dt) == 0
    return o


@pytest.mark.parametrize("nx,ny,nz", [(21, 17, 1), (13, 11, 9)])
@pytest.mark.parametrize("zt,zp", ZETA_CASES)
def test_bulk_extension_bitwise(nx, ny, nz, zt, zp):
    print(f"case bulk ext nx={nx} ny={ny} nz={nz} zeta_Tpeak={zt} zeta_peak={zp}", flush=True)
    dx, dt, t0, etabar, T_reg, nsteps = 0.1, 0.02, 0.6, 0.08, 0.1, 6
    ic = V.random_ic(nx, ny, nz, dx, t0, seed=11, lattice=True, bulk=True)
    g = Hydro(nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=T_reg, bulk=True, eos_table=V.eos_table(),
              etas_slope_lo=0.5, etas_slope_hi=1.0, cs2_mode=1, zeta_Tpeak_GeV=zt, zeta_peak=zp)
    g.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic["Pi"])
    g.init_conserved(t0)
    g.steps(1, nsteps)
    g.sync()
    o = _oracle_run(nx, ny, nz, dx, dt, t0, etabar, ic, nsteps, T_reg, 1, zt, zp)
    for f in FIELDS:
        a, b = o.get(f), g.get(f)
        assert V.bits_equal(a, b), f"{f}: max rel {V.max_rel(a, b):.3e}"
    # the extension acts: the reference's sound speed gives another Pi
    ref = _oracle_run(nx, ny, nz, dx, dt, t0, etabar, ic, nsteps, T_reg, 0, zt, zp)
    assert not np.array_equal(ref.get("Pi"), g.get("Pi"))


EVENTS = [(17, 15, 1, 0.1), (12, 14, 1, 0.05), (11, 9, 5, 0.0)]


def _batch(ics, precision, nsteps):
    cap = sum(nx * ny * nz for nx, ny, nz, _ in EVENTS)
    b = Batch(cap, max_events=len(EVENTS), max_steps=nsteps, bulk=True, eos_table=V.eos_table(),
              precision=precision, cs2_mode=1)
    ids = [b.add(nx, ny, nz, 0.1, 0.1, 0.1, 0.02, 0.6, 0.08, ic, T_reg_GeV=0.1, etas_slope_lo=0.5,
                 etas_slope_hi=1.0, zeta_Tpeak_GeV=0.16, zeta_peak=zp)
           for (nx, ny, nz, zp), ic in zip(EVENTS, ics)]
    assert None not in ids
    b.steps(nsteps)
    b.sync()
    return b, ids


def test_batch_per_event_zeta_bitwise():
    """Three events with their own zeta/s heights in one FP64 batch, each
    equal to the oracle bit for bit."""
    nsteps = 5
    ics = [V.random_ic(nx, ny, nz, 0.1, 0.6, seed=20 + k, lattice=True, bulk=True)
           for k, (nx, ny, nz, _) in enumerate(EVENTS)]
    b, ids = _batch(ics, 0, nsteps)
    for (nx, ny, nz, zp), ic, eid in zip(EVENTS, ics, ids):
        print(f"case batch event {nx}x{ny}x{nz} zeta_peak={zp}", flush=True)
        o = _oracle_run(nx, ny, nz, 0.1, 0.02, 0.6, 0.08, ic, nsteps, 0.1, 1, 0.16, zp)
        for f in FIELDS:
            a, c = o.get(f), b.get(eid, f)
            assert V.bits_equal(a, c), f"event {eid} {f}: max rel {V.max_rel(a, c):.3e}"


def test_mixed_bulk_extension_within_tolerance():
    """The mixed form with the extension against the FP64 batch after 30
    steps: e and u^mu pointwise within 1e-3, pi^munu and Pi within 1e-2 of
    their largest magnitude."""
    nsteps = 30
    ics = [V.random_ic(nx, ny, nz, 0.1, 0.6, seed=40 + k, lattice=True, bulk=True, pi_scale=0.005)
           for k, (nx, ny, nz, _) in enumerate(EVENTS)]
    b64, ids64 = _batch(ics, 0, nsteps)
    bmx, idsmx = _batch(ics, PREC_MIXED, nsteps)
    for ev, i64, imx in zip(EVENTS, ids64, idsmx):
        ref = {f: b64.get(i64, f) for f in FIELDS}
        got = {f: bmx.get(imx, f) for f in FIELDS}
        errs = {f: float(np.max(np.abs(ref[f] - got[f]) / np.abs(ref[f]))) for f in ("e", "ut")}
        for f in ("pixx", "pixy", "pinn", "Pi", "ux", "uy"):
            errs[f] = float(np.max(np.abs(ref[f] - got[f]))) / (float(np.max(np.abs(ref[f]))) or 1.0)
        print(ev, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert errs["e"] < 1e-3 and errs["ut"] < 1e-3, errs
        assert errs["ux"] < 1e-3 and errs["uy"] < 1e-3, errs
        for f in ("pixx", "pixy", "pinn", "Pi"):
            assert errs[f] < 1e-2, (f, errs[f])
