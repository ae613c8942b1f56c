"""Mixed-precision form (FP32 storage and flux/source arithmetic; FP64 root
solve, EoS lookups, accumulation and regulator) against the FP64 path on
identical inputs: relative error within 1e-3 in e and u^mu (pointwise for e
and u^tau, against the field's largest magnitude for u^x, u^y, u^eta and
pi^{mu nu}), every tile variant bit-identical to every other, and an event
bit-identical alone and inside a mixed batch."""
import numpy as np
import pytest

import vhtest as V
from cpuvh_amd import Batch, PREC_MIXED

pytestmark = pytest.mark.gpu

EVENTS = [
    (41, 37, 1, 0.10, 0.020, 0.60, 0.12, 0.10, (0.5, 1.0), 11),
    (19, 17, 13, 0.10, 0.020, 0.60, 0.08, 0.10, (0.5, 1.0), 12),
    (50, 9, 1, 0.08, 0.010, 1.00, 0.20, 0.00, (0.0, 0.0), 13),
]
PRIM = ["e", "ut", "ux", "uy", "un"]
PI = ["pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"]


def _ic(ev, lattice, bulk):
    """A smooth seeded fireball: Gaussian e over a floor, Hubble-like flow and
    a smooth shear stress (the mixed form's error bar is stated for resolved
    flow; the FP64 tests use noisy profiles)."""
    nx, ny, nz, dx, dt, t0, etabar, treg, slopes, seed = ev
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) - (nx - 1) / 2) * dx
    y = (np.arange(ny) - (ny - 1) / 2) * dx
    z = (np.arange(nz) - (nz - 1) / 2) * dx
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    w = rng.uniform(0.8, 1.2, 3)
    g = np.exp(-(w[0] * X ** 2 + w[1] * Y ** 2) / 2 - w[2] * Z ** 2 / 2)
    e = 20.0 * g + 0.5
    ux, uy = 0.05 * X, 0.04 * Y
    un = 0.01 * Z if nz > 1 else np.zeros_like(X)
    ut = np.sqrt(1 + ux ** 2 + uy ** 2 + t0 ** 2 * un ** 2)
    pi = 0.02 * g[None] * rng.uniform(-1, 1, (10, 1, 1, 1))
    p = e / 3.0 if not lattice else e / 3.4
    ic = {"e": e.ravel(), "p": p.ravel(), "ut": ut.ravel(), "ux": ux.ravel(), "uy": uy.ravel(),
          "un": un.ravel(), "pi": pi.reshape(10, -1).ravel()}
    if bulk:
        ic["Pi"] = (0.002 * g).ravel()
    return ic


def _batch(evs, ics, lattice, bulk, precision, nsteps, tiles=None):
    cap = sum(e[0] * e[1] * e[2] for e in evs)
    b = Batch(cap, max_events=len(evs), max_steps=max(nsteps, 16), bulk=bulk,
              eos_table=V.eos_table() if lattice else None, precision=precision)
    if tiles is not None:
        b.set_tiles(*tiles)
    ids = []
    for ev, ic in zip(evs, ics):
        nx, ny, nz, dx, dt, t0, etabar, treg, (slo, shi), seed = ev
        ids.append(b.add(nx, ny, nz, dx, dx, dx, dt, t0, etabar, ic, T_reg_GeV=treg, etas_slope_lo=slo,
                         etas_slope_hi=shi))
    b.steps(nsteps)
    b.sync()
    return b, ids


def _rel_errors(ref, got):
    out = {}
    for f in PRIM + PI:
        a, c = ref[f], got[f]
        scale = float(np.max(np.abs(a))) or 1.0
        out[f] = float(np.max(np.abs(a - c))) / scale
    out["e_pointwise"] = float(np.max(np.abs(ref["e"] - got["e"]) / np.abs(ref["e"])))
    out["ut_pointwise"] = float(np.max(np.abs(ref["ut"] - got["ut"]) / np.abs(ref["ut"])))
    return out


@pytest.mark.parametrize("lattice,bulk", [(False, False), (True, True)])
def test_mixed_against_fp64(lattice, bulk):
    print(f"case mixed vs fp64 lattice={lattice} bulk={bulk}", flush=True)
    nsteps = 30
    ics = [_ic(ev, lattice, bulk) for ev in EVENTS]
    b64, ids64 = _batch(EVENTS, ics, lattice, bulk, 0, nsteps)
    bmx, idsmx = _batch(EVENTS, ics, lattice, bulk, PREC_MIXED, nsteps)
    for ev, i64, imx in zip(EVENTS, ids64, idsmx):
        ref = {f: b64.get(i64, f) for f in PRIM + PI}
        got = {f: bmx.get(imx, f) for f in PRIM + PI}
        err = _rel_errors(ref, got)
        print(ev[:3], {k: f"{v:.2e}" for k, v in err.items()}, flush=True)
        assert err["e_pointwise"] < 1e-3 and err["ut_pointwise"] < 1e-3, err
        for f in ("ux", "uy", "un"):
            assert err[f] < 1e-3, (f, err[f])
        for f in PI:
            assert err[f] < 1e-2, (f, err[f])


def test_tile_variants_bit_identical():
    """Every face flux is the same arithmetic whichever tile evaluates it, so
    all tile variants give the same bits."""
    lattice, bulk, nsteps = True, False, 6
    ics = [_ic(ev, lattice, bulk) for ev in EVENTS]
    runs = []
    for tile in range(3):
        b, ids = _batch(EVENTS, ics, lattice, bulk, PREC_MIXED, nsteps, tiles=(tile, tile))
        runs.append([{f: b.get(i, f) for f in ["e", "ut", "ux", "ttt", "pixy"]} for i in ids])
    for r in runs[1:]:
        for ev_a, ev_b in zip(runs[0], r):
            for f in ev_a:
                assert V.bits_equal(ev_a[f], ev_b[f]), f


def test_mixed_alone_equals_in_mix():
    lattice, bulk, nsteps = True, True, 8
    ics = [_ic(ev, lattice, bulk) for ev in EVENTS]
    bmix, ids = _batch(EVENTS, ics, lattice, bulk, PREC_MIXED, nsteps)
    for k, ev in enumerate(EVENTS):
        solo, sid = _batch([ev], [ics[k]], lattice, bulk, PREC_MIXED, nsteps)
        for f in PRIM + PI + ["ttt", "ttx", "Pi"]:
            assert V.bits_equal(solo.get(sid[0], f), bmix.get(ids[k], f)), (k, f)


def test_mixed_autotune_keeps_state():
    lattice, bulk = False, False
    ics = [_ic(ev, lattice, bulk) for ev in EVENTS]
    b1, ids1 = _batch(EVENTS, ics, lattice, bulk, PREC_MIXED, 0)
    b2, ids2 = _batch(EVENTS, ics, lattice, bulk, PREC_MIXED, 0)
    t2, t3 = b2.autotune(reps=2)
    assert 0 <= t2 < 3 and 0 <= t3 < 3
    b1.steps(4)
    b2.steps(4)
    for i1, i2 in zip(ids1, ids2):
        for f in ("e", "ut", "pitt"):
            assert V.bits_equal(b1.get(i1, f), b2.get(i2, f)), f


BI_EVENTS = [
    (41, 37, 1, 0.10, 0.020, 0.60, 0.12, 0.10, (0.5, 1.0), 21),
    (50, 9, 1, 0.08, 0.010, 1.00, 0.20, 0.00, (0.0, 0.0), 22),
    (33, 45, 1, 0.10, 0.020, 0.40, 0.08, 0.15, (0.5, 1.0), 23),
]
ETA = ["ttn", "pitn", "pixn", "piyn", "un"]


def _bi_ic(ev, lattice, bulk, tiny=0.0):
    """A boost-invariant 2+1D state: u^eta and pi^{tau eta}, pi^{x eta},
    pi^{y eta} zero (tiny > 0 puts that much pi^{tau eta} in one cell, which
    makes the event general)."""
    ic = _ic(ev, lattice, bulk)
    N = ev[0] * ev[1] * ev[2]
    for c in (3, 6, 8):
        ic["pi"][c * N:(c + 1) * N] = 0.0
    ic["pi"][3 * N + N // 2] = tiny
    return ic


@pytest.mark.parametrize("lattice,bulk", [(False, False), (True, True)])
def test_boost_invariant_kernel(lattice, bulk):
    """Boost-invariant 2+1D events run the kernel that neither loads nor
    evolves the eta components: against FP64 to the mixed form's bar, their
    eta components stay exactly zero, and they agree with the general kernel
    (the same events made general by a 1e-30 pi^{tau eta} in one cell) to
    that bar."""
    print(f"case boost-invariant lattice={lattice} bulk={bulk}", flush=True)
    nsteps = 30
    ics = [_bi_ic(ev, lattice, bulk) for ev in BI_EVENTS]
    b64, ids64 = _batch(BI_EVENTS, ics, lattice, bulk, 0, nsteps)
    bmx, idsmx = _batch(BI_EVENTS, ics, lattice, bulk, PREC_MIXED, nsteps)
    gics = [_bi_ic(ev, lattice, bulk, tiny=1e-30) for ev in BI_EVENTS]
    bgen, idsgen = _batch(BI_EVENTS, gics, lattice, bulk, PREC_MIXED, nsteps)
    for ev, i64, imx, igen in zip(BI_EVENTS, ids64, idsmx, idsgen):
        ref = {f: b64.get(i64, f) for f in PRIM + PI}
        got = {f: bmx.get(imx, f) for f in PRIM + PI}
        gen = {f: bgen.get(igen, f) for f in PRIM + PI}
        for f in ETA:
            assert not np.any(bmx.get(imx, f)), f
        for name, other in (("fp64", ref), ("general", gen)):
            err = _rel_errors(other, got)
            print(ev[:3], name, {k: f"{v:.2e}" for k, v in err.items() if k not in ("un", "pitn", "pixn", "piyn")},
                  flush=True)
            assert err["e_pointwise"] < 1e-3 and err["ut_pointwise"] < 1e-3, (name, err)
            for f in ("ux", "uy"):
                assert err[f] < 1e-3, (name, f, err[f])
            for f in ("pitt", "pitx", "pity", "pixx", "pixy", "piyy", "pinn"):
                assert err[f] < 1e-2, (name, f, err[f])


def test_mixed_3d_at_precision_form_size():
    """The 3+1D mixed kernel on the precision form's lattice (201 x 201 x 101
    = 4,080,501 cells, dx = deta = 0.08, seeded MC-Glauber, table EoS, bulk,
    eta/s(T)) for 10 steps against the FP64 path on the same event: e and u^tau
    within 1e-3 pointwise, u^x, u^y, u^eta within 1e-3 of their largest
    magnitude, and the same freezeout cube set."""
    import cpuvh_amd as C
    nx = ny = 201
    nz, dx, dt, t0, nsteps = 101, 0.08, 0.02, 0.6, 10
    print(f"case mixed 3d {nx}x{ny}x{nz} steps={nsteps}", flush=True)
    table = V.eos_table()
    e, _ = C.mc_glauber(nx, ny, nz, dx, dx, dx, seed=4242, e0=10.0)
    N = e.size
    h = C.Hydro(2, 2, 1, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2, eos_table=table)
    eF = h.eos_energy_density(0.15 / V.HBARC)
    h.close()
    import sys
    sys.path.insert(0, V.REPO)
    import bench as B
    ic = {"e": e, "p": B.table_pressure(table, e), "ut": np.ones(N), "ux": np.zeros(N), "uy": np.zeros(N),
          "un": np.zeros(N), "pi": np.zeros(10 * N), "Pi": np.zeros(N)}
    out = {}
    for prec in (0, PREC_MIXED):
        b = Batch(N, max_events=1, max_steps=nsteps, bulk=True, eos_table=table, precision=prec)
        eid = b.add(nx, ny, nz, dx, dx, dx, dt, t0, 0.08, ic, T_reg_GeV=0.1, etas_slope_lo=0.5, etas_slope_hi=1.0)
        b.steps(nsteps)
        b.sync()
        out[prec] = ({f: b.get(eid, f) for f in PRIM}, b.fo_flag_ids(eid, eF))
        b.close()
    ref, got = out[0][0], out[PREC_MIXED][0]
    err = {f: float(np.max(np.abs(ref[f] - got[f]))) / (float(np.max(np.abs(ref[f]))) or 1.0) for f in PRIM}
    for f in ("e", "ut"):
        err[f + "_pointwise"] = float(np.max(np.abs(ref[f] - got[f]) / np.abs(ref[f])))
    print({k: f"{v:.2e}" for k, v in err.items()}, flush=True)
    assert err["e_pointwise"] < 1e-3 and err["ut_pointwise"] < 1e-3, err
    for f in ("ux", "uy", "un"):
        assert err[f] < 1e-3, (f, err[f])
    assert np.array_equal(out[0][1], out[PREC_MIXED][1])
