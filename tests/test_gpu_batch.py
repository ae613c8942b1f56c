"""Batches of events (vh_batch_*): every event of a mixed batch evolves bit
for bit as it does alone (same kernels, own lattice, spacings, clock and
transport parameters), including events admitted and removed mid-run; the
per-event counts and freezeout flags match the single-event calls."""
import numpy as np
import pytest

import vhtest as V
from cpuvh_amd import Batch, Hydro

pytestmark = pytest.mark.gpu

FIELDS = ["ttt", "ttx", "tty", "ttn", "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy",
          "piyn", "pinn", "e", "p", "ut", "ux", "uy", "un", "up_t", "up_x", "up_y", "up_n"]

# (nx, ny, nz, dx, dt, t0, etabar, T_reg, eta/s slopes, seed): 2+1D and 3+1D
# events of different sizes, spacings, clocks and transport parameters.
EVENTS = [
    (21, 17, 1, 0.10, 0.020, 0.60, 0.20, 0.00, (0.0, 0.0), 3),
    (13, 11, 9, 0.12, 0.015, 0.50, 0.08, 0.00, (0.0, 0.0), 4),
    (33, 29, 1, 0.08, 0.010, 1.00, 0.12, 0.25, (0.0, 0.0), 5),
    (9, 7, 5, 0.10, 0.020, 0.60, 0.08, 0.10, (0.5, 1.0), 6),
    (70, 3, 1, 0.05, 0.010, 0.40, 0.16, 0.00, (0.3, 0.7), 7),
]


def _ic(ev, lattice, bulk):
    nx, ny, nz, dx, dt, t0, etabar, treg, slopes, seed = ev
    return V.random_ic(nx, ny, nz, dx, t0, seed=seed, lattice=lattice, bulk=bulk)


def _solo(ev, ic, lattice, bulk, nsteps):
    nx, ny, nz, dx, dt, t0, etabar, treg, (slo, shi), seed = ev
    h = Hydro(nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=treg, bulk=bulk,
              eos_table=V.eos_table() if lattice else None, etas_slope_lo=slo, etas_slope_hi=shi)
    h.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic.get("Pi"))
    h.init_conserved(t0)
    if nsteps:
        h.steps(1, nsteps)
    h.sync()
    return h


def _add(b, ev, ic):
    nx, ny, nz, dx, dt, t0, etabar, treg, (slo, shi), seed = ev
    return b.add(nx, ny, nz, dx, dx, dx, dt, t0, etabar, ic, T_reg_GeV=treg, etas_slope_lo=slo,
                 etas_slope_hi=shi)


def _fields(bulk):
    return FIELDS + (["Pi"] if bulk else [])


def _same(b, eid, h, bulk):
    for f in _fields(bulk):
        a, c = h.get(f), b.get(eid, f)
        assert V.bits_equal(a, c), f"event {eid} {f}: max rel {V.max_rel(a, c):.3e}"


@pytest.mark.parametrize("lattice,bulk", [(False, False), (True, True), (False, True), (True, False)])
def test_mixed_batch_equals_solo(lattice, bulk):
    print(f"case batch lattice={lattice} bulk={bulk}", flush=True)
    cap = sum(e[0] * e[1] * e[2] for e in EVENTS) + 100
    b = Batch(cap, max_events=8, max_steps=64, bulk=bulk, eos_table=V.eos_table() if lattice else None)
    ics = [_ic(ev, lattice, bulk) for ev in EVENTS]
    ids = [_add(b, ev, ic) for ev, ic in zip(EVENTS, ics)]
    assert None not in ids
    b.steps(3)
    b.steps(4)
    b.sync()
    for ev, ic, eid in zip(EVENTS, ics, ids):
        assert b.info(eid)["steps_done"] == 7
        _same(b, eid, _solo(ev, ic, lattice, bulk, 7), bulk)


def test_batch_matches_oracle():
    """One event of a batch, placed after a neighbour in the pool, against the
    oracle directly (unregulated, table EoS, bulk)."""
    ev = EVENTS[1]
    nx, ny, nz, dx, dt, t0, etabar, treg, slopes, seed = ev
    ic = _ic(ev, True, True)
    n0 = EVENTS[0][0] * EVENTS[0][1] * EVENTS[0][2]
    b = Batch(n0 + nx * ny * nz, max_events=3, max_steps=16, bulk=True, eos_table=V.eos_table())
    _add(b, EVENTS[0], _ic(EVENTS[0], True, True))
    eid = _add(b, ev, ic)
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=treg, bulk=True, table=V.eos_table())
    o.set_ic(ic)
    o.init_conserved(t0)
    for n in range(1, 6):
        assert o.step(t0 + (n - 1) * dt) == 0
    b.steps(5)
    b.sync()
    for f in _fields(True):
        a, c = o.get(f), b.get(eid, f)
        assert V.bits_equal(a, c), f"{f}: max rel {V.max_rel(a, c):.3e}"


def test_admission_and_removal_mid_run():
    """Events join and leave between step chunks; each keeps its own clock and
    the freed cells are reused (first fit)."""
    lattice, bulk = False, False
    A, B, C = EVENTS[0], EVENTS[1], EVENTS[3]
    nA, nB, nC = (e[0] * e[1] * e[2] for e in (A, B, C))
    b = Batch(nA + nB, max_events=4, max_steps=32)
    icA, icB, icC = (_ic(e, lattice, bulk) for e in (A, B, C))
    a = _add(b, A, icA)
    b.steps(3)
    bb = _add(b, B, icB)
    b.steps(4)
    assert _add(b, EVENTS[2], _ic(EVENTS[2], lattice, bulk)) is None  # no room: refused
    b.remove(a)
    free, largest = b.free_cells()
    assert free == nA and largest == nA
    c = _add(b, C, icC)  # nC <= nA: fits into A's old range
    assert c is not None
    b.steps(2)
    b.sync()
    assert b.info(bb)["steps_done"] == 6 and b.info(c)["steps_done"] == 2
    _same(b, bb, _solo(B, icB, lattice, bulk, 6), bulk)
    _same(b, c, _solo(C, icC, lattice, bulk, 2), bulk)


def test_counts_and_freezeout_flags():
    lattice, bulk = True, False
    evs = [EVENTS[0], EVENTS[1], EVENTS[4]]
    ics = [_ic(e, lattice, bulk) for e in evs]
    b = Batch(sum(e[0] * e[1] * e[2] for e in evs), max_events=3, max_steps=16, eos_table=V.eos_table())
    ids = [_add(b, e, ic) for e, ic in zip(evs, ics)]
    solos = [_solo(e, ic, lattice, bulk, 0) for e, ic in zip(evs, ics)]
    b.steps(2)
    b.steps(3)
    for h in solos:
        h.steps(1, 5)
        h.sync()
    eF = 1.5
    counts = b.counts_above(eF)
    for eid, h in zip(ids, solos):
        assert counts[eid] == h.count_above(eF)
        assert np.array_equal(b.fo_flag_ids(eid, eF), h.fo_flag_ids(eF))
        assert np.array_equal(b.fo_flag_ids(eid, eF), h.fo_flag_ids(eF))


def test_step_budget_and_bad_arguments():
    b = Batch(1000, max_events=2, max_steps=4)
    ev = EVENTS[4]
    eid = _add(b, ev, _ic(ev, False, False))
    b.steps(4)
    with pytest.raises(Exception):
        b.steps(1)  # past max_steps
    with pytest.raises(Exception):
        b.remove(eid + 1)
    b.remove(eid)
    assert b.free_cells() == (1000, 1000)


@pytest.mark.parametrize("precision", [0, 1])
def test_device_stop_rule_equals_host_loop(precision):
    """vh_batch_run (the stop rule on the device, one graph for all checks)
    against the host loop of steps, counts and strikes with removal at the
    third zero count: the same events stop at the same checks, with the same
    bits, and the others run on to the same bits."""
    lattice, bulk, coarse, nchecks = True, True, 5, 12
    print(f"case device stop rule precision={precision}", flush=True)
    ics = [_ic(ev, lattice, bulk) for ev in EVENTS]
    kw = dict(max_events=len(EVENTS), max_steps=200, bulk=bulk, eos_table=V.eos_table(), precision=precision)
    cap = sum(ev[0] * ev[1] * ev[2] for ev in EVENTS)
    eF = 9.0  # two of the five events fall below it within the 60 steps (stopping at 55 and 60)
    fields = ["e", "ut", "ux", "uy", "pixx", "Pi", "ttt"]

    hb = Batch(cap, **kw)
    hid = [_add(hb, ev, ic) for ev, ic in zip(EVENTS, ics)]
    strikes = {i: 0 for i in hid}
    host_stop, host_state = {}, {}
    for c in range(nchecks):
        hb.steps(coarse)
        counts = hb.counts_above(eF)
        for i in list(strikes):
            strikes[i] += counts[i] == 0
            if strikes[i] >= 3:
                host_stop[i] = hb.info(i)["steps_done"]
                host_state[i] = {f: hb.get(i, f) for f in fields}
                hb.remove(i)
                del strikes[i]
    for i in strikes:
        host_state[i] = {f: hb.get(i, f) for f in fields}

    db = Batch(cap, **kw)
    did = [_add(db, ev, ic) for ev, ic in zip(EVENTS, ics)]
    st, stopped, steps = db.run(eF, coarse, nchecks // 2)
    assert st == 0
    st, stopped, steps = db.run(eF, coarse, nchecks - nchecks // 2)  # a second call continues the rule
    assert st == 0
    assert 0 < len(host_stop) < len(EVENTS), host_stop  # the case exercises both outcomes
    for h, d in zip(hid, did):
        assert bool(stopped[d]) == (h in host_stop), (h, d)
        assert steps[d] == host_stop.get(h, coarse * nchecks)
        for f in fields:
            assert V.bits_equal(host_state[h][f], db.get(d, f)), (h, f)
    # a stopped event is no longer stepped by vh_batch_steps either, and its
    # state reads the same after an odd number of further steps (the state
    # buffers swap roles every step)
    h_stopped, stopped_id = next((h, d) for h, d in zip(hid, did) if h in host_stop)
    db.steps(coarse)
    db.sync()
    for f in fields:
        assert V.bits_equal(host_state[h_stopped][f], db.get(stopped_id, f)), f


@pytest.mark.parametrize("precision", [0, 1])
def test_nan_event_is_isolated(precision):
    """An event whose initial state holds a NaN (set from the host) reports
    VH_ERR_NAN through vh_batch_run with its id; removed, the others run on
    and end bit-identical to a batch that never held it."""
    lattice, bulk, coarse = True, True, 5
    print(f"case nan isolation precision={precision}", flush=True)
    evs = [EVENTS[0], EVENTS[1], EVENTS[2]]
    ics = [_ic(ev, lattice, bulk) for ev in evs]
    bad_ic = dict(ics[1], e=ics[1]["e"].copy())
    bad_ic["e"][len(bad_ic["e"]) // 2] = np.nan
    kw = dict(max_events=3, max_steps=100, bulk=bulk, eos_table=V.eos_table(), precision=precision)
    cap = sum(ev[0] * ev[1] * ev[2] for ev in evs)
    fields = ["e", "ut", "ux", "pixx", "Pi", "ttt"]

    b = Batch(cap, **kw)
    ids = [_add(b, ev, ic) for ev, ic in zip(evs, [ics[0], bad_ic, ics[2]])]
    st, stopped, steps = b.run(1e-9, coarse, 2)
    assert st == -4  # VH_ERR_NAN
    site, bad, cell = b.nan_info()
    assert bad == ids[1] and site != 0, (site, bad, cell)
    b.remove(bad)
    b.clear_error()
    st, stopped, steps = b.run(1e-9, coarse, 2)
    assert st == 0

    ref = Batch(cap, **kw)
    rid = [_add(ref, ev, ic) for ev, ic in zip([evs[0], evs[2]], [ics[0], ics[2]])]
    st, _, rsteps = ref.run(1e-9, coarse, 4)
    assert st == 0
    for i, r in zip([ids[0], ids[2]], rid):
        assert steps[i] == rsteps[r] == 4 * coarse
        for f in fields:
            assert V.bits_equal(b.get(i, f), ref.get(r, f)), (i, f)


def test_workload_events_fp64_bitwise_and_mixed_within_envelope():
    """Eight events of the batching workload's own generator (bench.py: MC-Glauber
    Au+Au, grids spread over 81^2..201^2, Latin-hypercube eta/s, tau0; table
    EoS, bulk, eta/s(T), regulation below 100 MeV): the FP64 batch equals the
    oracle bit for bit after 20 steps, every field of every event; the mixed
    batch is within the 1e-3 relative error the throughput metric is quoted
    at, in e and u^mu at every cell; and, on the 99.9th percentile of each
    event's pointwise errors, within twice the FP64 path's own sensitivity
    (FP64 from the FP32-rounded and two 1e-7-perturbed starts, largest over
    the events), with no more differing freezeout cubes.  The percentile, not
    the maximum: zeta/s(T) jumps by about 1 % at the joins of its
    parametrization (T / T_peak = 0.995 and 1.05), and a cell crossing a join
    one stage earlier or later than in FP64 moves that cell's Pi by ~5e-4
    relative.  In this set, event 3 (130^2) has such a cell at step 11
    (T = 168.03 MeV) for almost any FP32 arithmetic, a 2-ulp change of the
    face speeds included; its u^y maximum then reaches ~2.6 x the envelope."""
    import sys
    sys.path.insert(0, V.REPO)
    import bench as B
    table = V.eos_table()
    allev = B.make_batch_events(table, 256)
    events = sorted(allev, key=lambda ev: ev["N"])[::32][:8]
    assert events[0]["n"] < 100 and events[-1]["n"] > 180
    nsteps = 20
    print("case workload events", [ev["n"] for ev in events], flush=True)
    b = Batch(sum(ev["N"] for ev in events), max_events=8, max_steps=nsteps, bulk=True, eos_table=table)
    ids = [B.add_event(b, ev) for ev in events]
    b.steps(nsteps)
    b.sync()
    for ev, eid in zip(events, ids):
        n, N = ev["n"], ev["N"]
        o = V.Oracle(n, n, 1, B.BATCH["dx"], B.BATCH["dx"], B.BATCH["dx"], B.BATCH["dt"], ev["t0"], ev["etabar"],
                     T_reg_GeV=B.BATCH["T_reg_GeV"], bulk=True, table=table, etas_Tk_GeV=B.BATCH["etas_Tk_GeV"],
                     etas_slope_lo=B.BATCH["etas_slope_lo"], etas_slope_hi=B.BATCH["etas_slope_hi"])
        o.set_ic(B.full_ic(ev["ic"], N))
        o.init_conserved(ev["t0"])
        for s in range(1, nsteps + 1):
            assert o.step(ev["t0"] + (s - 1) * B.BATCH["dt"]) == 0
        for f in _fields(True):
            a, c = o.get(f), b.get(eid, f)
            assert V.bits_equal(a, c), f"event {n}^2 {f}: max rel {V.max_rel(a, c):.3e}"
    b.close()
    acc = B.accuracy_all(events, table, checkpoints=(nsteps,))
    r = acc["report"][0]
    print(r["mixed_vs_fp64_max_over_events"], r["fp64_envelope_max_over_events"], flush=True)
    print({k: v for k, v in r["mixed_vs_fp64_max_over_events"].items() if k.endswith("_q999")},
          {k: v for k, v in r["fp64_envelope_max_over_events"].items() if k.endswith("_q999")}, flush=True)
    for f in ("e_pt", "ut_pt", "ux", "uy"):
        assert r["mixed_vs_fp64_max_over_events"][f] <= 1e-3, f
    for f in ("e", "e_pt", "ut_pt", "ux", "uy"):
        q = f + "_q999"
        assert r["mixed_vs_fp64_max_over_events"][q] <= 2 * r["fp64_envelope_max_over_events"][q], q
    assert r["fo_cubes_differing_mixed"] <= r["fo_cubes_differing_fp64_envelope"]


PI_NAMES = ["pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"]


@pytest.mark.parametrize("precision", [0, 1])
def test_admission_state_before_any_step(precision):
    """vh_batch_add enqueues its uploads and the initial-state kernel and
    returns without waiting.  Events admitted back to back (sizes that grow,
    so the staging buffers are replaced; more than two, so the pinned slots
    are reused; some with every array, some with e and p only), one removed
    and admitted again into the freed slot and cells, all read back before
    any step: e, p, u and the previous u, pi and Pi as given (u^tau one and
    the rest zero where an array is not given), and T^{tau mu} =
    (e + p + Pi) u^tau u^mu - (p + Pi) g^{tau mu} + pi^{tau mu} as
    setConservedVariables forms it (DynamicalVariables.cpp:169-218) - bit for
    bit in FP64; in the mixed form the fields rounded to FP32 and T^{tau mu}
    to 1e-13 (its FP32 high and low parts)."""
    print(f"case admission precision={precision}", flush=True)
    order = sorted(range(len(EVENTS)), key=lambda k: np.prod(EVENTS[k][:3]))
    evs = [EVENTS[k] for k in order]
    ics = []
    for k, ev in enumerate(evs):
        ic = _ic(ev, True, True)
        ics.append(ic if k % 2 == 0 else {"e": ic["e"], "p": ic["p"]})
    cap = sum(ev[0] * ev[1] * ev[2] for ev in evs) + evs[1][0] * evs[1][1] * evs[1][2]
    b = Batch(cap, max_events=len(evs) + 1, max_steps=16, bulk=True, eos_table=V.eos_table(), precision=precision)
    ids = [_add(b, ev, ic) for ev, ic in zip(evs, ics)]
    b.remove(ids[1])
    ids[1] = _add(b, evs[1], ics[1])
    f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if precision else (lambda a: a)
    for ev, ic, eid in zip(evs, ics, ids):
        N = ev[0] * ev[1] * ev[2]
        z = np.zeros(N)
        u = [ic.get(k, np.ones(N) if k == "ut" else z) for k in ("ut", "ux", "uy", "un")]
        pi = ic["pi"].reshape(10, N) if "pi" in ic else np.zeros((10, N))
        Pi = ic.get("Pi", z)
        want = {"e": ic["e"], "p": ic["p"], "Pi": Pi}
        for n, k in enumerate(("t", "x", "y", "n")):
            want["u" + k] = u[n]
            want["up_" + k] = u[n]
        for n, name in enumerate(PI_NAMES):
            want[name] = pi[n]
        for f, a in want.items():
            assert V.bits_equal(b.get(eid, f), f32(a)), (eid, f)
        pp = ic["p"] + Pi
        T = [(ic["e"] + pp) * u[0] * u[0] - pp + pi[0]] + [(ic["e"] + pp) * u[0] * u[n] + pi[n] for n in (1, 2, 3)]
        for n, name in enumerate(("ttt", "ttx", "tty", "ttn")):
            got = b.get(eid, name)
            if precision:
                assert np.all(np.abs(got - T[n]) <= 1e-13 * np.maximum(np.abs(T[n]), 1e-300)), (eid, name)
            else:
                assert V.bits_equal(got, T[n]), (eid, name)
    b.close()
