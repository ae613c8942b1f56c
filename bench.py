#!/usr/bin/env python3
"""Benchmark of the MI355X viscous-hydro step (one GPU).  Prints one JSON line.

Headline (value): events/s of the batched mixed-precision form on the
batching workload -- 256 seeded 2+1D boost-invariant MC-Glauber events with
square grids drawn from 81^2..201^2 (dx = 0.1 fm, dtau = 0.02 fm/c) and
Latin-hypercube per-event parameters ((eta/s)_min in [0.04, 0.2], zeta/s peak
in [0, 0.1] at 160 MeV, tau0 in [0.3, 1.0] fm/c; bulk viscosity acting through
the EoS table's sound speed), all resident in one batch and stepped
together until each is below T_F = 150 MeV (the third coarse check, every 5
steps, with no cell above e_F -- the reference never resets that count;
finished events leave the batch).  A step of --steps/--warmup is one pass of
that workload: admission of every event (upload and conserved variables),
then the run to the last event's freezeout; value leaves out the
admission's host-side staging of the initial states (inputs resident in HBM),
which batched_mixed.events_per_s includes.
Beside it: the plain FP64 drop-in looping over events one at a time (a
sample of the same events, creation to release inside the timed region), the
same batch in FP64, the mixed form's error against FP64 on every event beside
the FP64 path's own sensitivity, the serving sweep, the 201x201x101
precision form, the 3+1D single event at real size (131x131x61, 600 steps)
through the plain FP64 path, the dominant kernel's roofline and the reference
CPU build timed on this host's cores.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cpu-vh_amd"))

METRIC = "events/s batched-FP32 vs plain FP64 (≤1e-3 rel err), zone-upd/s, p50/p99"
HBM_PEAK_GBS = 8000.0
# VALU issue: one wave instruction per CU per clock (four SIMDs, a wave64
# instruction every four clocks each), 256 CUs at 2.4 GHz
VALU_ISSUE_PER_S = 256 * 2.4e9
HBARC = 0.197326938

# Bulk viscosity that acts: c_s^2 from the EoS table (cs2_mode 1; the
# reference's own c_s^2 is 1/3 on this table, which zeroes every bulk term)
# and zeta/s(T) peaked at 160 MeV, with the batching workload's per-event
# height and DROPIN_ZETA_PEAK for the 3+1D events.  --reference-bulk turns it
# off (the reference's transport).
BULK_EXT = True
ZETA_TPEAK_GEV = 0.16
DROPIN_ZETA_PEAK = 0.1


def cs2_mode(reference=False):
    return 0 if reference or not BULK_EXT else 1


def zeta_kw(peak, reference=False):
    if reference or not BULK_EXT:
        return {}
    return dict(zeta_Tpeak_GeV=ZETA_TPEAK_GEV, zeta_peak=peak)


def bulk_text():
    return (f"bulk zeta/s(T) peaked at {ZETA_TPEAK_GEV * 1e3:.0f} MeV with c_s^2 from the table" if BULK_EXT
            else "bulk on with the reference's c_s^2 = 1/3 (bulk terms zero)")

WORKLOADS = {
    # 3+1D single event at real size (the plain FP64 drop-in configuration)
    "glauber3d": dict(nx=131, ny=131, nz=61, dx=0.1, dt=0.02, t0=0.6, steps_per_event=600, lattice=True,
                      bulk=True, etabar=0.08, etas_Tk_GeV=0.154, etas_slope_lo=0.5, etas_slope_hi=1.0,
                      T_reg_GeV=0.1, seed=12345, e0=10.0),
}


def table_pressure(table, e):
    """The reference's table pressure (EquationOfState.cpp:44-107,329-360),
    vectorised, for the initial p of the Glauber state."""
    raw = table.reshape(-1, 4)
    ev = raw[:, 0] / HBARC
    pt = raw[:, 1] / HBARC
    e_min, de = ev[0], ev[1] - ev[0]
    idx = np.clip(((e - e_min) / de).astype(np.int64), 0, len(ev) - 2)
    frac = (e - (idx * de + e_min)) / de
    return np.maximum(1e-15, pt[idx] * (1.0 - frac) + pt[idx + 1] * frac)


def make_ic(w, table):
    import cpuvh_amd as C
    nx, ny, nz, dx = w["nx"], w["ny"], w["nz"], w["dx"]
    e, info = C.mc_glauber(nx, ny, nz, dx, dx, dx, seed=w["seed"], e0=w["e0"])
    N = e.size
    p = table_pressure(table, e) if w["lattice"] else e / 3.0
    ic = {"e": e, "p": p, "ut": np.ones(N), "ux": np.zeros(N), "uy": np.zeros(N), "un": np.zeros(N),
          "pi": np.zeros(10 * N), "Pi": np.zeros(N)}
    return ic, info


def full_ic(ic, N):
    """Every initial-state array, the missing ones at rest (u^tau 1, others 0)."""
    d = {"ut": np.ones(N), "ux": np.zeros(N), "uy": np.zeros(N), "un": np.zeros(N), "pi": np.zeros(10 * N),
         "Pi": np.zeros(N)}
    d.update({k: v for k, v in ic.items() if v is not None})
    return d


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cases, seconds=10.0, max_steps=400):
    """The reference build (oracle/_ref, -O3 -ffp-contract=off) on this
    host: one single-threaded process per case, all running at once, each
    timing up to max_steps RK2 steps of its event for about `seconds` after
    one untimed step.  cases: [(workload dict, initial state)].  Returns the
    summed zone-updates/s and the per-process results."""
    procs = []
    with tempfile.TemporaryDirectory() as td:
        for k, (w, ic) in enumerate(cases):
            variant = ("lat" if w["lattice"] else "conf") + ("_bulk" if w["bulk"] else "")
            lib = os.path.join(REPO, "oracle", "_ref", f"libcpuvh_ref_{variant}.so")
            if not os.path.exists(lib):
                return None
            path = os.path.join(td, f"ic{k}.npz")
            np.savez(path, **ic)
            env = dict(os.environ, OMP_NUM_THREADS="1")
            cmd = [sys.executable, os.path.join(REPO, "oracle", "cpu_baseline.py"), path, variant, str(w["nx"]),
                   str(w["ny"]), str(w["nz"]), str(w["dx"]), str(w["dt"]), str(w["t0"]), str(w["etabar"]),
                   str(w["T_reg_GeV"]), str(seconds), str(max_steps)]
            procs.append(subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
        results = []
        for pr in procs:
            out, err = pr.communicate(timeout=600)
            lines = [l for l in out.splitlines() if l.startswith("{")]
            if pr.returncode != 0 or not lines:
                return {"error": (err or out)[-300:]}
            results.append(json.loads(lines[-1]))
    return {"zone_upd_per_s": sum(r["zone_upd_per_s"] for r in results), "processes": results}


def valu_issue(name, cell_stages, seconds):
    """The kernel's VALU issue rate against the chip's: wave instructions per
    cell-stage measured with rocprofv3 (SQ_INSTS_VALU, committed under
    profiles/) x the cell-stages the timed launches covered, over their time."""
    per = load_traffic(name, "valu_wave_insts_per_cell_stage")
    if not per or not seconds:
        return None
    rate = per * cell_stages / seconds
    return {"valu_wave_insts_per_cell_stage": per, "achieved_per_s": rate, "peak_per_s": VALU_ISSUE_PER_S,
            "frac": rate / VALU_ISSUE_PER_S, "source": name}


def load_traffic(name, key="hbm_bytes_per_launch"):
    """HBM bytes measured with rocprofv3 FETCH_SIZE/WRITE_SIZE (scaled by the
    copy calibration), committed under profiles/."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get(key)


# ------------------------------------------------------- batching workload --
BATCH = dict(n_events=256, n_min=81, n_max=201, dx=0.1, dt=0.02, seed=20240601, e0=10.0, e_floor=0.01, T_F_GeV=0.150,
             etas_Tk_GeV=0.154, etas_slope_lo=0.5, etas_slope_hi=1.0, T_reg_GeV=0.1, coarse=5, max_steps=3000)


def lhs(n, bounds, rng):
    """Latin-hypercube sample: one point per stratum in every dimension."""
    out = np.empty((n, len(bounds)))
    for d, (lo, hi) in enumerate(bounds):
        strata = (rng.permutation(n) + rng.random(n)) / n
        out[:, d] = lo + (hi - lo) * strata
    return out


TILES_OVERRIDE = None  # (2+1D, 3+1D) tile variants of the mixed kernel, when set from the command line
FACE_MODE = None  # face semantics of the mixed kernel when set (vh_batch_set_face_mode); default: the library's


def make_batch_events(table, n_events):
    """Seeded 2+1D boost-invariant MC-Glauber events of the batching workload."""
    import cpuvh_amd as C
    rng = np.random.default_rng(BATCH["seed"])
    sides = rng.integers(BATCH["n_min"], BATCH["n_max"] + 1, n_events)
    params = lhs(n_events, [(0.04, 0.2), (0.0, 0.1), (0.3, 1.0)], rng)
    events = []
    for k in range(n_events):
        n = int(sides[k])
        e, info = C.mc_glauber(n, n, 1, BATCH["dx"], BATCH["dx"], BATCH["dx"], seed=BATCH["seed"] + 7919 * k,
                               e0=BATCH["e0"], e_floor=BATCH["e_floor"])
        N = e.size
        # at rest without dissipative currents: only e and p cross the boundary
        ic = {"e": e, "p": table_pressure(table, e)}
        events.append(dict(n=n, N=N, etabar=float(params[k, 0]), zeta_peak=float(params[k, 1]),
                           t0=float(params[k, 2]), ic=ic, b=info["b"]))
    return events


def add_event(b, ev):
    d = BATCH["dx"]
    const = ev.get("const_etas", False)
    return b.add(ev["n"], ev["n"], 1, d, d, d, BATCH["dt"], ev["t0"], ev["etabar"], ev["ic"],
                 T_reg_GeV=BATCH["T_reg_GeV"], etas_Tk_GeV=BATCH["etas_Tk_GeV"],
                 etas_slope_lo=0.0 if const else BATCH["etas_slope_lo"],
                 etas_slope_hi=0.0 if const else BATCH["etas_slope_hi"], **zeta_kw(ev["zeta_peak"], const))


CHECKS_PER_SYNC = 8


def one_pass(b, events, coarse, max_batched_steps=None):
    """One pass of the batching workload through batch b: every event
    admitted (upload of its initial state, conserved variables formed on the
    device), all stepped together until each meets the stop rule, finished
    events removed.  Returns the pass's timings."""
    import cpuvh_amd as C
    t_start = time.perf_counter()
    ids = [add_event(b, ev) for ev in events]
    t_admitted = time.perf_counter()
    eF = b.eos_energy_density(BATCH["T_F_GeV"] / HBARC)
    cells = {i: ev["N"] for i, ev in zip(ids, events)}
    running = set(ids)
    done_steps = {i: 0 for i in ids}
    steps_done = {}
    zone_updates = 0
    graph_ms = 0.0
    batched_steps = 0
    failed = []
    # the stop rule runs on the device: one call holds CHECKS_PER_SYNC rounds
    # of (coarse steps, count, strike), and the host only removes the events
    # that stopped
    while running and (max_batched_steps is None or batched_steps < max_batched_steps):
        room = (BATCH["max_steps"] - max(done_steps[i] for i in running)) // coarse
        if room <= 0:  # step budget spent: the remaining events stop here
            for i in list(running):
                steps_done[i] = done_steps[i]
                b.remove(i)
            running.clear()
            break
        n = min(CHECKS_PER_SYNC, room)
        if max_batched_steps is not None:
            n = max(1, min(n, (max_batched_steps - batched_steps) // coarse))
        st, stopped, steps = b.run(eF, coarse, n)
        graph_ms += b.last_ms()
        batched_steps += coarse * n
        for i in running:
            zone_updates += cells[i] * (int(steps[i]) - done_steps[i])
            done_steps[i] = int(steps[i])
        if st == C.VH_ERR_NAN:  # an event hit a NaN: it leaves the batch, the others go on
            site, bad, cell = b.nan_info()
            b.clear_error()
            if bad in running:
                failed.append(ids.index(bad))
                b.remove(bad)
                running.discard(bad)
        for i in list(running):
            if stopped[i] == 1:
                steps_done[i] = done_steps[i]
                b.remove(i)
                running.discard(i)
    for i in list(running):  # max_batched_steps reached
        steps_done[i] = done_steps[i]
        b.remove(i)
    wall = time.perf_counter() - t_start
    return dict(wall_s=wall, admission_s=t_admitted - t_start, events=len(events) - len(failed),
                events_failed=len(failed), failed=sorted(failed), batched_steps=batched_steps,
                zone_updates=zone_updates, graph_ms=graph_ms,
                steps_to_freezeout=[None if k in failed else steps_done.get(i) for k, i in enumerate(ids)])


def run_batched(events, table, precision, passes, warmup, max_batched_steps=None, const_etas=False):
    """The batching workload: `warmup` untimed passes, then `passes` timed
    passes, each from admission of all events to the last one's freezeout,
    through one batch object.  const_etas: the reference's transport (each
    event's (eta/s)_min as a constant, the reference's c_s^2 and zeta/s)."""
    import cpuvh_amd as C
    if const_etas:
        events = [dict(ev, const_etas=True) for ev in events]
    cap = sum(ev["N"] for ev in events)
    kw = dict(max_events=len(events), max_steps=BATCH["max_steps"], bulk=True, eos_table=table,
              precision=precision, cs2_mode=cs2_mode(const_etas))
    tiles = (-1, -1)
    if precision == C.PREC_MIXED:
        if TILES_OVERRIDE is not None:
            tiles = TILES_OVERRIDE
        else:  # the tile autotuner on a throwaway batch holding the workload
            wb = C.Batch(cap, **kw)
            for ev in events:
                add_event(wb, ev)
            tiles = wb.autotune(reps=5)
            wb.close()
    b = C.Batch(cap, **kw)
    if precision == C.PREC_MIXED:
        b.set_tiles(max(tiles[0], 0), max(tiles[1], 0))
        if FACE_MODE is not None:
            b.set_face_mode(FACE_MODE)
    for _ in range(warmup):
        one_pass(b, events, BATCH["coarse"], max_batched_steps)
    runs = [one_pass(b, events, BATCH["coarse"], max_batched_steps) for _ in range(passes)]
    b.close()
    wall = sum(r["wall_s"] for r in runs)
    adm = sum(r["admission_s"] for r in runs)
    done = sum(r["events"] for r in runs)
    steps = sum(r["batched_steps"] for r in runs)
    zu = sum(r["zone_updates"] for r in runs)
    return dict(passes=passes, warmup_passes=warmup, wall_s=wall, admission_s=adm, events=done,
                events_failed=sum(r["events_failed"] for r in runs), failed_ids=runs[-1]["failed"],
                events_per_s=done / wall, events_per_s_excl_admission=done / (wall - adm),
                batched_steps=steps, batched_steps_per_pass=runs[-1]["batched_steps"],
                ms_per_batched_step=(wall - adm) / steps * 1e3, zone_updates=zu, zone_upd_per_s=zu / wall,
                graph_ms=sum(r["graph_ms"] for r in runs), steps_to_freezeout=runs[-1]["steps_to_freezeout"],
                pass_wall_s=[r["wall_s"] for r in runs], tiles=tiles)


def run_plain_loop(events, table):
    """The plain FP64 drop-in, one event at a time, same stop rule; the
    timed region holds each event's handle creation, upload, conserved
    variables, steps and release (setup_s: all but the steps and checks)."""
    import cpuvh_amd as C
    coarse = BATCH["coarse"]
    d = BATCH["dx"]
    zone_updates = 0
    steps = []
    failed = 0
    setup = 0.0
    t_start = time.perf_counter()
    for ev in events:
        t0 = time.perf_counter()
        h = C.Hydro(ev["n"], ev["n"], 1, d, d, d, BATCH["dt"], ev["t0"], ev["etabar"], T_reg_GeV=BATCH["T_reg_GeV"],
                    bulk=True, eos_table=table, etas_Tk_GeV=BATCH["etas_Tk_GeV"],
                    etas_slope_lo=BATCH["etas_slope_lo"], etas_slope_hi=BATCH["etas_slope_hi"],
                    cs2_mode=cs2_mode(), **zeta_kw(ev["zeta_peak"]))
        ic = ev["ic"]
        h.upload_ic(ic["e"], ic["p"], ic.get("ut"), ic.get("ux"), ic.get("uy"), ic.get("un"), ic.get("pi"),
                    ic.get("Pi"))
        h.init_conserved(ev["t0"])
        eF = h.eos_energy_density(BATCH["T_F_GeV"] / HBARC)
        setup += time.perf_counter() - t0
        n, strikes = 1, 0
        try:
            while strikes < 3 and n + coarse - 1 <= BATCH["max_steps"]:
                h.steps(n, coarse)
                n += coarse
                strikes += h.count_above(eF) == 0
            steps.append(n - 1)
        except C.VhError as ex:
            if ex.status != C.VH_ERR_NAN:
                raise
            failed += 1
        zone_updates += ev["N"] * (n - 1)
        t1 = time.perf_counter()
        h.close()
        setup += time.perf_counter() - t1
    wall = time.perf_counter() - t_start
    done = len(events) - failed
    return dict(wall_s=wall, setup_s=setup, events=done, events_failed=failed, events_per_s=done / wall,
                events_per_s_excl_setup=done / (wall - setup), zone_updates=zone_updates,
                zone_upd_per_s=zone_updates / wall, steps_to_freezeout=steps)


ACC_FIELDS = ("e", "ut", "ux", "uy")
ENVELOPE_K = 2.0  # the attainable accuracy bar: mixed error <= K x the FP64 perturbed-start envelope


def accuracy_all(events, table, checkpoints=(50, 100, 200, 300)):
    """The mixed form against FP64 on every event of the workload, at fixed
    steps, beside the FP64 path's own sensitivity: FP64 from the initial state
    rounded to FP32 (the mixed form's stored state at t0) and from two states
    with e scaled by 1 + 1e-7 u (u uniform in [-1, 1]).  The envelope of an
    event's field is the largest error of those three runs against FP64 from
    the exact state.  The reference's root solver stops at a relative change
    of 1e-3 and switches formula at e = 0.1, so differences at the rounding
    level grow by orders of magnitude over hundreds of steps in any arithmetic
    that is not bit-identical to it.  All five batches stay resident and step
    together; the freezeout cubes are flagged at each checkpoint (previous
    slice: the last checkpoint's) and compared per event."""
    import cpuvh_amd as C
    cap = sum(ev["N"] for ev in events)
    nsteps = max(checkpoints)
    variants = {"fp64": (C.PREC_FP64, None), "mixed": (C.PREC_MIXED, None), "f32_ic": (C.PREC_FP64, "f32"),
                "pert1": (C.PREC_FP64, 1), "pert2": (C.PREC_FP64, 2)}
    batches = {}
    for name, (prec, pert) in variants.items():
        b = C.Batch(cap, max_events=len(events), max_steps=nsteps, bulk=True, eos_table=table, precision=prec,
                    cs2_mode=cs2_mode())
        if prec == C.PREC_MIXED and FACE_MODE is not None:
            b.set_face_mode(FACE_MODE)
        ids = []
        for k, ev in enumerate(events):
            if pert == "f32":
                ev = dict(ev, ic={f: v.astype(np.float32).astype(np.float64) for f, v in ev["ic"].items()})
            elif pert is not None:
                r = np.random.default_rng(1000 * pert + k)
                e = ev["ic"]["e"] * (1 + 1e-7 * r.uniform(-1, 1, ev["ic"]["e"].size))
                ev = dict(ev, ic=dict(ev["ic"], e=e, p=table_pressure(table, e)))
            ids.append(add_event(b, ev))
        batches[name] = (b, ids)
    eF = batches["fp64"][0].eos_energy_density(BATCH["T_F_GeV"] / HBARC)
    envs = ("f32_ic", "pert1", "pert2")
    report, done = [], 0
    dropped = {}  # event index -> (variant, step): a NaN in one run drops the event from the comparison
    for stop in checkpoints:
        for name, (b, ids) in batches.items():
            b.steps(stop - done)
            try:
                b.sync()
            except C.VhError as ex:
                if ex.status != C.VH_ERR_NAN:
                    raise
                _, bad, _ = b.nan_info()
                b.clear_error()
                b.remove(bad)
                dropped.setdefault(ids.index(bad), (name, stop))
        done = stop
        per_event = []
        for k in range(len(events)):
            if k in dropped:
                continue
            got = {name: {f: b.get(ids[k], f) for f in ACC_FIELDS} for name, (b, ids) in batches.items()}
            cubes = {name: set(b.fo_flag_ids(ids[k], eF).tolist()) for name, (b, ids) in batches.items()}
            ref = got["fp64"]
            row = {}
            for name in ("mixed",) + envs:
                err = {}
                for f in ACC_FIELDS:
                    a, c = ref[f], got[name][f]
                    d, scale = np.abs(a - c), float(np.max(np.abs(a))) or 1.0
                    err[f] = float(np.max(d)) / scale
                    err[f + "_q999"] = float(np.quantile(d, 0.999)) / scale
                    if f in ("e", "ut"):
                        rel = d / np.abs(a)
                        err[f + "_pt"] = float(np.max(rel))
                        err[f + "_pt_q999"] = float(np.quantile(rel, 0.999))
                err["cubes"] = len(cubes["fp64"] ^ cubes[name])
                row[name] = err
            row["cubes_fp64"] = len(cubes["fp64"])
            per_event.append(row)
        keys = [f for f in ACC_FIELDS] + ["e_pt", "ut_pt"]
        keys += [k + "_q999" for k in keys]
        mixed_max = {f: max(r["mixed"][f] for r in per_event) for f in keys}
        env = [{f: max(r[v][f] for v in envs) for f in keys + ["cubes"]} for r in per_event]
        env_max = {f: max(x[f] for x in env) for f in keys}
        within = {f: sum(r["mixed"][f] <= ENVELOPE_K * x[f] for r, x in zip(per_event, env)) for f in keys}
        cubes_mixed = [r["mixed"]["cubes"] for r in per_event]
        cubes_env = [x["cubes"] for x in env]
        report.append({
            "step": stop,
            "mixed_vs_fp64_max_over_events": mixed_max,
            "fp64_envelope_max_over_events": env_max,
            "events_within_k_envelope": within,
            "fo_cubes_fp64": sum(r["cubes_fp64"] for r in per_event),
            "fo_cubes_differing_mixed": sum(cubes_mixed),
            "fo_cubes_differing_fp64_envelope": sum(cubes_env),
            "events_with_differing_cubes_mixed": sum(c > 0 for c in cubes_mixed),
            "events_with_differing_cubes_fp64_envelope": sum(c > 0 for c in cubes_env),
            "events_mixed_cubes_beyond_envelope": sum(m > x for m, x in zip(cubes_mixed, cubes_env)),
            "fo_cubes_differing_mixed_per_event": cubes_mixed,
        })
    for b, _ in batches.values():
        b.close()
    n = len(events) - len(dropped)
    pass_metric = all(r["mixed_vs_fp64_max_over_events"][f] <= 1e-3 for r in report for f in ("e_pt", "ut_pt", "ux", "uy"))
    pass_metric = pass_metric and all(r["fo_cubes_differing_mixed"] == 0 for r in report)
    pass_env = all(r["events_within_k_envelope"][f] == n for r in report for f in ("e", "e_pt", "ut_pt", "ux", "uy"))
    pass_env = pass_env and all(r["events_mixed_cubes_beyond_envelope"] == 0 for r in report)
    pass_env_q = all(r["events_within_k_envelope"][f + "_q999"] == n for r in report
                     for f in ("e", "e_pt", "ut_pt", "ux", "uy"))
    return {"events": n, "events_dropped_nan": {str(k): v for k, v in dropped.items()}, "checkpoints": list(checkpoints),
            "bar_metric": "max relative error <= 1e-3 in e and u^mu (e, u^tau pointwise; u^x, u^y against the "
                          "field's largest magnitude) and identical freezeout cube sets, every event, every checkpoint",
            "pass_metric_bar": bool(pass_metric),
            "bar_attainable": f"per event and field: mixed error <= {ENVELOPE_K:g} x the largest error of FP64 "
                              "from three perturbed starts (FP32-rounded, two 1e-7 e perturbations), and no more "
                              "differing freezeout cubes than the largest of those runs",
            "pass_attainable_bar": bool(pass_env),
            "bar_attainable_q999": f"as bar_attainable on the 99.9th percentile of each event's pointwise errors "
                                   "(a cell whose temperature crosses a join of the zeta/s parametrization, where "
                                   "it jumps by about 1 %, in a different stage than FP64 changes the maximum only)",
            "pass_attainable_bar_q999": bool(pass_env_q), "report": report}


PRECISION_3D = dict(nx=201, ny=201, nz=101, dx=0.08, dt=0.02, t0=0.6, steps=600, etabar=0.08, etas_Tk_GeV=0.154,
                    etas_slope_lo=0.5, etas_slope_hi=1.0, T_reg_GeV=0.1, seed=777, e0=10.0)


def run_precision_3d(table, steps, checkpoints=(100, 300)):
    """The precision form's configuration: one 3+1D Glauber event on
    201 x 201 x 101 cells (dx = deta = 0.08) through the FP64 and the mixed
    form; per-field max and RMS relative errors and the freezeout cube sets
    at each checkpoint and at the end."""
    import cpuvh_amd as C
    w = PRECISION_3D
    nx, ny, nz, dx = w["nx"], w["ny"], w["nz"], w["dx"]
    e, info = C.mc_glauber(nx, ny, nz, dx, dx, dx, seed=w["seed"], e0=w["e0"])
    N = e.size
    ic = {"e": e, "p": table_pressure(table, e), "ut": np.ones(N), "ux": np.zeros(N), "uy": np.zeros(N),
          "un": np.zeros(N), "pi": np.zeros(10 * N), "Pi": np.zeros(N)}
    fields = ("e", "ut", "ux", "uy", "un", "pitt", "pixx", "pixy", "piyy", "pinn", "Pi")
    batches, wall = {}, {}
    for prec in (C.PREC_FP64, C.PREC_MIXED):
        b = C.Batch(N, max_events=1, max_steps=steps, bulk=True, eos_table=table, precision=prec,
                    cs2_mode=cs2_mode())
        eid = b.add(nx, ny, nz, dx, dx, dx, w["dt"], w["t0"], w["etabar"], ic, T_reg_GeV=w["T_reg_GeV"],
                    etas_Tk_GeV=w["etas_Tk_GeV"], etas_slope_lo=w["etas_slope_lo"], etas_slope_hi=w["etas_slope_hi"],
                    **zeta_kw(DROPIN_ZETA_PEAK))
        batches[prec] = (b, eid)
        wall[prec] = 0.0
    eF = batches[C.PREC_FP64][0].eos_energy_density(BATCH["T_F_GeV"] / HBARC)
    report, done = [], 0
    for stop in [c for c in checkpoints if c < steps] + [steps]:
        runs, cubes = {}, {}
        for prec, (b, eid) in batches.items():
            t = time.perf_counter()
            b.steps(stop - done)
            b.sync()
            wall[prec] += time.perf_counter() - t
            runs[prec] = {f: b.get(eid, f) for f in fields}
            cubes[prec] = set(b.fo_flag_ids(eid, eF).tolist())  # cubes cut by e_F (previous slice: this state)
        done = stop
        hot = runs[C.PREC_FP64]["e"] > eF
        out = {}
        for f in fields:
            a, c = runs[C.PREC_FP64][f], runs[C.PREC_MIXED][f]
            scale = float(np.max(np.abs(a))) or 1.0
            norm = float(np.sqrt(np.mean(a * a))) or 1.0
            r = {"max_abs_over_max": float(np.max(np.abs(a - c))) / scale,
                 "rms_rel": float(np.sqrt(np.mean((a - c) ** 2))) / norm}
            if f in ("e", "ut"):
                rel = np.abs(a - c) / np.abs(a)
                r["max_pointwise_rel"] = float(np.max(rel))
                r["max_pointwise_rel_above_eF"] = float(np.max(rel[hot])) if hot.any() else None
            out[f] = r
        sym = cubes[C.PREC_FP64] ^ cubes[C.PREC_MIXED]
        report.append({"step": stop, "cells_above_eF": int(hot.sum()), "fields": out,
                       "fo_cubes_fp64": len(cubes[C.PREC_FP64]), "fo_cubes_mixed": len(cubes[C.PREC_MIXED]),
                       "fo_cubes_differing": len(sym)})
    for b, _ in batches.values():
        b.close()
    return {"workload": "3+1D MC-Glauber event 201x201x101, dx=deta=0.08, tau0=0.6, dtau=0.02, table EoS, "
                        f"eta/s(T) min 0.08, {bulk_text()}"
                        + (f" (height {DROPIN_ZETA_PEAK})" if BULK_EXT else ""), "steps": steps, "cells": N, "glauber": info,
            "fp64_s": wall[C.PREC_FP64], "mixed_s": wall[C.PREC_MIXED], "checkpoints": report}


SERVE = dict(frac_3d=0.2, side2_min=101, side2_max=201, side3_min=101, side3_max=161, nz3_min=41, nz3_max=81,
             dx=0.1, dt=0.02, t0=0.6, etabar=0.08, etas_Tk_GeV=0.154, etas_slope_lo=0.5, etas_slope_hi=1.0,
             T_reg_GeV=0.1, T_F_GeV=0.150, coarse=5, e0=10.0, e_floor=0.01, hbm_budget_bytes=8 << 30,
             bulk=1, max_steps=3000, max_resident=256, collect_surface=2)


def run_serving(table, rates, n_requests, share_3d, budget, surface=2):
    """Open-loop serving at each arrival rate (vh_serve_run)."""
    import cpuvh_amd as C
    out = []
    for k, (rate, nreq) in enumerate(zip(rates, n_requests)):
        stats, reqs = C.serve(table, rate_per_s=float(rate), n_requests=nreq, seed=1000 + k,
                              precision=C.PREC_MIXED, **dict(SERVE, frac_3d=share_3d, hbm_budget_bytes=budget,
                                                             collect_surface=surface))
        done = [r for r in reqs if r["status"] == C.VH_OK]
        out.append({"rate_per_s": rate, "requests": nreq, "completed": stats["completed"],
                    "failed": stats["failed"], "events_per_s": stats["events_per_s"],
                    "zone_upd_per_s": stats["zone_upd_per_s"], "p50_s": stats["p50_s"], "p99_s": stats["p99_s"],
                    "mean_s": stats["mean_s"], "hbm_occupancy_mean": stats["occupancy_mean"],
                    "hbm_occupancy_peak": stats["occupancy_peak"],
                    "mean_wait_for_admission_s": float(np.mean([r["admit_s"] - r["arrival_s"] for r in done]))
                    if done else None,
                    "share_3d": float(np.mean([r["nz"] > 1 for r in reqs])),
                    "surface_elements_per_request_mean": float(np.mean([r["fo_elements"] for r in done]))
                    if done else None,
                    "failed_requests": [{"nx": r["nx"], "nz": r["nz"], "glauber_seed": r["glauber_seed"],
                                         "status": r["status"]} for r in reqs if r["status"] != C.VH_OK]})
    return out


def run_dropin_3d(w, table, steps, warmup):
    import cpuvh_amd as C
    ic, info = make_ic(w, table)
    nx, ny, nz, dx = w["nx"], w["ny"], w["nz"], w["dx"]
    N = nx * ny * nz
    h = C.Hydro(nx, ny, nz, dx, dx, dx, w["dt"], w["t0"], w["etabar"], T_reg_GeV=w["T_reg_GeV"], bulk=w["bulk"],
                eos_table=table if w["lattice"] else None, etas_Tk_GeV=w["etas_Tk_GeV"],
                etas_slope_lo=w["etas_slope_lo"], etas_slope_hi=w["etas_slope_hi"], cs2_mode=cs2_mode(),
                **zeta_kw(DROPIN_ZETA_PEAK))
    h.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic["Pi"])
    h.init_conserved(w["t0"])
    if warmup > 0:
        h.steps(1, warmup)
    h.sync()
    h.prepare_steps(steps)
    t_start = time.perf_counter()
    h.steps(1 + warmup, steps)
    h.sync()
    elapsed = time.perf_counter() - t_start
    kernel_ms = h.last_kernel_ms()
    h.close()
    nc = 15 if w["bulk"] else 14
    bytes_per_zone = (5 * nc + 32) * 8  # read/write of the fused two-kernel step
    bytes_per_launch = bytes_per_zone / 2 * N
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else None
    return ic, {
        "workload": "3+1D single event 131x131x61, dx=deta=0.1, tau0=0.6, dtau=0.02, MC-Glauber Au+Au, "
                    f"HotQCD+HRG table EoS, eta/s(T) min 0.08 @154 MeV, {bulk_text()}"
                    + (f" (height {DROPIN_ZETA_PEAK})" if BULK_EXT else ""),
        "cells": N, "glauber": info, "steps": steps, "warmup": warmup, "ms_per_step": elapsed / steps * 1e3,
        "zone_upd_per_s": N * steps / elapsed, "events_per_s_at_600_steps": 1.0 / (elapsed / steps * 600),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": load_traffic("stage_kernel_glauber3d_pmc.json"),
                     "kernel": "stage_kernel (one RK stage, FP64, fused)", "avg_launch_ms": kernel_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "valu_issue": valu_issue("stage_kernel_glauber3d_pmc.json", N, kernel_ms * 1e-3)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed passes of the batching workload (each: admission of every event, all stepped "
                         "to freezeout)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before them")
    ap.add_argument("--events", type=int, default=BATCH["n_events"])
    ap.add_argument("--plain-sample", type=int, default=32, help="events of the plain FP64 loop")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-procs", type=int, default=0, help="concurrent single-thread reference processes")
    ap.add_argument("--no-fp64-batch", action="store_true")
    ap.add_argument("--skip-dropin", action="store_true")
    ap.add_argument("--dropin-steps", type=int, default=WORKLOADS["glauber3d"]["steps_per_event"],
                    help="timed steps of the 131x131x61 FP64 line (600: the whole run; fewer for profiling)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-accuracy", action="store_true")
    ap.add_argument("--skip-precision-3d", action="store_true")
    ap.add_argument("--max-batch-steps", type=int, default=None,
                    help="stop the batched runs after this many steps (profiling; events/s then not to freezeout)")
    ap.add_argument("--serve-rates", default="1,10,50,200", help="arrival rates of the serving sweep (events/s)")
    ap.add_argument("--serve-requests", default="6,16,32,48", help="requests at each rate")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--serve-3d-share", type=float, default=0.2, help="share of 3+1D requests in the serving sweep")
    ap.add_argument("--serve-budget-gib", type=float, default=8.0, help="HBM budget of the serving loop")
    ap.add_argument("--serve-surface", type=int, default=2,
                    help="1: flag the freezeout cubes of each request at every check; 2: also build its surface "
                         "elements (surface.dat's rows) on the device and return them")
    ap.add_argument("--tiles", default=None,
                    help="mixed-kernel tile variants 'T2,T3' for the timed batch (counter runs without the autotuner)")
    ap.add_argument("--face-mode", type=int, default=None, help="face semantics of the mixed kernel (0, 1, 2)")
    ap.add_argument("--reference-bulk", action="store_true",
                    help="the reference's c_s^2 (1/3 on this table: bulk terms zero) instead of the table's sound "
                         "speed with zeta/s(T) peaked at 160 MeV")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 512 MiB device to device (scale of the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 8, "copy_probe_bytes_written": dst.numel() * 8}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(events shard across GPUs with no communication)", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: one GPU only", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("bench.py: --steps must be at least 1 and --warmup at least 0", file=sys.stderr)
        return 2

    import cpuvh_amd as C
    global TILES_OVERRIDE, FACE_MODE, BULK_EXT
    BULK_EXT = not args.reference_bulk
    if args.face_mode is not None:
        FACE_MODE = args.face_mode
    if args.tiles:
        TILES_OVERRIDE = tuple(int(x) for x in args.tiles.split(","))
    table = C.load_eos_table(os.path.join(REPO, "EoS", "hrg_hotqcd_eos_binary.dat"))
    log = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
    log("building initial states")
    events = make_batch_events(table, args.events)
    log(f"batched mixed: {args.warmup} untimed + {args.steps} timed passes")
    mixed = run_batched(events, table, C.PREC_MIXED, args.steps, args.warmup, args.max_batch_steps)
    fp64b = None
    if not args.no_fp64_batch:
        log("batched fp64")
        fp64b = run_batched(events, table, C.PREC_FP64, 1, 0, args.max_batch_steps)
    stride = max(1, len(events) // max(1, args.plain_sample))
    sample = events[::stride][:args.plain_sample]
    plain = None
    if not args.skip_plain:
        log("plain fp64 loop")
        plain = run_plain_loop(sample, table)
    acc = None
    if not args.skip_accuracy:
        log("accuracy over every event")
        acc = accuracy_all(events, table)
    serving = None
    if not args.skip_serve:
        log("serving sweep")
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        nreq = [int(x) for x in args.serve_requests.split(",") if x]
        nreq = (nreq + nreq[-1:] * len(rates))[:len(rates)]
        serving = run_serving(table, rates, nreq, args.serve_3d_share, int(args.serve_budget_gib * (1 << 30)),
                              args.serve_surface)
    precision3d = None
    if not args.skip_precision_3d:
        log("precision form 3+1D 201x201x101")
        precision3d = run_precision_3d(table, PRECISION_3D["steps"])
    dropin = None
    if not args.skip_dropin:
        log(f"3+1D drop-in 131x131x61, {args.dropin_steps} steps")
        _, dropin = run_dropin_3d(WORKLOADS["glauber3d"], table, args.dropin_steps, 20 if args.dropin_steps >= 100 else 0)

    # dominant kernel: the mixed stage kernel over the batch (two launches per
    # step).  Algorithmic bytes of one zone-update (two stages): each stage
    # reads q (nl FP32 fields + the low parts of T^{tau mu}), u, the previous
    # u, e and p, and writes q, u, e and p; stage 2 also reads q^n.  The
    # boost-invariant kernel moves the 11 fields without eta components, 3
    # low parts and 3 velocity components; the general one 15, 4 and 4.
    bi = all((ev["ic"].get("un") is None or not np.any(ev["ic"]["un"])) and
             (ev["ic"].get("pi") is None or not np.any(ev["ic"]["pi"].reshape(10, -1)[[3, 6, 8]]))
             for ev in events)
    nl, nlo, nu = (11, 3, 3) if bi else (15, 4, 4)
    bytes_per_zone_mx = (2 * (nl + nlo + 2 * nu + 2) + 2 * (nl + nlo + nu + 2) + (nl + nlo)) * 4
    achieved = mixed["zone_updates"] * bytes_per_zone_mx / (mixed["graph_ms"] * 1e-3) / 1e9
    launches = 2 * mixed["batched_steps"]
    per_cell = load_traffic("stage_mixed_kernel_batch_pmc.json", "hbm_bytes_per_cell_stage")
    traffic_mx = per_cell * mixed["zone_updates"] / mixed["batched_steps"] if per_cell else None
    cpu, gpu_const = None, None
    if not args.no_cpu:
        nproc = args.cpu_procs or min(16, len(os.sched_getaffinity(0)))
        log(f"cpu reference: {nproc} single-thread processes")
        by_size = sorted(events, key=lambda ev: ev["N"])
        picks = [by_size[int((k + 0.5) * len(by_size) / nproc)] for k in range(nproc)]
        cases = [(dict(nx=ev["n"], ny=ev["n"], nz=1, dx=BATCH["dx"], dt=BATCH["dt"], t0=ev["t0"], lattice=True,
                       bulk=True, etabar=ev["etabar"], T_reg_GeV=BATCH["T_reg_GeV"]), full_ic(ev["ic"], ev["N"]))
                 for ev in picks]
        res = cpu_baseline(cases, seconds=10.0, max_steps=400)
        log("batched mixed with the reference's transport (the CPU comparison's)")
        gpu_const = run_batched(events, table, C.PREC_MIXED, 1, 0, args.max_batch_steps, const_etas=True)
        if res and "zone_upd_per_s" in res:
            work = float(np.mean([ev["N"] * s for ev, s in zip(events, gpu_const["steps_to_freezeout"])
                                  if s is not None]))
            cpu = {"value": res["zone_upd_per_s"], "unit": "zone-upd/s", "cores": nproc, "kind": "reference",
                   "cpu_model": cpu_model(), "logical_cpus_available": len(os.sched_getaffinity(0)),
                   "events_per_s": res["zone_upd_per_s"] / work,
                   "sample": f"{nproc} single-thread processes of the reference build (oracle/_ref, -O3 "
                             "-ffp-contract=off) at once, one event of the workload each (spread over its sizes), "
                             "constant eta/s = the event's (eta/s)_min and c_s^2 = 1/3 (the reference has neither "
                             "eta/s(T) nor an acting bulk term), ~10 s of "
                             "steps each after one untimed step; events/s = summed zone-upd/s / the workload's mean "
                             f"zone-updates per event to freezeout ({work:.3g}, from the GPU run with the same "
                             "transport)",
                   "per_process": [{"cells": c[0]["nx"] ** 2, "steps": r["steps"], "zone_upd_per_s": r["zone_upd_per_s"]}
                                   for c, r in zip(cases, res["processes"])]}
        elif res:
            cpu = res
    # value: the initial states already on the device -- the passes' wall time
    # less the admission's host-side staging (vh_batch_add copies the host
    # arrays to pinned memory and queues their upload); the plain loop's ratio
    # leaves out its set-up alike.  With them: *_incl_setup.
    value = mixed["events_per_s_excl_admission"]
    ratio = None if plain is None else value / plain["events_per_s_excl_setup"]
    out = {
        "metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": (mixed["wall_s"] - mixed["admission_s"]) / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 storage/flux/source + f64 root solve/EoS/accumulation",
        "data": "synthetic (seeded MC-Glauber initial states)",
        "config": {"workload": f"{len(events)} 2+1D MC-Glauber Au+Au events, grids {BATCH['n_min']}^2..{BATCH['n_max']}^2, "
                               "dx=0.1, dtau=0.02, LHS (eta/s)_min [0.04,0.2], zeta/s peak [0,0.1], tau0 [0.3,1.0], "
                               f"HotQCD+HRG table EoS, {bulk_text()}, eta/s(T) min at 154 MeV, "
                               "to freezeout at T=150 MeV "
                               "(coarse check every 5 steps, 3 strikes), one batch, mixed precision",
                   "step": "one pass of the workload: every event admitted (initial state uploaded, conserved "
                           "variables formed), all advanced together to freezeout, finished events removed; "
                           "value leaves out the admission's host-side staging of the initial states (inputs "
                           "resident in HBM), batched_mixed.events_per_s holds it",
                   "cells_total": int(sum(ev["N"] for ev in events)), "form": "batched mixed precision",
                   "tiles": mixed["tiles"]},
        "accuracy": None if acc is None else {k: acc[k] for k in ("bar_metric", "pass_metric_bar", "bar_attainable",
                                                                   "pass_attainable_bar", "bar_attainable_q999",
                                                                   "pass_attainable_bar_q999")},
        "batched_mixed": {k: v for k, v in mixed.items() if k != "steps_to_freezeout"},
        "plain_fp64_loop": None if plain is None else {k: v for k, v in plain.items() if k != "steps_to_freezeout"},
        "speedup_vs_plain_fp64": ratio,
        "speedup_vs_plain_fp64_incl_setup": None if plain is None else
        mixed["events_per_s"] / plain["events_per_s"],
        "batched_fp64": None if fp64b is None else {k: v for k, v in fp64b.items() if k != "steps_to_freezeout"},
        "mixed_vs_fp64_error": acc,
        "precision_form_3d": precision3d,
        "serving": None if serving is None else {
            "workload": f"Poisson arrivals of seeded MC-Glauber Au+Au events, {100 * (1 - args.serve_3d_share):.0f}% "
                        f"2+1D (101^2-201^2), {100 * args.serve_3d_share:.0f}% 3+1D (101x101x41 - 161x161x81), "
                        f"dx=0.1, table EoS, bulk, mixed precision, admitted into one batch by a "
                        f"{args.serve_budget_gib:g} GiB HBM budget, resident until no cell is above T=150 MeV, "
                        + ("freezeout surface (surface.dat's rows) built on the device and returned per request at "
                           "every check" if args.serve_surface == 2 else "freezeout cubes flagged per request at every check"),
                        "sweep": serving},
        "fp64_dropin_3d": dropin,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_mx,
                     "traffic_note": "HBM bytes per average launch: bytes per cell-stage measured with rocprofv3 "
                                     "(profiles/stage_mixed_kernel__batch256_2d.md) x the run's mean cells per launch",
                     "kernel": "stage_mixed_kernel (one RK stage, all 2+1D events"
                               + (", boost-invariant)" if bi else ")"),
                     "avg_launch_ms": mixed["graph_ms"] / launches,
                     "algorithmic_bytes_per_zone_update": bytes_per_zone_mx,
                     "valu_issue": valu_issue("stage_mixed_kernel_batch_pmc.json", 2 * mixed["zone_updates"],
                                              mixed["graph_ms"] * 1e-3)},
        "cpu_baseline": cpu,
        "batched_mixed_reference_transport": None if gpu_const is None else
        {k: v for k, v in gpu_const.items() if k != "steps_to_freezeout"},
    }
    if cpu and "events_per_s" in cpu and gpu_const:
        out["speedup_vs_cpu_reference"] = gpu_const["events_per_s_excl_admission"] / cpu["events_per_s"]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
