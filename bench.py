#!/usr/bin/env python3
"""Benchmark of the MI355X viscous-hydro step (one GPU).

A step is one RK2 step (rungeKutta2 + buffer swap) of the plain FP64 drop-in
over one event.  Default workload: the 3+1D single event at real size,
131x131x61 cells (dx=dy=deta=0.1), tau0=0.6 fm/c, dtau=0.02 fm/c, seeded
MC-Glauber Au+Au initial energy density, HotQCD+HRG table EoS, eta/s(T) with
minimum 0.08 at 154 MeV, bulk pressure evolved, regulation below 100 MeV.
Prints one JSON line.
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
HBARC = 0.197326938

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


def cpu_baseline(w, ic, threads, min_seconds=12.0, max_steps=40):
    variant = ("lat" if w["lattice"] else "conf") + ("_bulk" if w["bulk"] else "")
    lib = os.path.join(REPO, "oracle", "_ref", f"libcpuvh_ref_{variant}.so")
    if not os.path.exists(lib):
        return None
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "ic.npz")
        np.savez(path, **ic)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        cmd = [sys.executable, os.path.join(REPO, "oracle", "cpu_baseline.py"), path, variant, str(w["nx"]),
               str(w["ny"]), str(w["nz"]), str(w["dx"]), str(w["dt"]), str(w["t0"]), str(w["etabar"]),
               str(w["T_reg_GeV"]), str(min_seconds), str(max_steps)]
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
        lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode != 0 or not lines:
            return {"error": (r.stderr or r.stdout)[-300:]}
        res = json.loads(lines[-1])
    return {"value": res["zone_upd_per_s"], "unit": "zone-upd/s", "cores": threads, "kind": "reference",
            "sample": f"{res['steps']} RK2 steps of the same event after one untimed step, "
                      f"{res['seconds']:.1f} s wall, OpenMP reference build (-O3 -ffp-contract=off), "
                      f"constant eta/s={w['etabar']} (the reference has no eta/s(T))"}


def load_traffic(name):
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="glauber3d", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(events shard across GPUs with no communication, see DESIGN.md)", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: one GPU only", file=sys.stderr)
        return 2

    import cpuvh_amd as C
    w = WORKLOADS[args.workload]
    table = C.load_eos_table(os.path.join(REPO, "EoS", "hrg_hotqcd_eos_binary.dat"))
    ic, info = make_ic(w, table)
    nx, ny, nz, dx = w["nx"], w["ny"], w["nz"], w["dx"]
    N = nx * ny * nz
    h = C.Hydro(nx, ny, nz, dx, dx, dx, w["dt"], w["t0"], w["etabar"], T_reg_GeV=w["T_reg_GeV"], bulk=w["bulk"],
                eos_table=table if w["lattice"] else None, etas_Tk_GeV=w["etas_Tk_GeV"],
                etas_slope_lo=w["etas_slope_lo"], etas_slope_hi=w["etas_slope_hi"])
    h.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic["Pi"])
    h.init_conserved(w["t0"])
    if args.warmup > 0:
        h.steps(1, args.warmup)
    h.sync()
    h.prepare_steps(args.steps)
    t_start = time.perf_counter()
    h.steps(1 + args.warmup, args.steps)
    h.sync()
    elapsed = time.perf_counter() - t_start
    kernel_ms = h.last_kernel_ms()
    h.close()

    nc = 15 if w["bulk"] else 14
    bytes_per_zone = (5 * nc + 32) * 8  # read/write of the fused two-kernel step
    bytes_per_launch = bytes_per_zone / 2 * N
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else None
    ms_per_step = elapsed / args.steps * 1e3
    zone_rate = N * args.steps / elapsed
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(w, ic, threads)
    traffic = load_traffic(f"stage_kernel_{args.workload}_pmc.json")
    out = {
        "metric": METRIC, "value": zone_rate, "unit": "zone-upd/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded MC-Glauber initial state)",
        "config": {"workload": "3+1D single event 131x131x61, dx=deta=0.1, tau0=0.6, dtau=0.02, 600 steps/event, "
                               "MC-Glauber Au+Au, HotQCD+HRG table EoS, eta/s(T) min 0.08 @154 MeV, bulk on, "
                               "FP64 plain drop-in",
                   "cells": N, "glauber": info, "form": "plain FP64"},
        "events_per_s": 1.0 / (ms_per_step * 1e-3 * w["steps_per_event"]),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "kernel": "stage_kernel (one RK stage, fused)", "avg_launch_ms": kernel_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
