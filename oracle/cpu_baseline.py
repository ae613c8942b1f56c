"""CPU baseline leg of bench.py (test infrastructure; run as a child process).

Times the reference's own rungeKutta2 (oracle/_ref/libcpuvh_ref_<variant>.so,
built from the cpu-vh sources with OpenMP) on the initial state bench.py
hands over, for a bounded number of steps, and prints one JSON line.  It runs
in its own process because the reference calls exit(-1) on a NaN.

usage: python oracle/cpu_baseline.py IC.npz VARIANT NX NY NZ DX DT T0 ETABAR TREG MIN_SECONDS MAX_STEPS
"""
import ctypes
import json
import os
import sys
import time

import numpy as np


def main():
    (ic_path, variant, nx, ny, nz, dx, dt, t0, etabar, treg, min_s, max_steps) = sys.argv[1:13]
    nx, ny, nz, max_steps = int(nx), int(ny), int(nz), int(max_steps)
    dx, dt, t0, etabar, treg, min_s = map(float, (dx, dt, t0, etabar, treg, min_s))
    here = os.path.dirname(os.path.abspath(__file__))
    repo = os.path.dirname(here)
    L = ctypes.CDLL(os.path.join(here, "_ref", f"libcpuvh_ref_{variant}.so"))
    D, I, VP = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    L.ref_setup.argtypes = [I, I, I, I, D, D, D, D, I, D, D, D, D, ctypes.c_char_p]
    L.ref_set_ic.argtypes = [VP] * 8
    L.ref_init_conserved.argtypes = [D]
    L.ref_step.argtypes = [D]
    ic = np.load(ic_path)
    if L.ref_setup(nx, ny, nz, 100000, dx, dx, dx, dt, 1, t0, etabar, 0.15, treg, repo.encode()) != 0:
        raise SystemExit("ref_setup failed")
    arrs = [np.ascontiguousarray(ic[k]) for k in ("e", "p", "ut", "ux", "uy", "un", "pi", "Pi")]
    L.ref_set_ic(*[a.ctypes.data for a in arrs])
    L.ref_init_conserved(t0)
    L.ref_step(t0)  # first step untimed (page faults, thread start-up)
    n, t = 1, time.perf_counter()
    start = t
    while True:
        L.ref_step(t0 + n * dt)
        n += 1
        el = time.perf_counter() - start
        if (el >= min_s and n - 1 >= 2) or n - 1 >= max_steps:
            break
    steps = n - 1
    cells = nx * ny * nz
    print(json.dumps({"steps": steps, "seconds": el, "zone_upd_per_s": cells * steps / el,
                      "threads": int(os.environ.get("OMP_NUM_THREADS", "0") or 0)}))


if __name__ == "__main__":
    main()
