"""Writes the full-run fixtures of the plain FP64 path: the oracle's state
after every checkpoint step and its freezeout cube sets at every coarse step,
for a whole run of a workload, as hashes plus sampled values.

    python tests/golden/make_fullrun.py glauber3d    # 131x131x61, 600 steps
    python tests/golden/make_fullrun.py gubser       # 201x201x1, 400 steps

The run follows the reference's loop (rhic/src/HydroPlugin.cpp:226-308): for
n = 1 ... steps, when n % coarse == 0 the previous coarse slice of e is
replaced by the current one (the first previous slice is zero,
HydroPlugin.cpp:159), the cubes cut by e_F are flagged and the cells above
e_F counted; then one RK2 step at t = t0 + (n-1) dt.  All `steps` steps are
taken (the stop rule is recorded, not applied), so every checkpoint exists.

Output: tests/golden/fullrun_<name>.npz (numpy arrays only; load with
allow_pickle=False).  tests/test_gpu_fullrun.py replays the same run on the
GPU and compares field hashes, samples and cube sets.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, REPO)

import fullrun_cases as FC  # noqa: E402
import vhtest as V  # noqa: E402


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


def oracle_run(name, made=None, log=print):
    """The oracle's run of case `name` as the fixture's arrays."""
    case = FC.CASES[name]
    ic, kw, eos = made if made is not None else FC.make_case(name)
    nx, ny, nz = case["nx"], case["ny"], case["nz"]
    dx, dt, t0, steps, coarse = case["dx"], case["dt"], case["t0"], case["steps"], case["coarse"]
    table = V.eos_table() if eos == "table" else None
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, case["etabar"], table=table, **kw)
    eF = FC.oracle_eF(case, table)
    o.set_ic(ic)
    o.init_conserved(t0)
    N = nx * ny * nz
    rng = np.random.default_rng(20240607)
    samples = np.sort(rng.choice(N, size=min(256, N), replace=False)).astype(np.int64)
    fields = FC.FIELDS_BULK if kw.get("bulk") else FC.FIELDS
    ck_hash, ck_samp, ck_max = [], [], []
    fo_steps, fo_count, fo_hash, above = [], [], [], []
    prev = np.zeros(N)
    strikes, stop_at = 0, -1
    t_start = time.time()
    for n in range(1, steps + 1):
        if n % coarse == 0:
            cur = o.get("e")
            ids = V.oracle_fo_flag(prev, cur, nx, ny, nz, eF)
            prev = cur
            fo_steps.append(n)
            fo_count.append(ids.size)
            fo_hash.append(sha(ids.astype("<i8")))
            c = int(np.count_nonzero(cur > eF))
            above.append(c)
            if c == 0:
                strikes += 1
                if strikes >= 3 and stop_at < 0:
                    stop_at = n
        st = o.step(t0 + (n - 1) * dt)
        if st != 0:
            raise RuntimeError(f"oracle step {n} failed: {st}")
        if n in case["checkpoints"]:
            hs, sm, mx = [], [], []
            for f in fields:
                a = o.get(f)
                hs.append(sha(a.view("<u8")))
                sm.append(a[samples])
                mx.append(np.max(np.abs(a)))
            ck_hash.append(hs)
            ck_samp.append(sm)
            ck_max.append(mx)
            log(f"{name}: step {n} ({time.time() - t_start:.0f} s)")
    meta = dict(case=name, fields=fields, checkpoints=list(case["checkpoints"]), coarse=coarse, stop_at=stop_at,
                generator="tests/golden/make_fullrun.py", oracle="oracle/vh_oracle.cpp")
    log(f"{name}: done in {time.time() - t_start:.0f} s, stop rule at step {stop_at}")
    return dict(
        meta=np.array(json.dumps(meta)),
        ic_hash=FC.ic_hash(ic),
        eF=np.array([eF]),
        samples=samples,
        ck_hash=np.array(ck_hash, dtype=np.uint8),
        ck_samp=np.array(ck_samp),
        ck_max=np.array(ck_max),
        fo_steps=np.array(fo_steps, dtype=np.int64),
        fo_count=np.array(fo_count, dtype=np.int64),
        fo_hash=np.array(fo_hash, dtype=np.uint8),
        above=np.array(above, dtype=np.int64),
    )


def main(name):
    out = oracle_run(name, log=lambda m: print(m, flush=True))
    np.savez_compressed(FC.fixture_path(name), **out)


if __name__ == "__main__":
    main(sys.argv[1])
