"""Whole runs of the plain FP64 path against the oracle (tests/fullrun_cases.py):
the 400-step Gubser configuration (201 x 201) and the 600-step 3+1D drop-in
configuration (131 x 131 x 61, table EoS, bulk, eta/s(T), regulation below
100 MeV).  At every checkpoint each field must equal the oracle's bit for bit
(sha256 of the whole array), and at every coarse step the flagged freezeout
cubes must be the oracle's set (count and sha256 of the ordered ids) and the
count of cells above e_F the oracle's count.

The oracle's results come from tests/golden/fullrun_<case>.npz, written in
the build container by tests/golden/make_fullrun.py.  The fixture records a
hash of the initial state it was written from; the Gubser state is computed
here with numpy/scipy, so where this host's libraries round it differently the
oracle is run here instead (about 20 s), on the state this host computes.
"""
import json
import os
import sys

import numpy as np
import pytest

import fullrun_cases as FC
import vhtest as V
from cpuvh_amd import Hydro

pytestmark = pytest.mark.gpu


def _sha(a):
    import hashlib
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), dtype=np.uint8)


@pytest.mark.parametrize("name", ["gubser", "glauber3d"])
def test_full_run_bitwise(name):
    case = FC.CASES[name]
    made = FC.make_case(name)
    ic, kw, eos = made
    fx = dict(np.load(FC.fixture_path(name), allow_pickle=False))
    if not np.array_equal(FC.ic_hash(ic), fx["ic_hash"]):
        assert name == "gubser", "the seeded MC-Glauber state differs from the one the fixture was written from"
        sys.path.insert(0, os.path.join(V.REPO, "tests", "golden"))
        import make_fullrun as MF
        print("gubser state rounds differently on this host: running the oracle here", flush=True)
        fx = MF.oracle_run(name, made=made, log=lambda m: print(m, flush=True))
    meta = json.loads(str(fx["meta"]))
    nx, ny, nz = case["nx"], case["ny"], case["nz"]
    dx, dt, t0, steps, coarse = case["dx"], case["dt"], case["t0"], case["steps"], case["coarse"]
    print(f"case full run {name} {nx}x{ny}x{nz} steps={steps}", flush=True)
    table = V.eos_table() if eos == "table" else None
    g = Hydro(nx, ny, nz, dx, dx, dx, dt, t0, case["etabar"], eos_table=table, **kw)
    eF = g.eos_energy_density(case["T_F_GeV"] / V.HBARC)
    assert eF == fx["eF"][0]
    g.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic.get("Pi"))
    g.init_conserved(t0)
    fields = meta["fields"]
    for f in ("regfac", "regfac_bulk"):
        if f in fields:
            g.get(f)  # the first request enables the regulation-factor output
    samples = fx["samples"]
    fo_i = 0
    ck_i = 0
    report = []
    for n in range(1, steps + 1):
        if n % coarse == 0:
            ids = g.fo_flag_ids(eF)
            assert fx["fo_steps"][fo_i] == n
            assert ids.size == fx["fo_count"][fo_i], f"step {n}: {ids.size} cubes, oracle {fx['fo_count'][fo_i]}"
            assert np.array_equal(_sha(ids.astype("<i8")), fx["fo_hash"][fo_i]), f"step {n}: cube ids differ"
            assert g.count_above(eF) == fx["above"][fo_i], f"step {n}: count above e_F"
            fo_i += 1
        g.step(t0 + (n - 1) * dt)
        if n in meta["checkpoints"]:
            g.sync()
            bad = []
            for k, f in enumerate(fields):
                a = g.get(f)
                if not np.array_equal(_sha(a.view("<u8")), fx["ck_hash"][ck_i][k]):
                    ref = fx["ck_samp"][ck_i][k]
                    err = float(np.max(np.abs(a[samples] - ref)) / max(fx["ck_max"][ck_i][k], 1e-300))
                    bad.append(f"{f} (sampled max |d|/max {err:.2e})")
            report.append((n, bad))
            print(f"step {n}: {len(fields) - len(bad)} of {len(fields)} fields bit-identical", flush=True)
            ck_i += 1
    g.close()
    assert fo_i == len(fx["fo_steps"]) and ck_i == len(meta["checkpoints"])
    assert all(not bad for _, bad in report), report
