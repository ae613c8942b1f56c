"""The branches of the primitive recovery (getInferredVariables,
rhic/src/EnergyMomentumTensor.cpp:22-233) on crafted states
(vhtest.recovery_states): the oracle takes the intended branch for every
state (checked from its iteration counts and outputs), and agrees bit for bit
with the reference's own getInferredVariables compiled here (oracle/_ref),
NaN states aside (the reference exits on them; the oracle records the site).
The GPU side of the same states is tests/test_gpu_recovery.py."""
import ctypes

import numpy as np
import pytest

import vhtest as V

L = V.oracle_lib()
L.orc_last_root.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
L.orc_clear_error.argtypes = [ctypes.c_void_p]
T = 1.0


def oracle_inferred(o, q, eprev):
    n = eprev.size
    out = np.empty((6, n))
    err = np.zeros(n, dtype=np.int32)
    nr = np.zeros(n, dtype=np.int64)
    bi = np.zeros(n, dtype=np.int64)
    for i in range(n):
        qc = np.zeros(15)
        qc[:q.shape[0]] = q[:, i]
        r = np.empty(6)
        L.orc_clear_error(o.h)
        err[i] = L.orc_inferred(o.h, T, qc.ctypes.data, float(eprev[i]), r.ctypes.data)
        a, b = ctypes.c_int(), ctypes.c_int()
        L.orc_last_root(o.h, ctypes.byref(a), ctypes.byref(b))
        out[:, i], nr[i], bi[i] = r, a.value, b.value
    return out, err, nr, bi


def _oracle(table, bulk):
    return V.Oracle(3, 3, 1, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2, bulk=bulk, table=V.eos_table() if table else None)


@pytest.mark.parametrize("table", [True, False])
@pytest.mark.parametrize("bulk", [False, True])
def test_oracle_takes_each_branch(table, bulk):
    nc = 15 if bulk else 14
    q, ep, labels = V.recovery_states(nc, T, table=table)
    out, err, nr, bi = oracle_inferred(_oracle(table, bulk), q, ep)
    seen = set()
    for i, lab in enumerate(labels):
        if lab == "closed":
            assert nr[i] == -1 and ep[i] <= 0.1
        elif lab == "newton":
            assert 0 < nr[i] < 10000 and bi[i] == 0, (i, nr[i], bi[i])
        elif lab == "stuck":
            assert nr[i] == 10000 and bi[i] == 1000000 and out[0, i] == ep[i], (i, nr[i], bi[i])
        elif lab == "floor":
            assert out[0, i] == 1e-7 and out[1, i] == 1e-7
        elif lab == "ut_reset":
            assert q[1, i] != 0 or q[2, i] != 0
            assert tuple(out[2:, i]) == (1.0, 0.0, 0.0, 0.0)
        elif lab == "nan_e":
            assert err[i] == 1
        elif lab == "nan_root":
            assert err[i] == 2 and out[0, i] == ep[i]
        if not lab.startswith("nan"):
            assert err[i] == 0, (i, lab)
        seen.add(lab)
    want = {"closed", "floor", "ut_reset", "nan_e"} | ({"newton", "stuck", "nan_root"} if table else set())
    assert want <= seen


@pytest.mark.parametrize("table", [True, False])
@pytest.mark.parametrize("bulk", [False, True])
def test_oracle_equals_reference_on_each_branch(table, bulk):
    variant = ("lat" if table else "conf") + ("_bulk" if bulk else "")
    if not V.ref_available(variant):
        pytest.skip("reference build not present")
    nc = 15 if bulk else 14
    q, ep, labels = V.recovery_states(nc, T, table=table)
    out, err, _, _ = oracle_inferred(_oracle(table, bulk), q, ep)
    ref = V.Reference(variant, 3, 3, 1, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2)
    ref.L.ref_get_inferred.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    checked = 0
    for i, lab in enumerate(labels):
        if lab.startswith("nan"):
            continue  # the reference prints and exits
        qc = np.zeros(15)
        qc[:nc] = q[:, i]
        r = np.empty(6)
        ref.L.ref_get_inferred(T, qc.ctypes.data, float(ep[i]), r.ctypes.data)
        assert V.bits_equal(r, out[:, i]), (i, lab, r, out[:, i])
        checked += 1
    assert checked == sum(not lab.startswith("nan") for lab in labels)
