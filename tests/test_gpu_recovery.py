"""Each branch of the primitive recovery through the device code
(vh_inferred_variables, the recover() of the stage kernels) against the
oracle on the crafted states of vhtest.recovery_states: the closed form below
ePrev = 0.1, Newton convergence, Newton failing into a bisection without a
sign change (the device returns ePrev at once where the reference iterates
10^6 times to the same value), the e = p = 1e-7 floor, the u^tau > 1e3
velocity reset, a NaN e and a NaN Newton step (reported as site 1 and 2,
VH_ERR_NAN).  Bit for bit, NaN outputs as NaN.  tests/test_recovery_edges.py
shows on the CPU that the oracle takes each branch and equals the reference's
own getInferredVariables there."""
import numpy as np
import pytest

import vhtest as V
from cpuvh_amd import Hydro
from test_recovery_edges import T, _oracle, oracle_inferred

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("table", [True, False])
@pytest.mark.parametrize("bulk", [False, True])
def test_recovery_branches_match_oracle(table, bulk):
    print(f"case recovery branches table={table} bulk={bulk}", flush=True)
    nc = 15 if bulk else 14
    q, ep, labels = V.recovery_states(nc, T, table=table)
    want, err, _, _ = oracle_inferred(_oracle(table, bulk), q, ep)
    g = Hydro(3, 3, 1, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2, bulk=bulk, eos_table=V.eos_table() if table else None)
    got, nan = g.inferred(T, q, ep)
    g.close()
    for i, lab in enumerate(labels):
        assert nan[i] == err[i], (i, lab, nan[i], err[i])
        assert V.same_or_both_nan(got[:, i], want[:, i]).all(), (i, lab, got[:, i], want[:, i])
    assert {"closed", "floor", "ut_reset", "nan_e"} <= set(labels)
    if table:
        assert {"newton", "stuck", "nan_root"} <= set(labels)
