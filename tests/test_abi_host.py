"""C-ABI surface and host-side functions (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import vhtest as V
import cpuvh_amd as C

HEADER = os.path.join(V.REPO, "include", "cpuvh_amd.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return set(re.findall(r"\b(vh_[a-z_0-9]+)\s*\(", src))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(C.LIB_PATH)
    names = _declared()
    assert names, "no declarations parsed"
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared in include/cpuvh_amd.h but not exported"
    assert names == set(C.EXPORTS)


def test_status_strings():
    L = C._lib()
    for st in (C.VH_OK, C.VH_ERR_ARG, C.VH_ERR_NAN, C.VH_ERR_STATE, C.VH_ERR_CAPACITY, C.VH_ERR_NOMEM):
        assert L.vh_status_string(st)
    assert L.vh_build_info()


def test_invalid_configuration_is_rejected_before_device_use():
    cfg = C.VhConfig(0, 5, 5, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2, 0.0, 0, 0, None, 0.154, 0.0, 0.0, 0)
    h = ctypes.c_void_p()
    assert C._lib().vh_create(ctypes.byref(cfg), ctypes.byref(h)) == C.VH_ERR_ARG
    cfg = C.VhConfig(5, 5, 5, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2, 0.0, 0, C.EOS_TABLE, None, 0.154, 0.0, 0.0, 0)
    assert C._lib().vh_create(ctypes.byref(cfg), ctypes.byref(h)) == C.VH_ERR_ARG


def test_null_handle_calls_fail_cleanly():
    L = C._lib()
    assert L.vh_step(None, 0.6) == C.VH_ERR_ARG
    assert L.vh_sync(None) == C.VH_ERR_ARG
    assert L.vh_destroy(None) == C.VH_ERR_ARG


def test_mc_glauber_seeded_and_shaped():
    nx, ny, nz, dx = 41, 41, 15, 0.3
    e1, i1 = C.mc_glauber(nx, ny, nz, dx, dx, 0.2, seed=7)
    e2, i2 = C.mc_glauber(nx, ny, nz, dx, dx, 0.2, seed=7)
    e3, i3 = C.mc_glauber(nx, ny, nz, dx, dx, 0.2, seed=8)
    assert np.array_equal(e1, e2) and i1 == i2
    assert not np.array_equal(e1, e3)
    assert 0 <= i1["b"] <= 10 and 2 <= i1["n_part"] <= 2 * 197
    assert np.all(e1 >= 1e-3)
    e = e1.reshape(nz, ny, nx)
    # longitudinal profile of InitialConditions.cpp:1046-1064 with rapidityMean 4, variance 1.96
    eta = (np.arange(nz) - (nz - 1) // 2) * 0.2
    es = np.abs(eta) - 2.0
    prof = np.exp(-es * es / 1.96 / 2 * (es >= 0))
    ratio = (e - 1e-3) / (e[(nz - 1) // 2] - 1e-3)[None]
    mask = e[(nz - 1) // 2] > 1.0
    for k in range(nz):
        assert np.allclose(ratio[k][mask], prof[k], rtol=1e-12)


def test_tmunu_round_trip_through_float(tmp_path):
    nx, ny, nz = 5, 4, 3
    ic = V.random_ic(nx, ny, nz, 0.1, 0.6, seed=3, bulk=True)
    path = str(tmp_path / "Tmunu.dat")
    C.write_tmunu(path, nx, ny, nz, 0.1, 0.1, 0.1, ic)
    back = C.read_tmunu(path, nx, ny, nz)
    for k in ("e", "p", "ut", "ux", "uy", "un", "pi", "Pi"):
        want = np.asarray(ic[k], dtype=np.float32).astype(np.float64)
        assert np.array_equal(back[k], want), k
    # 20 columns per line, i outer and k inner (InitialConditions.cpp:68-71)
    rows = np.loadtxt(path)
    assert rows.shape == (nx * ny * nz, 20)
    assert rows[1, 2] > rows[0, 2] and rows[nz, 1] > rows[0, 1]
