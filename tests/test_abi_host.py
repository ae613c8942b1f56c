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


def test_classic_glauber_sum_rule_and_participants():
    """The reference's type-4 profile: a fixed-seed event (the same call gives
    the same bits), a central Au+Au participant count in the physical range,
    and the transverse integral of e - 1e-3 equal to e0 N_w 2 pi sigma^2 with
    the reference's sigma = 0.96 fm (a Gaussian sum sampled at dx << sigma
    integrates exactly); the longitudinal plateau leaves mid-rapidity at 1 and
    the tails follow exp(-(|eta| - eta_flat/2)^2 / (2 var))."""
    n, dx = 161, 0.2
    e, nw = C.mc_glauber_classic(n, n, 1, dx, dx, 0.1, e0=2.0)
    e2, nw2 = C.mc_glauber_classic(n, n, 1, dx, dx, 0.1, e0=2.0)
    assert nw == nw2 and V.bits_equal(e, e2)
    assert 330 <= nw <= 394, nw
    integral = float(np.sum(e - 1e-3)) * dx * dx
    assert abs(integral / (2.0 * nw * 2 * np.pi * 0.96 ** 2) - 1) < 1e-6
    # no overlap, no participants: the floor only
    e0, nw0 = C.mc_glauber_classic(21, 21, 1, dx, dx, 0.1, b=40.0)
    assert nw0 == 0 and np.all(e0 == 1e-3)
    # longitudinal profile on an odd grid: eta = (k - (nz-1)/2) deta
    nz, deta = 41, 0.25
    e3, _ = C.mc_glauber_classic(31, 31, nz, dx, dx, deta, e0=1.0, eta_flat=4.0, eta_variance=1.96)
    col = e3.reshape(nz, 31, 31)[:, 15, 15] - 1e-3
    eta = (np.arange(nz) - (nz - 1) // 2) * deta
    ratio = col / col[(nz - 1) // 2]
    tail = np.abs(eta) > 2.0
    assert np.allclose(ratio[~tail], 1.0, rtol=0, atol=1e-12)
    assert np.allclose(ratio[tail], np.exp(-(np.abs(eta[tail]) - 2.0) ** 2 / (2 * 1.96)), rtol=1e-12)


def test_driver_binary_parses_and_exits_without_hydro(tmp_path):
    import subprocess
    exe = os.path.join(os.path.dirname(C.LIB_PATH), "cpuvh_run")
    (tmp_path / "lattice.properties").write_text("numLatticePointsX=5\n")
    r = subprocess.run([exe, "--config", str(tmp_path), "-o", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert "runHydro = False" in r.stdout and "Hydro took" in r.stdout
    assert "No configuration file found for initial condition parameters" in r.stderr


def _print_config(conf_dir):
    import subprocess
    exe = os.path.join(os.path.dirname(C.LIB_PATH), "cpuvh_run")
    r = subprocess.run([exe, "--config", str(conf_dir), "--print-config"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return {ln.split()[0]: ln.split()[1:] for ln in r.stdout.splitlines() if ln.split()[:1] in (["lattice"], ["ic"], ["hydro"], ["gpu"])}


def test_driver_reads_reference_rhic_conf():
    """The reference's own rhic-conf files (tests/golden/rhic-conf, copied as
    data): values by position, keys ignored -- the misspelt
    coarseProperTimeFOFactorFO key still gives the FO factor."""
    cfg = _print_config(os.path.join(V.REPO, "tests", "golden", "rhic-conf"))
    lat = cfg["lattice"]
    assert [int(x) for x in lat[:4]] == [101, 101, 1, 10000]
    assert [float(x) for x in lat[4:8]] == [0.2, 0.2, 0.1, 0.05] and int(lat[8]) == 2
    ic = cfg["ic"]
    assert [int(ic[0]), int(ic[1]), int(ic[2])] == [2, 0, 208]
    assert [float(x) for x in ic[3:9]] == [70.0, 62.0, 0.0, 1.0, 0.16, 5.9] and int(ic[9]) == 0
    hy = cfg["hydro"]
    assert [float(hy[0]), float(hy[1]), float(hy[2]), int(hy[3]), float(hy[4]), int(hy[5])] == [0.5, 0.2, 0.155, 0,
                                                                                                  0.1, 1]
    assert cfg["gpu"] == ["0", "0", "EoS/hrg_hotqcd_eos_binary.dat", "0", "100"]


def test_driver_reader_short_file_and_spacing(tmp_path):
    """Whitespace anywhere is dropped; once a file runs out, the last line's
    value repeats (std::getline leaves the line buffer as it was); a missing
    file leaves zeros; gpu.properties is keyed."""
    (tmp_path / "lattice.properties").write_text("numLatticePointsX=7\nnumLatticePointsY = 9\n numLatticePointsRapidity=3")
    (tmp_path / "gpu.properties").write_text("# compile-time options\nbulk = 1\neos=table\nsnapshotEvery=25\n")
    cfg = _print_config(tmp_path)
    assert cfg["lattice"] == ["7", "9", "3", "3", "3", "3", "3", "3", "3"]
    assert [float(x) for x in cfg["hydro"]] == [0.0] * 6
    assert cfg["gpu"][:2] == ["1", "1"] and cfg["gpu"][4] == "25"


def test_optical_glauber_closed_forms():
    """The reference's IC type 2: normalised to e0 (+ the 1e-3 floor) at the
    centre of a central collision for any binary fraction; for wounded
    nucleons only (alpha = 0) the profile is T(x, y)/T(0, 0), whose integral
    is the Fermi integral of the Woods-Saxon density, (4 pi/3) n0 R^3
    (1 + (pi d/R)^2) - 8 pi n0 d^3 Li3(-e^(-R/d)), over T(0, 0) =
    2 n0 d ln(1 + e^(R/d))."""
    for alpha in (0.0, 0.3, 1.0):
        e = C.optical_glauber(41, 41, 1, 0.4, 0.4, 0.1, alpha=alpha, e0=5.0)
        assert abs(e.reshape(41, 41)[20, 20] - 5.001) < 1e-12
    n, dx = 221, 0.2  # +-22 fm: the density is down by e^-29 at the edge
    e = C.optical_glauber(n, n, 1, dx, dx, 0.1, alpha=0.0, e0=1.0)
    A, n0, d = 197.0, 0.17, 0.54
    R = 1.12 * A ** (1 / 3) - 0.86 * A ** (-1 / 3)
    li3 = sum((-np.exp(-R / d)) ** k / k ** 3 for k in range(1, 8))
    volume = (4 * np.pi / 3) * n0 * R ** 3 * (1 + (np.pi * d / R) ** 2) - 8 * np.pi * n0 * d ** 3 * li3
    want = volume / (2 * n0 * d * np.log1p(np.exp(R / d)))
    got = float(np.sum(e - 1e-3)) * dx * dx
    assert abs(got / want - 1) < 1e-8, (got, want)
    # an off-centre collision is symmetric under x -> -x and y -> -y
    e = C.optical_glauber(31, 29, 1, 0.5, 0.5, 0.1, b=6.0, alpha=0.5, e0=1.0).reshape(29, 31)
    assert np.allclose(e, e[::-1, :], rtol=0, atol=1e-14) and np.allclose(e, e[:, ::-1], rtol=0, atol=1e-14)
