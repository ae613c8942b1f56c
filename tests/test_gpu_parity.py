"""HIP path vs the oracle on identical seeded inputs: bit-exact, regulated
runs included (the regulator's tanh is the host C library's algorithm,
vh::glibc_tanh, checked against it on the host in tests/test_libm_host.py)."""
import numpy as np
import pytest

import vhtest as V
from cpuvh_amd import Hydro, FIELD

pytestmark = pytest.mark.gpu

ALL_FIELDS = ["ttt", "ttx", "tty", "ttn", "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy",
              "piyn", "pinn", "e", "p", "ut", "ux", "uy", "un", "up_t", "up_x", "up_y", "up_n"]


def _pair(nx, ny, nz, lattice, bulk, T_reg, dx=0.1, dt=0.02, t0=0.6, etabar=0.2, seed=3, **ext):
    table = V.eos_table() if lattice else None
    ic = V.random_ic(nx, ny, nz, dx, t0, seed=seed, lattice=lattice, bulk=bulk)
    g = Hydro(nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=T_reg, bulk=bulk, eos_table=table, **ext)
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=T_reg, bulk=bulk, table=table, **ext)
    g.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic.get("Pi"))
    o.set_ic(ic)
    g.init_conserved(t0)
    o.init_conserved(t0)
    return g, o


def _run(g, o, nsteps, t0, dt):
    for n in range(1, nsteps + 1):
        t = t0 + (n - 1) * dt
        g.step(t)
        assert o.step(t) == 0
    g.sync()


def _fields(bulk):
    return ALL_FIELDS + (["Pi"] if bulk else [])


@pytest.mark.parametrize("nx,ny,nz", [(21, 17, 1), (13, 11, 9)])
@pytest.mark.parametrize("lattice", [False, True])
@pytest.mark.parametrize("bulk", [False, True])
def test_bitwise_unregulated(nx, ny, nz, lattice, bulk):
    print(f"case nx={nx} ny={ny} nz={nz} lattice={lattice} bulk={bulk}", flush=True)
    g, o = _pair(nx, ny, nz, lattice, bulk, T_reg=0.0)
    _run(g, o, 6, 0.6, 0.02)
    for f in _fields(bulk):
        a, b = o.get(f), g.get(f)
        assert V.bits_equal(a, b), f"{f}: max rel {V.max_rel(a, b):.3e}"


@pytest.mark.parametrize("lattice", [False, True])
@pytest.mark.parametrize("bulk", [False, True])
def test_regulated_bitwise(lattice, bulk):
    print(f"case regulated lattice={lattice} bulk={bulk}", flush=True)
    nx, ny, nz = 15, 13, 7
    g, o = _pair(nx, ny, nz, lattice, bulk, T_reg=0.25)
    assert o.e_reg() > 1.0  # most cells regulated
    g.get("regfac")  # the first request enables the regulation-factor output
    if bulk:
        g.get("regfac_bulk")
    _run(g, o, 6, 0.6, 0.02)
    reg = g.get("regfac")
    assert np.count_nonzero(reg != 1.0) > reg.size // 2  # the regulator acted on most cells
    for f in _fields(bulk) + ["regfac"] + (["regfac_bulk"] if bulk else []):
        a, b = o.get(f), g.get(f)
        assert V.bits_equal(a, b), f"{f}: max rel {V.max_rel(a, b):.3e}"


def test_graph_steps_equal_single_steps():
    nx, ny, nz = 19, 15, 5
    t0, dt = 0.6, 0.02
    g1, o = _pair(nx, ny, nz, False, False, T_reg=0.0)
    g2, _ = _pair(nx, ny, nz, False, False, T_reg=0.0)
    for n in range(1, 8):
        g1.step(t0 + (n - 1) * dt)
    g2.steps(1, 3)
    g2.steps(4, 4)
    for f in ALL_FIELDS:
        assert V.bits_equal(g1.get(f), g2.get(f)), f


def test_eta_over_s_extension_reduces_and_differs():
    nx, ny, nz = 15, 11, 3
    g, o = _pair(nx, ny, nz, False, False, T_reg=0.0, etas_Tk_GeV=0.154, etas_slope_lo=0.0, etas_slope_hi=0.0)
    _run(g, o, 3, 0.6, 0.02)
    assert V.bits_equal(o.get("pixx"), g.get("pixx"))
    g2, o2 = _pair(nx, ny, nz, False, False, T_reg=0.0, etas_Tk_GeV=0.154, etas_slope_lo=0.5, etas_slope_hi=1.5)
    _run(g2, o2, 3, 0.6, 0.02)
    assert V.bits_equal(o2.get("pixx"), g2.get("pixx"))
    assert not V.bits_equal(o.get("pixx"), o2.get("pixx"))


def test_count_and_freezeout_flags_match_oracle():
    nx, ny, nz = 23, 19, 6
    t0, dt = 0.6, 0.02
    g, o = _pair(nx, ny, nz, True, False, T_reg=0.1)
    eF = g.eos_energy_density(0.15 / V.HBARC)
    prev = np.zeros(g.N)
    for n in range(1, 5):
        ids = g.fo_flag_ids(eF)
        cur = o.get("e")
        want = V.oracle_fo_flag(prev, cur, nx, ny, nz, eF)
        assert np.array_equal(ids, want)
        assert g.count_above(eF) == int(np.count_nonzero(cur > eF))
        prev = cur
        t = t0 + (n - 1) * dt
        g.step(t)
        o.step(t)


@pytest.mark.parametrize("nx,ny,nz", [(71, 71, 61), (601, 601, 1)])
def test_freezeout_flags_many_blocks(nx, ny, nz):
    """Over 1024 blocks of 256 cubes (294,000 and 360,000 cubes), so the
    offsets scan of the ordered flag compaction walks several chunks: the ids
    equal the oracle's crossing test on the same two slices."""
    g, _ = _pair(nx, ny, nz, False, False, T_reg=0.1)
    eF = float(np.median(g.get("e")))
    prev = np.zeros(g.N)
    for n in range(3):
        ids = g.fo_flag_ids(eF)
        cur = g.get("e")
        want = V.oracle_fo_flag(prev, cur, nx, ny, nz, eF)
        assert len(want) > 0
        assert np.array_equal(ids, want), n
        prev = cur
        g.step(0.6 + n * 0.02)


FO_VARS = ["ux", "uy", "un", "e", "pixx", "pixy", "pixn", "piyy", "piyn", "Pi"]


def _lin(f, a, dim):
    """interpolateVariable3D/4D (freezeout/FreezeOut.cpp:13-56) with numpy
    float64 operations in the reference's term order; a[t][x][y][z] (n, ...)."""
    x = [f[:, k] for k in range(dim)]
    w = [(1 - xk, xk) for xk in x]
    if dim == 3:
        order = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
    else:
        order = [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1), (1, 1, 0, 0), (1, 0, 1, 0),
                 (1, 0, 0, 1), (0, 1, 1, 0), (0, 1, 0, 1), (0, 0, 1, 1), (1, 1, 1, 0), (1, 1, 0, 1), (1, 0, 1, 1),
                 (0, 1, 1, 1), (1, 1, 1, 1)]
    r = None
    for idx in order:
        term = w[0][idx[0]]
        for k in range(1, dim):
            term = term * w[k][idx[k]]
        term = term * a[(slice(None),) + idx]
        r = term if r is None else r + term
    return r


@pytest.mark.parametrize("nx,ny,nz", [(17, 13, 1), (13, 11, 7)])
def test_freezeout_interpolation(nx, ny, nz):
    """vh_fo_interp on the device's previous/current slices: bit-identical to
    the reference's interpolation of the oracle's fields at random points, and
    exactly the corner value at every corner of the cube (which pins the
    corner indexing and the slice order)."""
    dim = 4 if nz > 1 else 3
    g, o = _pair(nx, ny, nz, True, True, T_reg=0.1)
    eF = float(np.median(o.get("e")))  # a level that cuts cubes at every step
    rng = np.random.default_rng(11)
    prev = {v: np.zeros(g.N) for v in FO_VARS}
    t0, dt = 0.6, 0.02
    for n in range(1, 4):
        ids = g.fo_flag_ids(eF)
        cur = {v: o.get(v) for v in FO_VARS}
        assert len(ids) > 0
        cubes = np.concatenate([ids, ids[:1].repeat(2 ** dim)])
        frac = np.zeros((len(cubes), 4))
        frac[:len(ids), :dim] = rng.random((len(ids), dim))
        corners = np.array(list(np.ndindex(*([2] * dim))), dtype=np.float64)
        frac[len(ids):, :dim] = corners
        got = g.fo_interp(cubes, frac)
        mx, my = nx - 1, ny - 1
        ix, iy, iz = cubes % mx, (cubes // mx) % my, cubes // (mx * my)
        for v, name in enumerate(FO_VARS):
            sl = (prev[name], cur[name])
            shape = (len(cubes),) + (2,) * dim
            a = np.empty(shape)
            for c in np.ndindex(*([2] * dim)):
                s = (ix + c[1]) + nx * ((iy + c[2]) + ny * ((iz + c[3]) if dim == 4 else iz))
                a[(slice(None),) + c] = sl[c[0]][s]
            want = _lin(frac, a, dim)
            assert V.bits_equal(got[:, v], want), (n, name)
            for q, c in enumerate(np.ndindex(*([2] * dim))):
                assert got[len(ids) + q, v] == a[len(ids) + q][c], (n, name, c)
        prev = cur
        t = t0 + (n - 1) * dt
        g.step(t)
        o.step(t)
    with pytest.raises(RuntimeError):
        g.fo_interp(np.array([(nx - 1) * (ny - 1) * max(nz - 1, 1)]), np.zeros((1, 4)))


def test_gubser_configuration():
    """The Gubser configuration (201 x 201, dx = 0.05 fm, dtau = 0.0125 fm/c,
    conformal EoS, eta/s = 0.2, regulation below 100 MeV) on the GPU for 40
    steps against the semi-analytic solution with the oracle's tolerances
    (tests/test_gubser.py); the whole 400-step run against the oracle is
    tests/test_gpu_fullrun.py."""
    import gubser as Gb
    n, dx, dt, nsteps, etas = 201, 0.05, 0.0125, 40, 0.2
    print(f"case gubser {n}x{n} steps={nsteps}", flush=True)
    x = (np.arange(n) - (n - 1) / 2) * dx
    Y, X = np.meshgrid(x, x, indexing="ij")
    sol = Gb.solve(etas)
    e, u, pi = Gb.fields(sol, 1.0, X, Y)
    ic = {"e": e.ravel(), "p": (e / 3).ravel(), "ut": u[0].ravel(), "ux": u[1].ravel(), "uy": u[2].ravel(),
          "un": u[3].ravel(), "pi": np.concatenate([p.ravel() for p in pi])}
    g = Hydro(n, n, 1, dx, dx, dx, dt, 1.0, etas, T_reg_GeV=0.1)
    o = V.Oracle(n, n, 1, dx, dx, dx, dt, 1.0, etas, T_reg_GeV=0.1)
    g.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"])
    o.set_ic(ic)
    g.init_conserved(1.0)
    o.init_conserved(1.0)
    g.steps(1, nsteps)
    for s in range(1, nsteps + 1):
        assert o.step(1.0 + (s - 1) * dt) == 0
    g.sync()
    tau = 1.0 + nsteps * dt
    e2, u2, pi2 = Gb.fields(sol, tau, X, Y)
    core = (X * X + Y * Y < 1.0).ravel()
    ge, gux = g.get("e"), g.get("ux")
    assert np.max(np.abs(ge - e2.ravel())[core] / e2.ravel()[core]) < 2.5e-2
    assert np.max(np.abs(gux - u2[1].ravel())[core]) < 1e-2
    c = (n // 2) * n + n // 2
    assert abs(g.get("pinn")[c] / pi2[9].ravel()[c] - 1) < 4e-2
    for f in ("e", "ux", "uy", "pitt", "pixx", "pinn"):
        assert V.bits_equal(o.get(f), g.get(f)), f
