"""The oracle against the reference's own rungeKutta2, built from the cpu-vh
sources (oracle/ref.mk), on identical seeded states: every field bit for bit
after several steps, for conformal / table EoS x shear / shear+bulk x 2+1D /
3+1D, with and without active regulation.  This is what pins the oracle."""
import numpy as np
import pytest

import vhtest as V

pytestmark = pytest.mark.skipif(not V.ref_available("conf"), reason="reference build (oracle/_ref) not present")

FIELDS = ["ttt", "ttx", "tty", "ttn", "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn",
          "pinn", "e", "p", "ut", "ux", "uy", "un", "up_t", "up_x", "up_y", "up_n"]


def _run_both(variant, nx, ny, nz, T_reg, steps, dx=0.1, dt=0.02, t0=0.6, etabar=0.2, seed=5):
    lattice = variant.startswith("lat")
    bulk = variant.endswith("bulk")
    table = V.eos_table() if lattice else None
    ic = V.random_ic(nx, ny, nz, dx, t0, seed=seed, lattice=lattice, bulk=bulk)
    ref = V.Reference(variant, nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=T_reg)
    orc = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, etabar, T_reg_GeV=T_reg, bulk=bulk, table=table)
    ref.set_ic(ic)
    orc.set_ic(ic)
    ref.init_conserved(t0)
    orc.init_conserved(t0)
    for n in range(1, steps + 1):
        t = t0 + (n - 1) * dt
        ref.step(t)
        assert orc.step(t) == 0
    return ref, orc


@pytest.mark.parametrize("variant", ["conf", "conf_bulk", "lat", "lat_bulk"])
@pytest.mark.parametrize("dims", [(19, 15, 1), (11, 9, 7)])
@pytest.mark.parametrize("T_reg", [0.0, 0.25])
def test_oracle_bitwise_equals_reference(variant, dims, T_reg):
    ref, orc = _run_both(variant, *dims, T_reg=T_reg, steps=4)
    fields = FIELDS + (["Pi"] if variant.endswith("bulk") else [])
    for f in fields:
        a, b = ref.get(f), orc.get(f)
        assert V.bits_equal(a, b), f"{variant} {dims} T_reg={T_reg} {f}: max rel {V.max_rel(a, b):.3e}"


def test_regulation_is_exercised():
    """With T_reg = 0.25 GeV most cells sit below e_reg and are regulated."""
    nx, ny, nz = 11, 9, 5
    table = None
    orc = V.Oracle(nx, ny, nz, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2, T_reg_GeV=0.25, table=table)
    ic = V.random_ic(nx, ny, nz, 0.1, 0.6, seed=2)
    orc.set_ic(ic)
    orc.init_conserved(0.6)
    orc.step(0.6)
    fac = orc.get("regfac")
    assert np.mean(fac < 1.0) > 0.5


def test_root_solve_takes_the_table_branch():
    """ePrev > 0.1 cells go through Newton-Raphson on the table EoS: the
    recovered e differs from the conformal closed form."""
    table = V.eos_table()
    L = V.oracle_lib()
    o = V.Oracle(3, 3, 1, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2, table=table)
    q = np.zeros(15)
    q[0], q[1] = 12.0, 1.5
    out = np.zeros(6)
    L.orc_inferred(o.h, 0.6, q.ctypes.data, 10.0, out.ctypes.data)
    e_table = out[0]
    oc = V.Oracle(3, 3, 1, 0.1, 0.1, 0.1, 0.02, 0.6, 0.2)
    L.orc_inferred(oc.h, 0.6, q.ctypes.data, 10.0, out.ctypes.data)
    assert abs(e_table - out[0]) > 1e-3 * out[0]
    # the root satisfies the table-EoS energy equation to the solver's tolerance:
    # M0 = (e + p) ut^2 - p with ut^2 = 1 + |M|^2/((e+p)^2 ut^2)  <=>  e = M0 - M/(M0 + p)
    M0, M = q[0], q[1] ** 2
    E = L.orc_eos_create(table.ctypes.data)
    p = L.orc_eos_pressure(E, e_table)
    L.orc_eos_destroy(E)
    assert abs(e_table - (M0 - M / (M0 + p))) < 2e-3 * e_table


def test_count_and_freezeout_cells_match_reference_cornelius(tmp_path):
    """Flagged cubes of the oracle == cubes in which the reference's Cornelius
    produced surface elements (read back from its surface.dat), over several
    coarse steps of a 2+1D table-EoS run, plus checkForCellsAboveTc."""
    nx, ny, nz, dx, dt, t0 = 23, 21, 1, 0.2, 0.02, 0.6
    table = V.eos_table()
    ic = V.random_ic(nx, ny, nz, dx, t0, seed=11, lattice=True, e_scale=4.0, e_floor=0.05)
    ref = V.Reference("lat", nx, ny, nz, dx, dx, dx, dt, t0, 0.2, T_reg_GeV=0.0, fo_factor=1)
    orc = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, 0.2, table=table)
    ref.set_ic(ic)
    orc.set_ic(ic)
    ref.init_conserved(t0)
    orc.init_conserved(t0)
    eF = ref.L.ref_eos_energy_density(0.15 / V.HBARC)
    prev = np.zeros(nx * ny * nz)
    flagged = set()
    surf = tmp_path / "surface.dat"
    assert ref.L.ref_fo_begin(str(surf).encode()) == 0
    t = t0
    for n in range(1, 7):
        cur = orc.get("e")
        ids = V.oracle_fo_flag(prev, cur, nx, ny, nz, eF)
        flagged |= {(n, int(c)) for c in ids}
        ref.L.ref_fo_call(n, t, eF)
        assert ref.L.ref_count_above(eF) == V.oracle_lib().orc_count_above(cur.ctypes.data, cur.size, eF)
        prev = cur
        ref.step(t)
        orc.step(t)
        t = t0 + n * dt
    ref.L.ref_fo_end()
    rows = np.loadtxt(surf, ndmin=2)
    assert rows.shape[0] > 0
    cubes = {c for (_, c) in flagged}
    matched = set()
    for r in rows:
        x, y = r[1], r[2]
        ix = int(np.floor((x + (nx - 1) / 2 * dx) / dx + 1e-12))
        iy = int(np.floor((y + (ny - 1) / 2 * dx) / dx + 1e-12))
        # a centroid on the cube's upper face belongs to the cube below it
        cands = [(ix - a) + (nx - 1) * (iy - b) for a in (0, 1) for b in (0, 1)]
        hit = [c for c in cands if c in cubes]
        assert hit, f"surface element at x={x}, y={y} outside every flagged cube"
        matched.add(hit[0])
    assert matched == cubes
