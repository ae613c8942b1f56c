"""Surface elements built on the device (vh_fo_surface) against the host
construction (vh_surface_cube / vh_surface_hypercube, pinned by the
reference's Cornelius fixtures in tests/test_surface.py) on the corner
energies vh_fo_flag returns, and the variables at the centroids against
vh_fo_interp: bit for bit, at several coarse steps including the first
(whose previous slice is zero, so every cube with a corner above e_F is cut)."""
import numpy as np
import pytest

import vhtest as V
import cpuvh_amd as C

pytestmark = pytest.mark.gpu


def _builder_order(ec, d4):
    half, nzc = (8, 2) if d4 else (4, 1)
    out = np.empty(2 * half)
    for it in range(2):
        for x in range(2):
            for y in range(2):
                for z in range(nzc):
                    out[((it * 2 + x) * 2 + y) * nzc + z] = ec[it * half + (z * 2 + y) * 2 + x]
    return out


@pytest.mark.parametrize("nx,ny,nz", [(25, 23, 1), (21, 19, 11)])
def test_device_elements_equal_host(nx, ny, nz):
    print(f"case fo_surface nx={nx} ny={ny} nz={nz}", flush=True)
    dx, dt, t0, coarse = 0.1, 0.02, 0.6, 2
    table = V.eos_table()
    ic = V.random_ic(nx, ny, nz, dx, t0, seed=9, lattice=True, bulk=True, e_scale=6.0, e_floor=0.3)
    h = C.Hydro(nx, ny, nz, dx, dx, dx, dt, t0, 0.1, T_reg_GeV=0.1, bulk=True, eos_table=table)
    h.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic["Pi"])
    h.init_conserved(t0)
    eF = h.eos_energy_density(0.15 / V.HBARC)
    d4 = nz > 1
    dim = 4 if d4 else 3
    widths = np.array([coarse * dt, dx, dx, dx])[:dim]
    n_step, total_el = 1, 0
    for k in range(4):
        if k:
            h.steps(n_step, coarse)
            n_step += coarse
        ids, corners = h.fo_flag_ids(eF, with_corners=True)
        counts, el, var = h.fo_surface(widths)
        assert counts.shape == ids.shape
        want_rows, want_ids = [], []
        for m in range(ids.size):
            rows = C.surface_elements(_builder_order(corners[m], d4), eF, widths)
            assert counts[m] == rows.shape[0], (k, m)
            want_rows.append(rows)
            want_ids += [ids[m]] * rows.shape[0]
        want = np.concatenate(want_rows) if want_rows else np.empty((0, 2 * dim))
        assert el.shape == want.shape
        assert V.bits_equal(el, want), f"coarse step {k}: elements differ"
        frac = np.zeros((el.shape[0], 4))
        frac[:, :dim] = el[:, :dim] / widths
        want_var = h.fo_interp(np.array(want_ids, dtype=np.int64), frac)
        assert V.bits_equal(var, want_var), f"coarse step {k}: variables differ"
        total_el += el.shape[0]
        print(f"coarse step {k}: {ids.size} cubes, {el.shape[0]} elements", flush=True)
    assert total_el > 100
    h.close()
