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


def _event_ic(nx, ny, nz, seed):
    return V.random_ic(nx, ny, nz, 0.1, 0.6, seed=seed, lattice=True, bulk=True, e_scale=6.0, e_floor=0.3)


BATCH_EVENTS = [(27, 25, 1), (19, 17, 9), (23, 21, 1)]


def test_batch_surface_equals_single_event():
    """An FP64 batch's events (neighbours in one pool) build the same surface,
    bit for bit, as the same events alone through the single-event path."""
    table = V.eos_table()
    ics = [_event_ic(nx, ny, nz, 30 + k) for k, (nx, ny, nz) in enumerate(BATCH_EVENTS)]
    b = C.Batch(sum(nx * ny * nz for nx, ny, nz in BATCH_EVENTS), len(BATCH_EVENTS), 64, bulk=True, eos_table=table,
                fo_vars=True)
    ids = [b.add(nx, ny, nz, 0.1, 0.1, 0.1, 0.02, 0.6, 0.1, ic, T_reg_GeV=0.1)
           for (nx, ny, nz), ic in zip(BATCH_EVENTS, ics)]
    hs = []
    for (nx, ny, nz), ic in zip(BATCH_EVENTS, ics):
        h = C.Hydro(nx, ny, nz, 0.1, 0.1, 0.1, 0.02, 0.6, 0.1, T_reg_GeV=0.1, bulk=True, eos_table=table)
        h.upload_ic(ic["e"], ic["p"], ic["ut"], ic["ux"], ic["uy"], ic["un"], ic["pi"], ic["Pi"])
        h.init_conserved(0.6)
        hs.append(h)
    eF = hs[0].eos_energy_density(0.15 / V.HBARC)
    n_step = 1
    for k in range(3):
        if k:
            b.steps(2)
            for h in hs:
                h.steps(n_step, 2)
            n_step += 2
        for (nx, ny, nz), eid, h in zip(BATCH_EVENTS, ids, hs):
            print(f"case batch surface {nx}x{ny}x{nz} check {k}", flush=True)
            widths = np.array([0.04, 0.1, 0.1, 0.1])[:4 if nz > 1 else 3]
            assert np.array_equal(b.fo_flag_ids(eid, eF), h.fo_flag_ids(eF))
            cb, eb, vb = b.fo_surface(eid, widths)
            ch, eh, vh = h.fo_surface(widths)
            assert np.array_equal(cb, ch) and V.bits_equal(eb, eh) and V.bits_equal(vb, vh)
    b.close()
    for h in hs:
        h.close()


def _same(a, b):
    return np.array_equal(a[0], b[0]) and V.bits_equal(a[1], b[1]) and V.bits_equal(a[2], b[2])


@pytest.mark.parametrize("prec", [C.PREC_FP64, C.PREC_MIXED])
def test_batch_flags_and_surfaces_in_one_pass_equal_per_event(prec):
    """vh_batch_fo_flags and vh_batch_fo_surfaces: the cut cubes and surfaces
    of several events (2+1D and 3+1D together) found in one pass each are,
    bit for bit, those of the same events in a twin batch flagged and built
    one event at a time — with an event below e_F everywhere (no cut cube)
    between them; an event given too little room for its elements reports
    its count and gets none while the others get theirs, and a call for it
    alone then builds it."""
    table = V.eos_table()
    events = BATCH_EVENTS[:2] + [(15, 13, 1)] + BATCH_EVENTS[2:]
    ics = [_event_ic(nx, ny, nz, 70 + k) for k, (nx, ny, nz) in enumerate(events)]
    ics[2] = V.random_ic(15, 13, 1, 0.1, 0.6, seed=77, lattice=True, bulk=True, e_scale=0.05, e_floor=0.01)
    cells = sum(nx * ny * nz for nx, ny, nz in events)
    twins = [C.Batch(cells, len(events), 64, bulk=True, eos_table=table, precision=prec, fo_vars=True)
             for _ in range(2)]
    ids = [[t.add(nx, ny, nz, 0.1, 0.1, 0.1, 0.02, 0.6, 0.1, ic, T_reg_GeV=0.1)
            for (nx, ny, nz), ic in zip(events, ics)] for t in twins]
    b, one = twins
    widths = [np.array([0.04, 0.1, 0.1, 0.1])[:4 if nz > 1 else 3] for nx, ny, nz in events]
    eF = b.eos_energy_density(0.15 / V.HBARC)
    checked = 0
    for k in range(3):
        if k:
            for t in twins:
                t.steps(2)
        print(f"case flags and surfaces in one pass, precision {prec}, check {k}", flush=True)
        flagged = b.fo_flags(ids[0], eF)
        for f, eid in zip(flagged, ids[1]):
            assert np.array_equal(f, one.fo_flag_ids(eid, eF))
        # the hot events lie above e_F on the whole grid: cut at the first check
        # only (the previous slice is zero there)
        assert flagged[2].size == 0 and (k > 0 or all(f.size > 0 for j, f in enumerate(flagged) if j != 2))
        st, tot, res = b.fo_surfaces(ids[0], widths)
        assert st == C.VH_OK and tot[2] == 0
        for eid, w, r in zip(ids[1], widths, res):
            assert _same(r, one.fo_surface(eid, w))
            checked += r[1].shape[0]
        if tot[1] > 0:
            caps = tot.copy()
            caps[1] -= 1
            st2, tot2, res2 = b.fo_surfaces(ids[0], widths, caps)
            assert st2 == C.VH_ERR_CAPACITY and res2[1] is None and np.array_equal(tot2, tot)
            assert all(_same(res2[j], res[j]) for j in (0, 2, 3))
            st3, _, res3 = b.fo_surfaces([ids[0][1]], [widths[1]])
            assert st3 == C.VH_OK and _same(res3[0], res[1])
    assert checked > 100
    for t in twins:
        t.close()


def test_mixed_batch_elements_from_its_corners():
    """Mixed-precision events: the device's elements are the host
    construction's on the corner energies the batch flags with (its FP32 e
    widened), and the interpolated e at each centroid lies between the
    cube's smallest and largest corner."""
    table = V.eos_table()
    ics = [_event_ic(nx, ny, nz, 50 + k) for k, (nx, ny, nz) in enumerate(BATCH_EVENTS)]
    b = C.Batch(sum(nx * ny * nz for nx, ny, nz in BATCH_EVENTS), len(BATCH_EVENTS), 64, bulk=True, eos_table=table,
                precision=C.PREC_MIXED, fo_vars=True)
    ids = [b.add(nx, ny, nz, 0.1, 0.1, 0.1, 0.02, 0.6, 0.1, ic, T_reg_GeV=0.1)
           for (nx, ny, nz), ic in zip(BATCH_EVENTS, ics)]
    eF = b.eos_energy_density(0.15 / V.HBARC)
    for k in range(3):
        if k:
            b.steps(2)
        for (nx, ny, nz), eid in zip(BATCH_EVENTS, ids):
            print(f"case mixed surface {nx}x{ny}x{nz} check {k}", flush=True)
            d4 = nz > 1
            widths = np.array([0.04, 0.1, 0.1, 0.1])[:4 if d4 else 3]
            cube_ids, corners = b.fo_flag_ids(eid, eF, with_corners=True)
            counts, el, var = b.fo_surface(eid, widths)
            rows = [C.surface_elements(_builder_order(corners[m], d4), eF, widths) for m in range(cube_ids.size)]
            assert np.array_equal(counts, [r.shape[0] for r in rows])
            want = np.concatenate(rows) if rows else np.empty((0, 2 * (4 if d4 else 3)))
            assert V.bits_equal(el, want)
            lo = np.repeat(corners.min(axis=1), counts)
            hi = np.repeat(corners.max(axis=1), counts)
            assert np.all((var[:, 3] >= lo - 1e-12 * hi) & (var[:, 3] <= hi * (1 + 1e-12)))
    b.close()


SERVE_CFG = dict(rate_per_s=4000.0, n_requests=6, seed=11, frac_3d=0.34, side2_min=31, side2_max=41, side3_min=17,
                 side3_max=21, nz3_min=9, nz3_max=11, dx=0.4, dt=0.05, t0=0.6, etabar=0.12, T_reg_GeV=0.1,
                 T_F_GeV=0.150, coarse=4, e0=10.0, e_floor=0.01, precision=C.PREC_MIXED, bulk=1, max_steps=400,
                 max_resident=8, collect_surface=2)


def _served_surfaces(budget):
    got = {}

    def on_surface(request, rows):
        got.setdefault(request, []).append(rows)

    stats, reqs = C.serve(None, on_surface=on_surface, hbm_budget_bytes=budget, **SERVE_CFG)
    return stats, reqs, got


def _cube_order_keys(rows, req, dx):
    """Each row's cube (x, then y, then eta: the reference's loop order) from
    its position: an element's centroid lies in its cube, which starts at
    i*dx - (n - 1)*dx/2.  Rows on a cube face (ambiguous) are left out."""
    dims = (req["nx"], req["ny"], req["nz"]) if req["nz"] > 1 else (req["nx"], req["ny"])
    frac = [(rows[:, 1 + k] + (n - 1) / 2 * dx) / dx for k, n in enumerate(dims)]
    inside = np.all([np.abs(f - np.round(f)) > 1e-9 for f in frac], axis=0)
    key = np.zeros(rows.shape[0], dtype=np.int64)
    for f, n in zip(frac, dims):
        key = key * n + np.floor(f).astype(np.int64)
    return key[inside], int(inside.sum())


def test_served_surfaces_same_alone_and_in_the_mix():
    """Serving with surface collection: each request's surface elements (in
    surface.dat's columns, handed out at every check) are the same, bit for
    bit, whether all requests share the batch at once or wait for the memory
    of earlier ones (room for about one 3+1D request), and each check's rows
    come in the reference's cube order (x outermost, eta innermost); fo_elements
    counts them."""
    per_cell = (3 * 15 + 3 * 4 + 3 * 4 + 4) * 4 + 2 * 8 + 20 * 8
    big = 21 * 21 * 11
    s_mix, r_mix, g_mix = _served_surfaces(64 * big * per_cell)
    s_one, r_one, g_one = _served_surfaces(int(1.05 * big) * per_cell)
    print(s_mix, s_one, flush=True)
    assert s_mix["completed"] == s_one["completed"] >= SERVE_CFG["n_requests"] - 1
    assert max(r["admit_s"] - r["arrival_s"] for r in r_one) > max(r["admit_s"] - r["arrival_s"] for r in r_mix)
    for k, (a, c) in enumerate(zip(r_mix, r_one)):
        if a["status"] != C.VH_OK:
            continue
        ra, rc = g_mix.get(k, []), g_one.get(k, [])
        assert len(ra) == len(rc) and len(ra) > 0, k
        for x, y in zip(ra, rc):
            assert V.bits_equal(x, y), k
        n = sum(x.shape[0] for x in ra)
        assert a["fo_elements"] == c["fo_elements"] == n > 0
        rows = np.concatenate(ra)
        assert np.all(np.isfinite(rows))
        placed = 0
        for x in ra:
            keys, m = _cube_order_keys(x, a, SERVE_CFG["dx"])
            assert np.all(np.diff(keys) >= 0), k
            placed += m
        assert placed > 0.99 * n, (k, placed, n)
