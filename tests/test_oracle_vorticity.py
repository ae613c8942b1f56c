"""The oracle's thermal vorticity (orc_vorticity, restating
rhic/src/Vorticity.cpp:6-171) against the reference's own
calculateThermalVorticity built from its sources (oracle/_ref), bit for bit,
after steps of 2+1D and 3+1D runs; and a closed form that fixes its
conventions: a shear flow u^x = a y at uniform temperature has
w_xy = -a / (2T) in the interior (central differences are exact for the
linear beta_x), w_ty = (d_y beta_t) / 2 = a^2 y / (2 T u^t) to the stencil's
O(dx^2), and no other component."""
import numpy as np
import pytest

import vhtest as V


@pytest.mark.parametrize("variant", ["conf", "lat", "lat_bulk"])
@pytest.mark.parametrize("nx,ny,nz", [(17, 15, 1), (13, 11, 9)])
def test_vorticity_matches_reference_build(variant, nx, ny, nz):
    if not V.ref_available(variant):
        pytest.skip("reference build not present")
    lattice, bulk = variant.startswith("lat"), variant.endswith("bulk")
    table = V.eos_table() if lattice else None
    dx, dt, t0 = 0.1, 0.02, 0.6
    ic = V.random_ic(nx, ny, nz, dx, t0, seed=4, lattice=lattice, bulk=bulk)
    ref = V.Reference(variant, nx, ny, nz, dx, dx, dx, dt, t0, 0.1)
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, 0.1, bulk=bulk, table=table)
    for x in (ref, o):
        x.set_ic(ic)
        x.init_conserved(t0)
    for n in range(1, 5):
        t = t0 + (n - 1) * dt
        ref.step(t)
        assert o.step(t) == 0
        want, got = ref.vorticity(t), o.vorticity()
        assert V.bits_equal(got, want), (n, np.argwhere(got.view(np.int64) != want.view(np.int64))[:5])
    # the time derivative and the eta derivatives act (3+1D), nothing is trivially zero
    assert np.count_nonzero(got[0]) > got.shape[1] // 2
    if nz > 1:
        assert np.count_nonzero(got[2]) > got.shape[1] // 2


def test_shear_flow_closed_form():
    nx, ny, nz, dx, a = 9, 11, 1, 0.1, 0.3
    N = nx * ny
    y = (np.arange(ny) - (ny - 1) / 2) * dx
    Y = np.repeat(y, nx)  # x fastest
    ux = a * Y
    e = np.full(N, 10.0)
    o = V.Oracle(nx, ny, nz, dx, dx, dx, 0.02, 0.6, 0.1)
    o.set_ic({"e": e, "p": e / 3, "ut": np.sqrt(1 + ux ** 2), "ux": ux, "uy": np.zeros(N), "un": np.zeros(N),
              "pi": np.zeros(10 * N)})
    w = o.vorticity()  # before any step: up = u, no time derivative
    T = float(np.float32(np.float32(10.0 / 13.8997) ** np.float32(0.25)))
    inner = np.zeros((ny, nx), bool)
    inner[1:-1, :] = True
    inner = inner.ravel()
    np.testing.assert_allclose(w[3][inner], -a / (2 * T), rtol=1e-12)
    ut = np.sqrt(1 + ux ** 2)
    np.testing.assert_allclose(w[1][inner], (a * a * Y / (2 * T * ut))[inner], rtol=2e-3, atol=1e-12)
    for c in (0, 2, 4, 5):
        assert np.all(w[c] == 0.0), c
    # at the clamped y edges the one-sided stencil sees half the slope
    np.testing.assert_allclose(w[3][~inner], -a / (4 * T), rtol=1e-12)
    np.testing.assert_allclose(w[7], -ux / T, rtol=1e-15)  # beta_x = u_x / T = -u^x / T
