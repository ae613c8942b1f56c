"""The command-line driver (cpu-vh_amd/cpuvh_run) against the oracle stepped
through the reference's run() loop (rhic/src/HydroPlugin.cpp:74-321): same
initial state from the same rhic-conf files, the same snapshot files (the
thermal vorticity's among them) byte for
byte (rhic/src/FileIO.cpp:14-59 format), the same surface.dat (every coarse
step; the previous slice starts at zero; freezeout/FreezeOut.cpp:200-361: the
cut cubes in the reference's order, elements from the surface finder pinned
by the Cornelius fixtures in test_surface.py, the ten hydro variables
interpolated at each centroid) and the same stopping step (nt+1 steps at
most; the third check with no cell above e_F ends the run)."""
import os
import re
import subprocess

import numpy as np
import pytest

import cpuvh_amd as C
import vhtest as V

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(C.LIB_PATH), "cpuvh_run")
SHEAR = [("e", "e"), ("ux", "ux"), ("uy", "uy"), ("ut", "ut"), ("pixx", "pixx"), ("pixy", "pixy"),
         ("pinn", "pinn"), ("regFacShear", 25)]
BULK = [("Pi", "Pi"), ("regFacBulk", 26)]


def _write_conf(d, lat, ic, hyd, gpu):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "lattice.properties"), "w") as f:
        f.write("numLatticePointsX={nx}\nnumLatticePointsY={ny}\nnumLatticePointsRapidity={nz}\n"
                "numProperTimePoints={nt}\nlatticeSpacingX={dx}\nlatticeSpacingY={dx}\n"
                "latticeSpacingRapidity={dx}\nlatticeSpacingProperTime={dt}\n"
                "coarseProperTimeFOFactorFO={coarse}".format(**lat))  # the reference's own misspelt last key
    with open(os.path.join(d, "ic.properties"), "w") as f:
        f.write("initialConditionType={type}\nnumberOfSourceFiles=0\nnumberOfNucleonsPerNuclei=197\n"
                "initialEnergyDensity=1.0\nscatteringCrossSectionNN=42.0\nimpactParameter=0.0\n"
                "fractionOfBinaryCollisions=1.0\nrapidityVariance=1.96\nrapidityMean=4.0\nsourceType=0"
                .format(**ic))
    with open(os.path.join(d, "hydro.properties"), "w") as f:
        f.write("initialProperTimePoint = {t0}\nshearViscosityToEntropyDensity = {etas}\n"
                "freezeoutTemperatureGeV = {TF}\ninitializePimunuNavierStokes=0\n"
                "regulationTemperatureGeV={Treg}\ndoFreezeOut=1".format(**hyd))
    with open(os.path.join(d, "gpu.properties"), "w") as f:
        f.write("# options the reference fixes at compile time\n")
        for k, v in gpu.items():
            f.write(f"{k}={v}\n")


def _read_tmunu_f32(path, nx, ny, nz):
    """Input/Tmunu.dat as the reference reads it: %f into float, i outer, k inner."""
    rows = np.loadtxt(path, dtype=np.float32).astype(np.float64)
    N = nx * ny * nz
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    s = (i + nx * (j + ny * k)).ravel()
    out = {}
    for c, name in enumerate(["e", "p", "ut", "ux", "uy", "un"]):
        a = np.empty(N)
        a[s] = rows[:, 3 + c]
        out[name] = a
    pi = np.empty((10, N))
    for c in range(10):
        pi[c, s] = rows[:, 9 + c]
    out["pi"] = pi.ravel()
    out["Pi"] = np.empty(N)
    out["Pi"][s] = rows[:, 19]
    return out


def _snapshot_text(v, nx, ny, nz, dx):
    lines = []
    s = 0
    for k in range(nz):
        z = (k - (nz - 1) / 2.) * dx
        for j in range(ny):
            y = (j - (ny - 1) / 2.) * dx
            for i in range(nx):
                x = (i - (nx - 1) / 2.) * dx
                lines.append("%.3f\t%.3f\t%.3f\t%.8f\n" % (x, y, z, v[s]))
                s += 1
    return "".join(lines)


def _interp(fr, a):
    """FreezeOut.cpp:22-56, the terms in the reference's order; a[t][x][y][z]."""
    if len(fr) == 3:
        x0, x1, x2 = fr
        return (((1 - x0) * (1 - x1) * (1 - x2) * a[0][0][0]) + ((x0) * (1 - x1) * (1 - x2) * a[1][0][0])
                + ((1 - x0) * (x1) * (1 - x2) * a[0][1][0]) + ((1 - x0) * (1 - x1) * (x2) * a[0][0][1])
                + ((x0) * (x1) * (1 - x2) * a[1][1][0]) + ((x0) * (1 - x1) * (x2) * a[1][0][1])
                + ((1 - x0) * (x1) * (x2) * a[0][1][1]) + ((x0) * (x1) * (x2) * a[1][1][1]))
    x0, x1, x2, x3 = fr
    terms = [((1 - x0) * (1 - x1) * (1 - x2) * (1 - x3) * a[0][0][0][0]),
             ((x0) * (1 - x1) * (1 - x2) * (1 - x3) * a[1][0][0][0]),
             ((1 - x0) * (x1) * (1 - x2) * (1 - x3) * a[0][1][0][0]),
             ((1 - x0) * (1 - x1) * (x2) * (1 - x3) * a[0][0][1][0]),
             ((1 - x0) * (1 - x1) * (1 - x2) * (x3) * a[0][0][0][1]),
             ((x0) * (x1) * (1 - x2) * (1 - x3) * a[1][1][0][0]),
             ((x0) * (1 - x1) * (x2) * (1 - x3) * a[1][0][1][0]),
             ((x0) * (1 - x1) * (1 - x2) * (x3) * a[1][0][0][1]),
             ((1 - x0) * (x1) * (x2) * (1 - x3) * a[0][1][1][0]),
             ((1 - x0) * (x1) * (1 - x2) * (x3) * a[0][1][0][1]),
             ((1 - x0) * (1 - x1) * (x2) * (x3) * a[0][0][1][1]),
             ((x0) * (x1) * (x2) * (1 - x3) * a[1][1][1][0]),
             ((x0) * (x1) * (1 - x2) * (x3) * a[1][1][0][1]),
             ((x0) * (1 - x1) * (x2) * (x3) * a[1][0][1][1]),
             ((1 - x0) * (x1) * (x2) * (x3) * a[0][1][1][1]),
             ((x0) * (x1) * (x2) * (x3) * a[1][1][1][1])]
    r = terms[0]
    for v in terms[1:]:
        r = r + v
    return r


FO_FIELDS = ["ux", "uy", "un", "e", "pixx", "pixy", "pixn", "piyy", "piyn", "Pi"]


def _surface_lines(ids, slices, dim, nx, ny, nz, dx, spacing, cell_tau, t, eF, L, eos):
    """callFOFinder2p1D/3p1D (FreezeOut.cpp:200-361) for the cut cubes, in the
    reference's order (x outer, eta inner)."""
    cx, cy = nx - 1, ny - 1
    cubes = sorted((c % cx, (c // cx) % cy, c // (cx * cy) if dim == 4 else 0) for c in ids.tolist())
    nzc = 2 if dim == 4 else 1
    lines = []
    for ix, iy, iz in cubes:
        def at(x, y, z):
            return (ix + x) + nx * ((iy + y) + ny * (iz + z))
        corners = [slices[it][3][at(x, y, z)] for it in range(2) for x in range(2) for y in range(2)
                   for z in range(nzc)]
        cell = [ix * dx - ((nx - 1) / 2.0 * dx), iy * dx - ((ny - 1) / 2.0 * dx),
                iz * dx - ((nz - 1) / 2.0 * dx) if dim == 4 else 0.0]
        for row in C.surface_elements(np.array(corners), eF, spacing):
            cen, nor = row[:dim], row[dim:]
            fr = [cen[k] / spacing[k] for k in range(dim)]
            var = []
            for v in range(10):
                if dim == 4:
                    a = [[[[slices[it][v][at(x, y, z)] for z in range(2)] for y in range(2)] for x in range(2)]
                         for it in range(2)]
                else:
                    a = [[[slices[it][v][at(x, y, 0)] for y in range(2)] for x in range(2)] for it in range(2)]
                var.append(_interp(fr, a))
            pos = [cen[0] + cell_tau, cen[1] + cell[0], cen[2] + cell[1], cen[3] + cell[2] if dim == 4 else 0.0]
            ds = [t * nor[k] for k in range(dim)] + ([0.0] if dim == 3 else [])
            eps = var[3]
            vals = pos + ds + var[0:3] + [eps, L.orc_eos_temperature(eos, eps), L.orc_eos_pressure(eos, eps)] + var[4:]
            lines.append(" ".join("%g" % x for x in vals) + "\n")
    return lines


def _oracle_loop(lat, hyd, bulk, table, ic):
    """The reference's run() loop with the oracle as the step."""
    nx, ny, nz, nt, dx, dt, coarse = (lat[k] for k in ("nx", "ny", "nz", "nt", "dx", "dt", "coarse"))
    t0 = hyd["t0"]
    o = V.Oracle(nx, ny, nz, dx, dx, dx, dt, t0, hyd["etas"], T_reg_GeV=hyd["Treg"], bulk=bulk, table=table)
    o.set_ic(ic)
    o.init_conserved(t0)
    L = V.oracle_lib()
    eos = L.orc_eos_create(None if table is None else table.ctypes.data)
    eF = L.orc_eos_energy_density(eos, hyd["TF"] / 0.197326938)
    dim = 4 if (nx > 1 and ny > 1 and nz > 1) else (3 if (nx > 1 and ny > 1) else 0)
    N = nx * ny * nz
    slices = [[np.zeros(N) for _ in FO_FIELDS], [np.zeros(N) for _ in FO_FIELDS]]
    spacing = np.array([dt * coarse, dx, dx, dx][:dim])
    surface, files = [], {}
    t, zero, nsteps = t0, 0, 0
    for n in range(1, nt + 2):
        if (n - 1) % lat["snap"] == 0:
            for name, f in SHEAR + (BULK if bulk else []):
                v = np.zeros(N) if (n == 1 and f in (25, 26)) else o.get(f)
                files["%s_%.3f.dat" % (name, t)] = _snapshot_text(v, nx, ny, nz, dx)
            # the thermal vorticity of the current state (zeros before the first step)
            w = np.zeros((10, N)) if n == 1 else o.vorticity()
            for name, c in (("wxy", 3), ("beta_x", 7), ("beta_y", 8)):
                files["%s_%.3f.dat" % (name, t)] = _snapshot_text(w[c], nx, ny, nz, dx)
        if n % coarse == 0:
            e = o.get("e")
            if dim:
                slices = [slices[1], [o.get(f) if (f != "Pi" or bulk) else np.zeros(N) for f in FO_FIELDS]]
                ids = V.oracle_fo_flag(slices[0][3], e, nx, ny, nz, eF)
                surface.extend(_surface_lines(ids, slices, dim, nx, ny, nz, dx, spacing, t0 + n * dt, t, eF, L, eos))
            if int(np.count_nonzero(e > eF)) == 0:
                zero += 1
            if zero >= 3:
                break
        assert o.step(t) == 0
        nsteps += 1
        t = t0 + n * dt
    L.orc_eos_destroy(eos)
    files["surface.dat"] = "".join(surface)
    return files, nsteps


def _run_case(tmp_path, lat, hyd, gpu, ic_fields, bulk, table, ic_type=11):
    """ic_type 11: ic_fields go through input/Tmunu.dat; otherwise the driver
    builds the initial state itself and ic_fields is what it should be."""
    conf, out = str(tmp_path / "conf"), str(tmp_path / "out")
    os.makedirs(out)
    _write_conf(conf, lat, {"type": ic_type}, hyd, gpu)
    nx, ny, nz, dx = lat["nx"], lat["ny"], lat["nz"], lat["dx"]
    if ic_type == 11:
        os.makedirs(tmp_path / "input")
        C.write_tmunu(str(tmp_path / "input" / "Tmunu.dat"), nx, ny, nz, dx, dx, dx, ic_fields)
    print(f"case driver type={ic_type} {nx}x{ny}x{nz} nt={lat['nt']} bulk={bulk} table={table is not None}",
          flush=True)
    r = subprocess.run([DRIVER, "--hydro", "--config", conf, "--output", out], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    steps = int(re.search(r"Steps taken: (\d+)", r.stdout).group(1))
    if ic_type == 11:
        ic = _read_tmunu_f32(str(tmp_path / "input" / "Tmunu.dat"), nx, ny, nz)
        if not bulk:
            ic.pop("Pi")
    else:
        ic = dict(ic_fields)
    want, want_steps = _oracle_loop(lat, hyd, bulk, table, ic)
    assert steps == want_steps, (steps, want_steps)
    got = sorted(os.listdir(out))
    assert got == sorted(want), (got, sorted(want))
    for name, text in want.items():
        with open(os.path.join(out, name)) as f:
            assert f.read() == text, name
    return steps, r.stdout


def test_driver_2d_stops_on_third_cold_check(tmp_path):
    lat = dict(nx=21, ny=17, nz=1, nt=60, dx=0.2, dt=0.05, coarse=2, snap=10)
    hyd = dict(t0=0.6, etas=0.2, TF=0.172, Treg=0.1)
    ic = V.random_ic(21, 17, 1, 0.2, 0.6, seed=31)
    steps, log = _run_case(tmp_path, lat, hyd, {"snapshotEvery": 10}, ic, False, None)
    assert steps < lat["nt"] + 1 and "All cells have dropped below freezeout energy density" in log


def test_driver_3d_table_bulk_runs_nt_plus_one_steps(tmp_path):
    lat = dict(nx=11, ny=9, nz=7, nt=12, dx=0.2, dt=0.02, coarse=3, snap=5)
    hyd = dict(t0=0.6, etas=0.12, TF=0.150, Treg=0.15)
    table = V.eos_table()
    ic = V.random_ic(11, 9, 7, 0.2, 0.6, seed=32, lattice=True, bulk=True)
    eos_file = os.path.join(os.path.dirname(os.path.dirname(C.LIB_PATH)), "EoS", "hrg_hotqcd_eos_binary.dat")
    steps, _ = _run_case(tmp_path, lat, hyd, {"bulk": 1, "eos": "table", "eosFile": eos_file, "snapshotEvery": 5},
                         ic, True, table)
    assert steps == lat["nt"] + 1


def test_driver_ideal_gubser_initial_state(tmp_path):
    """IC type 3 built by the driver (initialEnergyDensity = 1, the Gubser
    flow at tau = 1 with u = 0 at the centre cell): same files as the oracle
    loop started from the same state."""
    lat = dict(nx=21, ny=21, nz=1, nt=20, dx=0.2, dt=0.02, coarse=4, snap=10)
    hyd = dict(t0=1.0, etas=0.08, TF=0.06, Treg=0.0)
    ic = C.analytic_ic(3, 21, 21, 1, 0.2, 0.2, e0=1.0)
    ic["pi"] = np.zeros(10 * 21 * 21)
    steps, log = _run_case(tmp_path, lat, hyd, {"snapshotEvery": 10}, ic, False, None, ic_type=3)
    assert steps == lat["nt"] + 1 and "Ideal hydrodynamic Gubser flow test." in log


def test_driver_israel_stewart_gubser_file(tmp_path):
    """IC type 1: the driver reads
    rhic/rhic-trunk/src/test/resources/gubser/viscous/gubserIC.dat under the
    working directory (the reference does not ship the file; it is made here
    from the semi-analytic Israel-Stewart Gubser solution at tau = 1,
    eta/s = 0.2), with regulation below 100 MeV; same files as the oracle
    loop."""
    import gubser as Gb
    from test_gubser import write_gubser_ic
    d = tmp_path / "rhic" / "rhic-trunk" / "src" / "test" / "resources" / "gubser" / "viscous"
    d.mkdir(parents=True)
    write_gubser_ic(str(d / "gubserIC.dat"), Gb.solve(0.2), 21, 0.2)
    ic = C.read_gubser(str(d / "gubserIC.dat"), 21, 21, 1)
    lat = dict(nx=21, ny=21, nz=1, nt=20, dx=0.2, dt=0.02, coarse=4, snap=10)
    hyd = dict(t0=1.0, etas=0.2, TF=0.15, Treg=0.1)
    steps, log = _run_case(tmp_path, lat, hyd, {"snapshotEvery": 10}, ic, False, None, ic_type=1)
    assert steps == lat["nt"] + 1 and "Gubser" in log


def test_driver_type10_field_files_table_bulk(tmp_path):
    """IC type 10: input/e.dat, ux.dat, uy.dat, un.dat and the ten pi files
    (read through float, i outer), p from the table EoS, Pi = e/3 - p with
    bulk (rhic/src/InitialConditions.cpp:108-466), u^tau normalised; 3+1D
    with u^eta != 0.  Same files as the oracle loop from that state."""
    nx, ny, nz, dx, t0 = 11, 9, 5, 0.2, 0.6
    N = nx * ny * nz
    ref = V.random_ic(nx, ny, nz, dx, t0, seed=33, lattice=True, bulk=True)
    f32 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
    vals = {"e": f32(ref["e"]), "ux": f32(ref["ux"]), "uy": f32(ref["uy"]), "un": f32(ref["un"])}
    names = ["pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn", "piyy", "piyn", "pinn"]
    for c, n in enumerate(names):
        vals[n] = f32(ref["pi"][c * N:(c + 1) * N])
    os.makedirs(tmp_path / "input")
    for n, v in vals.items():
        with open(tmp_path / "input" / (n + ".dat"), "w") as f:
            for i in range(nx):
                for j in range(ny):
                    for k in range(nz):
                        f.write("%.3f %.3f %.3f %.9g\n" % (i, j, k, v[i + nx * (j + ny * k)]))
    table = V.eos_table()
    L = V.oracle_lib()
    eos = L.orc_eos_create(table.ctypes.data)
    p = np.array([L.orc_eos_pressure(eos, x) for x in vals["e"]])
    L.orc_eos_destroy(eos)
    ux, uy, un = vals["ux"], vals["uy"], vals["un"]
    ic = {"e": vals["e"], "p": p, "ux": ux, "uy": uy, "un": un,
          "ut": np.sqrt(((1.0 + ux * ux) + uy * uy) + ((t0 * t0) * un) * un),
          "pi": np.concatenate([vals[n] for n in names]), "Pi": vals["e"] / 3. - p}
    lat = dict(nx=nx, ny=ny, nz=nz, nt=10, dx=dx, dt=0.02, coarse=3, snap=5)
    hyd = dict(t0=t0, etas=0.12, TF=0.150, Treg=0.0)
    eos_file = os.path.join(os.path.dirname(os.path.dirname(C.LIB_PATH)), "EoS", "hrg_hotqcd_eos_binary.dat")
    steps, log = _run_case(tmp_path, lat, hyd, {"bulk": 1, "eos": "table", "eosFile": eos_file, "snapshotEvery": 5},
                           ic, True, table, ic_type=10)
    assert steps == lat["nt"] + 1 and "input/e.dat" in log


HYDRO_DEMO = os.path.join(os.path.dirname(C.LIB_PATH), "hydro_vectors")


@pytest.mark.parametrize("ic_type,reference_ut", [(13, False), (14, False), (14, True)])
def test_hydro_wrapper_vectors(tmp_path, ic_type, reference_ut):
    """The reference's HYDRO class (include/cpuvh_hydro.hpp) with its vectors
    handed over in memory (cpu-vh_amd/examples/hydro_vectors): type 13, an
    energy density in GeV/fm^3 at rest; type 14, all sixteen vectors; run by
    run_hydro_no_cli() from rhic-conf/ into cpu-vh_output/.  Same files as the
    oracle loop from the state the reference builds from those vectors
    (rhic/src/InitialConditions.cpp:1593-1711; u^tau normalised, or with
    HYDRO::reference_ut the reference's 1 + ux^2 + uy^2 + un^2 of :1681)."""
    nx, ny, nz, dx, t0 = 17, 15, 1, 0.2, 0.6
    N = nx * ny * nz
    ref = V.random_ic(nx, ny, nz, dx, t0, seed=40 + ic_type, bulk=True)
    hbarc = 0.197326938
    e_gev = ref["e"] * hbarc
    lat = dict(nx=nx, ny=ny, nz=nz, nt=12, dx=dx, dt=0.02, coarse=3, snap=5)
    hyd = dict(t0=t0, etas=0.2, TF=0.150, Treg=0.0)
    _write_conf(str(tmp_path / "rhic-conf"), lat, {"type": ic_type}, hyd, {"bulk": 1, "snapshotEvery": 5})
    (tmp_path / "cpu-vh_output").mkdir()
    e = e_gev / hbarc + 1.0e-3
    ic = {"e": e, "p": e / 3.0, "ut": np.ones(N), "ux": np.zeros(N), "uy": np.zeros(N), "un": np.zeros(N),
          "pi": np.zeros(10 * N), "Pi": np.zeros(N)}
    if ic_type == 13:
        rows = e_gev[:, None]
    else:
        ux = np.where(np.abs(ref["ux"]) < 1e-7, 0.0, ref["ux"])
        uy = np.where(np.abs(ref["uy"]) < 1e-7, 0.0, ref["uy"])
        un = np.zeros(N)
        pi_gev = ref["pi"].reshape(10, N) * hbarc
        Pi_gev = ref["Pi"] * hbarc
        rows = np.column_stack([e_gev, np.ones(N), ux, uy, un, pi_gev.T, Pi_gev])
        ut = (1.0 + ux * ux + uy * uy + un * un if reference_ut
              else np.sqrt(1.0 + ux * ux + uy * uy + t0 * t0 * un * un))
        ic.update(ux=ux, uy=uy, un=un, ut=ut, pi=(pi_gev / hbarc).ravel(), Pi=Pi_gev / hbarc)
    np.savetxt(tmp_path / "vectors.txt", rows, fmt="%.17g")
    print(f"case hydro wrapper type={ic_type} reference_ut={reference_ut}", flush=True)
    r = subprocess.run([HYDRO_DEMO, str(tmp_path / "vectors.txt")] + (["--reference-ut"] if reference_ut else []),
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    steps = int(re.search(r"Steps taken: (\d+)", r.stdout).group(1))
    want, want_steps = _oracle_loop(lat, hyd, True, None, ic)
    assert steps == want_steps == lat["nt"] + 1
    out = tmp_path / "cpu-vh_output"
    assert sorted(os.listdir(out)) == sorted(want)
    for name, text in want.items():
        with open(out / name) as f:
            assert f.read() == text, name


def test_driver_runs_reference_default_configuration(tmp_path):
    """The reference's own rhic-conf directory (copied as data): IC type 2
    (optical Glauber Pb+Pb, e0 = 70), 101 x 101, conformal EoS, regulation
    below 100 MeV.  The run ends on the third cold check well before nt, the
    first e snapshot is the optical Glauber profile, and surface.dat holds
    20-column elements whose d sigma_tau is negative (the surface closes in
    time) somewhere."""
    conf = os.path.join(V.REPO, "tests", "golden", "rhic-conf")
    out = tmp_path / "out"
    out.mkdir()
    print("case driver reference rhic-conf", flush=True)
    r = subprocess.run([DRIVER, "--hydro", "--config", conf, "--output", str(out)], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    steps = int(re.search(r"Steps taken: (\d+)", r.stdout).group(1))
    assert "All cells have dropped below freezeout energy density" in r.stdout and steps < 10001
    e0 = np.loadtxt(out / "e_0.500.dat")[:, 3]
    want = C.optical_glauber(101, 101, 1, 0.2, 0.2, 0.1, A=208, b=0.0, sigma_nn_mb=62.0, alpha=1.0, e0=70.0,
                             eta_flat=5.9, eta_variance=0.16)
    assert np.allclose(e0, want, rtol=0, atol=5.1e-9)  # written with %.8f
    surf = np.loadtxt(out / "surface.dat")
    assert surf.ndim == 2 and surf.shape[1] == 20 and surf.shape[0] > 100
    assert (surf[:, 4] < 0).any() and np.all(np.isfinite(surf))
