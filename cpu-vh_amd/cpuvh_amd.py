"""ctypes binding of libcpuvh_amd.so (include/cpuvh_amd.h).

Python-side convenience over the C ABI of the MI355X viscous-hydro step:
tests and bench.py drive the product through exactly the entry points a
C/C++ caller uses.  Loading fails loudly when the built library is missing;
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpuvh_amd.so")

FIELD_NAMES = ["ttt", "ttx", "tty", "ttn", "pitt", "pitx", "pity", "pitn", "pixx", "pixy", "pixn",
               "piyy", "piyn", "pinn", "Pi", "e", "p", "ut", "ux", "uy", "un",
               "up_t", "up_x", "up_y", "up_n", "regfac", "regfac_bulk",
               "wtx", "wty", "wtn", "wxy", "wxn", "wyn", "beta_t", "beta_x", "beta_y", "beta_n"]
FIELD = {n: i for i, n in enumerate(FIELD_NAMES)}

VH_OK, VH_ERR_ARG, VH_ERR_HIP, VH_ERR_NOMEM, VH_ERR_NAN, VH_ERR_STATE, VH_ERR_CAPACITY = 0, -1, -2, -3, -4, -5, -6
EOS_CONFORMAL, EOS_TABLE = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

# The exported entry points, as declared in include/cpuvh_amd.h.
EXPORTS = ["vh_create", "vh_destroy", "vh_upload_ic", "vh_init_conserved", "vh_step", "vh_steps", "vh_sync",
           "vh_nan_info", "vh_download", "vh_upload_field", "vh_count_above", "vh_fo_flag", "vh_fo_interp",
           "vh_eos_energy_density", "vh_eos_pressure", "vh_eos_temperature", "vh_last_kernel_ms",
           "vh_inferred_variables", "vh_fo_flag_noalloc", "vh_fo_surface", "vh_thermal_vorticity",
           "vh_status_string", "vh_build_info", "vh_ic_mc_glauber", "vh_ic_read_tmunu", "vh_ic_write_tmunu",
           "vh_prepare_steps", "vh_batch_create", "vh_batch_destroy", "vh_batch_add", "vh_batch_remove",
           "vh_batch_steps", "vh_batch_sync", "vh_batch_nan_info", "vh_batch_clear_error",
           "vh_batch_counts_above", "vh_batch_download", "vh_batch_fo_flag", "vh_batch_fo_flags", "vh_batch_fo_surface", "vh_batch_fo_surfaces",
           "vh_batch_event_info",
           "vh_batch_free_cells", "vh_batch_last_ms", "vh_batch_eos_energy_density", "vh_batch_autotune",
           "vh_batch_set_tiles", "vh_batch_set_face_mode", "vh_serve_run", "vh_ic_mc_glauber_classic",
           "vh_surface_square", "vh_surface_cube", "vh_surface_hypercube",
           "vh_ic_optical_glauber", "vh_ic_test_problem", "vh_ic_read_fields", "vh_ic_read_gubser",
           "vh_batch_run"]


class VhConfig(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("dt", ctypes.c_double), ("t0", ctypes.c_double), ("etabar", ctypes.c_double),
                ("T_reg_GeV", ctypes.c_double), ("bulk", ctypes.c_int), ("eos_kind", ctypes.c_int),
                ("eos_table", ctypes.c_void_p),
                ("etas_Tk_GeV", ctypes.c_double), ("etas_slope_lo", ctypes.c_double),
                ("etas_slope_hi", ctypes.c_double), ("device", ctypes.c_int), ("cs2_mode", ctypes.c_int),
                ("zeta_Tpeak_GeV", ctypes.c_double), ("zeta_peak", ctypes.c_double)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("capacity_cells", ctypes.c_int64), ("max_events", ctypes.c_int),
                ("max_steps", ctypes.c_int), ("bulk", ctypes.c_int), ("eos_kind", ctypes.c_int),
                ("eos_table", ctypes.c_void_p), ("precision", ctypes.c_int), ("cs2_mode", ctypes.c_int),
                ("fo_vars", ctypes.c_int)]


class EventParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("dt", ctypes.c_double), ("t0", ctypes.c_double), ("etabar", ctypes.c_double),
                ("T_reg_GeV", ctypes.c_double), ("etas_Tk_GeV", ctypes.c_double),
                ("etas_slope_lo", ctypes.c_double), ("etas_slope_hi", ctypes.c_double),
                ("zeta_Tpeak_GeV", ctypes.c_double), ("zeta_peak", ctypes.c_double)]


SURFACE_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                    ctypes.POINTER(ctypes.c_double))


class ServeConfig(ctypes.Structure):
    _fields_ = [("rate_per_s", ctypes.c_double), ("n_requests", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("frac_3d", ctypes.c_double), ("side2_min", ctypes.c_int), ("side2_max", ctypes.c_int),
                ("side3_min", ctypes.c_int), ("side3_max", ctypes.c_int), ("nz3_min", ctypes.c_int),
                ("nz3_max", ctypes.c_int), ("dx", ctypes.c_double), ("dt", ctypes.c_double), ("t0", ctypes.c_double),
                ("etabar", ctypes.c_double), ("etas_Tk_GeV", ctypes.c_double), ("etas_slope_lo", ctypes.c_double),
                ("etas_slope_hi", ctypes.c_double), ("T_reg_GeV", ctypes.c_double), ("T_F_GeV", ctypes.c_double),
                ("coarse", ctypes.c_int), ("e0", ctypes.c_double), ("e_floor", ctypes.c_double),
                ("hbm_budget_bytes", ctypes.c_int64), ("precision", ctypes.c_int), ("bulk", ctypes.c_int),
                ("max_steps", ctypes.c_int), ("max_resident", ctypes.c_int), ("collect_surface", ctypes.c_int),
                ("device", ctypes.c_int), ("digest", ctypes.c_int), ("on_surface", SURFACE_CALLBACK),
                ("user", ctypes.c_void_p)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("steps", ctypes.c_int),
                ("status", ctypes.c_int), ("glauber_seed", ctypes.c_uint64), ("fo_cubes", ctypes.c_int64),
                ("e_sum", ctypes.c_double), ("digest", ctypes.c_uint64), ("fo_elements", ctypes.c_int64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_int), ("failed", ctypes.c_int), ("wall_s", ctypes.c_double),
                ("events_per_s", ctypes.c_double), ("zone_upd_per_s", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("mean_s", ctypes.c_double), ("occupancy_mean", ctypes.c_double),
                ("occupancy_peak", ctypes.c_double)]


class GlauberParams(ctypes.Structure):
    _fields_ = [("A", ctypes.c_int), ("R", ctypes.c_double), ("a", ctypes.c_double),
                ("sigma_nn_fm2", ctypes.c_double), ("b", ctypes.c_double), ("b_max", ctypes.c_double),
                ("width", ctypes.c_double), ("eta_flat", ctypes.c_double), ("eta_variance", ctypes.c_double),
                ("e0", ctypes.c_double), ("e_floor", ctypes.c_double), ("seed", ctypes.c_uint64)]


class GlauberInfo(ctypes.Structure):
    _fields_ = [("b", ctypes.c_double), ("n_part", ctypes.c_int), ("n_coll", ctypes.c_int)]


class VhError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: status {status} ({_lib().vh_status_string(status).decode()})")


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        L.vh_create.argtypes = [ctypes.POINTER(VhConfig), ctypes.POINTER(P)]
        L.vh_destroy.argtypes = [P]
        L.vh_upload_ic.argtypes = [P] + [P] * 8
        L.vh_init_conserved.argtypes = [P, D]
        L.vh_step.argtypes = [P, D]
        L.vh_steps.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.vh_sync.argtypes = [P]
        L.vh_nan_info.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)]
        L.vh_download.argtypes = [P, ctypes.c_int, P]
        L.vh_upload_field.argtypes = [P, ctypes.c_int, P]
        L.vh_count_above.argtypes = [P, D, ctypes.POINTER(ctypes.c_int64)]
        L.vh_fo_flag.argtypes = [P, D, P, P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.vh_fo_interp.argtypes = [P, P, P, ctypes.c_int64, P]
        L.vh_fo_surface.argtypes = [P, P, P, P, P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.vh_thermal_vorticity.argtypes = [P]
        for f in ("vh_eos_energy_density", "vh_eos_pressure", "vh_eos_temperature"):
            getattr(L, f).argtypes = [P, D]
            getattr(L, f).restype = D
        L.vh_last_kernel_ms.argtypes = [P]
        L.vh_last_kernel_ms.restype = D
        L.vh_inferred_variables.argtypes = [P, D, P, P, ctypes.c_int64, P, P]
        L.vh_fo_flag_noalloc.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, P, ctypes.c_int64, P, P]
        L.vh_status_string.argtypes = [ctypes.c_int]
        L.vh_status_string.restype = ctypes.c_char_p
        L.vh_build_info.restype = ctypes.c_char_p
        L.vh_ic_mc_glauber.argtypes = [ctypes.POINTER(GlauberParams), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       D, D, D, P, ctypes.POINTER(GlauberInfo)]
        L.vh_ic_mc_glauber_classic.argtypes = [ctypes.c_int, D, D, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               D, D, D, P, ctypes.POINTER(ctypes.c_int)]
        L.vh_surface_square.argtypes = [P, D, P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.vh_surface_cube.argtypes = [P, D, P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.vh_surface_hypercube.argtypes = [P, D, P, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.vh_ic_optical_glauber.argtypes = [ctypes.c_int, D, D, D, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            D, D, D, P]
        L.vh_ic_test_problem.argtypes = [ctypes.c_int, D, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D] + [P] * 6
        L.vh_ic_read_fields.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, D] + [P] * 6
        L.vh_ic_read_gubser.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [P] * 7
        L.vh_ic_read_tmunu.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [P] * 8
        L.vh_ic_write_tmunu.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D, D] + [P] * 8
        L.vh_prepare_steps.argtypes = [P, ctypes.c_int]
        I, I64 = ctypes.c_int, ctypes.c_int64
        L.vh_batch_create.argtypes = [ctypes.POINTER(BatchConfig), ctypes.POINTER(P)]
        L.vh_batch_destroy.argtypes = [P]
        L.vh_batch_add.argtypes = [P, ctypes.POINTER(EventParams)] + [P] * 8 + [ctypes.POINTER(I)]
        L.vh_batch_remove.argtypes = [P, I]
        L.vh_batch_steps.argtypes = [P, I]
        L.vh_batch_sync.argtypes = [P]
        L.vh_batch_nan_info.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I64)]
        L.vh_batch_clear_error.argtypes = [P]
        L.vh_batch_counts_above.argtypes = [P, D, P]
        L.vh_batch_run.argtypes = [P, D, I, I, P, P]
        L.vh_batch_download.argtypes = [P, I, I, P]
        L.vh_batch_fo_flag.argtypes = [P, I, D, P, P, I64, ctypes.POINTER(I64)]
        L.vh_batch_fo_surface.argtypes = [P, I, P, P, P, P, I64, ctypes.POINTER(I64)]
        L.vh_batch_fo_surfaces.argtypes = [P, I, P, P, P, P, P, P, P]
        L.vh_batch_fo_flags.argtypes = [P, I, P, ctypes.c_double, P, P, P]
        L.vh_batch_event_info.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I64)]
        L.vh_batch_free_cells.argtypes = [P, ctypes.POINTER(I64), ctypes.POINTER(I64)]
        L.vh_batch_last_ms.argtypes = [P]
        L.vh_batch_last_ms.restype = D
        L.vh_batch_eos_energy_density.argtypes = [P, D]
        L.vh_batch_eos_energy_density.restype = D
        L.vh_batch_autotune.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(I)]
        L.vh_batch_set_tiles.argtypes = [P, I, I]
        L.vh_batch_set_face_mode.argtypes = [P, I]
        L.vh_serve_run.argtypes = [ctypes.POINTER(ServeConfig), P, ctypes.POINTER(ServeStats),
                                   ctypes.POINTER(ServeRequest)]
        _LIB = L
    return _LIB


def load_eos_table(path):
    raw = np.fromfile(path, dtype="<f8")
    if raw.size != 400000:
        raise ValueError(f"{path}: expected 100000 rows of (e, P, s, T)")
    return np.ascontiguousarray(raw)


def _ic_arrays(arrs, N):
    """e, p, ut, ux, uy, un, pi (10 N), Pi as contiguous float64 arrays of the
    lattice's size; None stays None (e and p are required)."""
    out = []
    for k, a in enumerate(arrs):
        if a is None:
            if k < 2:
                raise ValueError("e and p are required")
            out.append(None)
            continue
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != (10 * N if k == 6 else N):
            raise ValueError("initial-condition arrays do not match the lattice")
        out.append(a)
    return out


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Hydro:
    """One lattice on one GPU (a vh_handle)."""

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, t0, etabar, T_reg_GeV=0.0, bulk=False,
                 eos_table=None, etas_Tk_GeV=0.154, etas_slope_lo=0.0, etas_slope_hi=0.0, device=0,
                 cs2_mode=0, zeta_Tpeak_GeV=0.0, zeta_peak=0.0):
        L = _lib()
        self.shape = (nz, ny, nx)
        self.N = nx * ny * nz
        self.nc = 15 if bulk else 14
        self._table = None if eos_table is None else np.ascontiguousarray(eos_table, dtype=np.float64)
        cfg = VhConfig(nx, ny, nz, dx, dy, dz, dt, t0, etabar, T_reg_GeV, int(bool(bulk)),
                       EOS_CONFORMAL if self._table is None else EOS_TABLE,
                       None if self._table is None else self._table.ctypes.data,
                       etas_Tk_GeV, etas_slope_lo, etas_slope_hi, device, cs2_mode, zeta_Tpeak_GeV, zeta_peak)
        self.cfg = cfg
        h = ctypes.c_void_p()
        self._check(L.vh_create(ctypes.byref(cfg), ctypes.byref(h)), "vh_create")
        self.h = h

    @staticmethod
    def _check(st, what):
        if st != VH_OK:
            raise VhError(st, what)

    def close(self):
        if getattr(self, "h", None):
            _lib().vh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_ic(self, e, p, ut, ux, uy, un, pi10, Pi=None):
        """x-fastest arrays; any of ut .. Pi may be None (u^tau 1, the others 0)."""
        arrs = _ic_arrays([e, p, ut, ux, uy, un, pi10, Pi], self.N)
        self._check(_lib().vh_upload_ic(self.h, *[_ptr(a) for a in arrs]), "vh_upload_ic")

    def init_conserved(self, t):
        self._check(_lib().vh_init_conserved(self.h, t), "vh_init_conserved")

    def step(self, t):
        self._check(_lib().vh_step(self.h, t), "vh_step")

    def steps(self, n_first, count):
        self._check(_lib().vh_steps(self.h, n_first, count), "vh_steps")

    def prepare_steps(self, count):
        self._check(_lib().vh_prepare_steps(self.h, count), "vh_prepare_steps")

    def sync(self):
        self._check(_lib().vh_sync(self.h), "vh_sync")

    def nan_info(self):
        site = ctypes.c_int()
        cell = ctypes.c_int64()
        self._check(_lib().vh_nan_info(self.h, ctypes.byref(site), ctypes.byref(cell)), "vh_nan_info")
        return site.value, cell.value

    def get(self, field):
        f = FIELD[field] if isinstance(field, str) else field
        out = np.empty(self.N, dtype=np.float64)
        self._check(_lib().vh_download(self.h, f, _ptr(out)), "vh_download")
        return out

    def set(self, field, values):
        f = FIELD[field] if isinstance(field, str) else field
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._check(_lib().vh_upload_field(self.h, f, _ptr(a)), "vh_upload_field")

    def count_above(self, eF):
        c = ctypes.c_int64()
        self._check(_lib().vh_count_above(self.h, eF, ctypes.byref(c)), "vh_count_above")
        return c.value

    def fo_flag_ids(self, eF, cap=None, with_corners=False):
        nx, ny, nz = self.cfg.nx, self.cfg.ny, self.cfg.nz
        m = (nx - 1) * (ny - 1) * (nz - 1 if nz > 1 else 1)
        cap = m if cap is None else cap
        ids = np.empty(max(cap, 1), dtype=np.int64)
        per = 16 if nz > 1 else 8
        corners = np.empty(max(cap, 1) * per, dtype=np.float64) if with_corners else None
        n = ctypes.c_int64()
        self._check(_lib().vh_fo_flag(self.h, eF, _ptr(ids), _ptr(corners), cap, ctypes.byref(n)), "vh_fo_flag")
        self._fo_n = n.value
        ids = ids[:n.value].copy()
        if with_corners:
            return ids, corners[:n.value * per].reshape(n.value, per).copy()
        return ids

    def fo_interp(self, cubes, frac):
        """The ten freezeout variables at points inside cubes (vh_fo_interp):
        cubes (n,) int64 cube ids, frac (n, 4) offsets in units of the
        coarse time step and the cell sizes; returns (n, 10)."""
        cubes = np.ascontiguousarray(cubes, dtype=np.int64)
        frac = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 4)
        if frac.shape[0] != cubes.shape[0]:
            raise ValueError("fo_interp: one row of frac per cube id")
        out = np.empty((max(cubes.shape[0], 1), 10), dtype=np.float64)
        self._check(_lib().vh_fo_interp(self.h, _ptr(cubes), _ptr(frac), cubes.shape[0], _ptr(out)), "vh_fo_interp")
        return out[:cubes.shape[0]].copy()

    def thermal_vorticity(self):
        """calculateThermalVorticity on the current state (vh_thermal_vorticity):
        (10, N) wtx wty wtn wxy wxn wyn beta_t beta_x beta_y beta_n."""
        self._check(_lib().vh_thermal_vorticity(self.h), "vh_thermal_vorticity")
        return np.stack([self.get(f) for f in ("wtx", "wty", "wtn", "wxy", "wxn", "wyn", "beta_t", "beta_x",
                                               "beta_y", "beta_n")])

    def fo_surface(self, widths):
        """The surface elements of the cubes the last fo_flag_ids call flagged,
        built on the device (vh_fo_surface): returns (counts per cube,
        elements (n, 2*dim): centroid relative to the cube then normal,
        vars (n, 10) at the centroids)."""
        nz = self.cfg.nz
        dim = 4 if nz > 1 else 3
        w = np.ascontiguousarray(widths, dtype=np.float64)
        if w.size < dim:
            raise ValueError(f"fo_surface: {dim} cell widths")
        m = (self.cfg.nx - 1) * (self.cfg.ny - 1) * (nz - 1 if nz > 1 else 1)
        counts = np.zeros(max(m, 1), dtype=np.int32)
        total = ctypes.c_int64()
        st = _lib().vh_fo_surface(self.h, _ptr(w), _ptr(counts), None, None, 0, ctypes.byref(total))
        if st != VH_ERR_CAPACITY:
            self._check(st, "vh_fo_surface")
        cap = max(total.value, 1)
        el = np.empty((cap, 2 * dim), dtype=np.float64)
        var = np.empty((cap, 10), dtype=np.float64)
        self._check(_lib().vh_fo_surface(self.h, _ptr(w), _ptr(counts), _ptr(el), _ptr(var), cap,
                                         ctypes.byref(total)), "vh_fo_surface")
        n = total.value
        return counts[:getattr(self, "_fo_n", 0)].copy(), el[:n].copy(), var[:n].copy()

    def eos_energy_density(self, T):
        return _lib().vh_eos_energy_density(self.h, T)

    def eos_pressure(self, e):
        return _lib().vh_eos_pressure(self.h, e)

    def eos_temperature(self, e):
        return _lib().vh_eos_temperature(self.h, e)

    def last_kernel_ms(self):
        return _lib().vh_last_kernel_ms(self.h)

    def inferred(self, t, q, ePrev):
        """getInferredVariables on the device for n states: q is (nc, n),
        ePrev (n,).  Returns (out (6, n): e, p, u^tau, u^x, u^y, u^eta;
        NaN site per state, 0 if none)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        ePrev = np.ascontiguousarray(ePrev, dtype=np.float64).ravel()
        n = ePrev.size
        if q.shape != (self.nc, n):
            raise ValueError(f"q must be ({self.nc}, {n})")
        out = np.empty((6, n), dtype=np.float64)
        nan = np.zeros(n, dtype=np.int32)
        st = _lib().vh_inferred_variables(self.h, t, _ptr(q), _ptr(ePrev), n, _ptr(out), _ptr(nan))
        if st not in (VH_OK, VH_ERR_NAN):
            self._check(st, "vh_inferred_variables")
        return out, nan


class Batch:
    """Many events resident on one GPU (a vh_batch), stepped together."""

    def __init__(self, capacity_cells, max_events, max_steps, bulk=False, eos_table=None, device=0,
                 precision=PREC_FP64, cs2_mode=0, fo_vars=False):
        L = _lib()
        self.nc = 15 if bulk else 14
        self.precision = precision
        self.max_events = max_events
        self._table = None if eos_table is None else np.ascontiguousarray(eos_table, dtype=np.float64)
        cfg = BatchConfig(device, capacity_cells, max_events, max_steps, int(bool(bulk)),
                          EOS_CONFORMAL if self._table is None else EOS_TABLE,
                          None if self._table is None else self._table.ctypes.data, precision, cs2_mode,
                          int(bool(fo_vars)))
        b = ctypes.c_void_p()
        Hydro._check(L.vh_batch_create(ctypes.byref(cfg), ctypes.byref(b)), "vh_batch_create")
        self.b = b
        self.events = {}

    def close(self):
        if getattr(self, "b", None):
            _lib().vh_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, nx, ny, nz, dx, dy, dz, dt, t0, etabar, ic, T_reg_GeV=0.0, etas_Tk_GeV=0.154,
            etas_slope_lo=0.0, etas_slope_hi=0.0, zeta_Tpeak_GeV=0.0, zeta_peak=0.0):
        """Admits one event; returns its id, or None when the batch is full
        (VH_ERR_CAPACITY).  ic: dict of x-fastest arrays e, p, ut, ux, uy, un, pi (10*N), Pi (optional)."""
        N = nx * ny * nz
        arrs = _ic_arrays([ic.get(k) for k in ("e", "p", "ut", "ux", "uy", "un", "pi", "Pi")], N)
        prm = EventParams(nx, ny, nz, dx, dy, dz, dt, t0, etabar, T_reg_GeV, etas_Tk_GeV, etas_slope_lo,
                          etas_slope_hi, zeta_Tpeak_GeV, zeta_peak)
        eid = ctypes.c_int(-1)
        st = _lib().vh_batch_add(self.b, ctypes.byref(prm), *[_ptr(a) for a in arrs], ctypes.byref(eid))
        if st == VH_ERR_CAPACITY:
            return None
        Hydro._check(st, "vh_batch_add")
        self.events[eid.value] = (nx, ny, nz)
        return eid.value

    def remove(self, eid):
        Hydro._check(_lib().vh_batch_remove(self.b, eid), "vh_batch_remove")
        self.events.pop(eid, None)

    def steps(self, count):
        Hydro._check(_lib().vh_batch_steps(self.b, count), "vh_batch_steps")

    def sync(self):
        Hydro._check(_lib().vh_batch_sync(self.b), "vh_batch_sync")

    def nan_info(self):
        site, ev, cell = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        Hydro._check(_lib().vh_batch_nan_info(self.b, ctypes.byref(site), ctypes.byref(ev), ctypes.byref(cell)),
                     "vh_batch_nan_info")
        return site.value, ev.value, cell.value

    def clear_error(self):
        Hydro._check(_lib().vh_batch_clear_error(self.b), "vh_batch_clear_error")

    def counts_above(self, eF):
        out = np.empty(self.max_events, dtype=np.int64)
        Hydro._check(_lib().vh_batch_counts_above(self.b, eF, _ptr(out)), "vh_batch_counts_above")
        return out

    def run(self, eF, coarse, nchecks):
        """vh_batch_run: nchecks x (coarse steps + the stop-rule check) on the
        device.  Returns (status, stopped, steps_done) per id; status is
        VH_OK or VH_ERR_NAN (other errors raise)."""
        stopped = np.empty(self.max_events, dtype=np.int32)
        steps = np.empty(self.max_events, dtype=np.int32)
        st = _lib().vh_batch_run(self.b, eF, coarse, nchecks, _ptr(stopped), _ptr(steps))
        if st not in (VH_OK, VH_ERR_NAN):
            raise VhError(st, "vh_batch_run")
        return st, stopped, steps

    def get(self, eid, field):
        f = FIELD[field] if isinstance(field, str) else field
        nx, ny, nz = self.events[eid]
        out = np.empty(nx * ny * nz, dtype=np.float64)
        Hydro._check(_lib().vh_batch_download(self.b, eid, f, _ptr(out)), "vh_batch_download")
        return out

    def fo_flag_ids(self, eid, eF, with_corners=False):
        nx, ny, nz = self.events[eid]
        cap = max(1, (nx - 1) * (ny - 1) * (nz - 1 if nz > 1 else 1))
        ids = np.empty(cap, dtype=np.int64)
        per = 16 if nz > 1 else 8
        corners = np.empty(cap * per, dtype=np.float64) if with_corners else None
        n = ctypes.c_int64()
        Hydro._check(_lib().vh_batch_fo_flag(self.b, eid, eF, _ptr(ids), _ptr(corners), cap, ctypes.byref(n)),
                     "vh_batch_fo_flag")
        self._fo_n = getattr(self, "_fo_n", {})
        self._fo_n[eid] = n.value
        if with_corners:
            return ids[:n.value].copy(), corners[:n.value * per].reshape(n.value, per).copy()
        return ids[:n.value].copy()

    def fo_flags(self, eids, eF):
        """The cut cubes of several events in one pass (vh_batch_fo_flags):
        each event's ids, in increasing order."""
        n = len(eids)
        ids = np.asarray(eids, dtype=np.int32)
        bufs = []
        for e in eids:
            nx, ny, nz = self.events[e]
            bufs.append(np.empty(max(1, (nx - 1) * (ny - 1) * (nz - 1 if nz > 1 else 1)), dtype=np.int64))
        cap = np.array([x.size for x in bufs], dtype=np.int64)
        nf = np.zeros(n, dtype=np.int64)
        arr = (ctypes.c_void_p * max(n, 1))(*[x.ctypes.data for x in bufs])
        Hydro._check(_lib().vh_batch_fo_flags(self.b, n, _ptr(ids), eF, arr, _ptr(cap), _ptr(nf)), "vh_batch_fo_flags")
        self._fo_n = getattr(self, "_fo_n", {})
        for e, k in zip(eids, nf):
            self._fo_n[e] = int(k)
        return [x[:k].copy() for x, k in zip(bufs, nf)]

    def fo_surface(self, eid, widths):
        """As Hydro.fo_surface for one event (batch created with fo_vars=True):
        (counts per flagged cube, elements (n, 2*dim), vars (n, 10))."""
        nx, ny, nz = self.events[eid]
        dim = 4 if nz > 1 else 3
        w = np.ascontiguousarray(widths, dtype=np.float64)
        nf = getattr(self, "_fo_n", {}).get(eid, 0)
        counts = np.zeros(max(nf, 1), dtype=np.int32)
        total = ctypes.c_int64()
        st = _lib().vh_batch_fo_surface(self.b, eid, _ptr(w), _ptr(counts), None, None, 0, ctypes.byref(total))
        if st != VH_ERR_CAPACITY:
            Hydro._check(st, "vh_batch_fo_surface")
        cap = max(total.value, 1)
        el = np.empty((cap, 2 * dim), dtype=np.float64)
        var = np.empty((cap, 10), dtype=np.float64)
        Hydro._check(_lib().vh_batch_fo_surface(self.b, eid, _ptr(w), _ptr(counts), _ptr(el), _ptr(var), cap,
                                                ctypes.byref(total)), "vh_batch_fo_surface")
        n = total.value
        return counts[:nf].copy(), el[:n].copy(), var[:n].copy()

    def fo_surfaces(self, eids, widths, caps=None):
        """Several events' surfaces in one pass (vh_batch_fo_surfaces), each on
        the cubes its last fo_flag_ids flagged: widths[k] the k-th event's cell
        widths; caps[k] its room in elements (default: a first call with no
        room gives every count).  Returns (status, totals, results) with
        results[k] = (counts per flagged cube, elements (n, 2*dim), vars
        (n, 10)), or None for an event whose elements exceed its room."""
        n = len(eids)
        ids = np.asarray(eids, dtype=np.int32)
        w = np.zeros(4 * n, dtype=np.float64)
        for k, wk in enumerate(widths):
            w[4 * k:4 * k + len(wk)] = wk
        nfs = [getattr(self, "_fo_n", {}).get(e, 0) for e in eids]
        counts = [np.zeros(max(nf, 1), dtype=np.int32) for nf in nfs]
        tot = np.zeros(n, dtype=np.int64)

        def call(cap):
            el = [np.empty(max(int(c), 1) * 8, dtype=np.float64) for c in cap]
            var = [np.empty(max(int(c), 1) * 10, dtype=np.float64) for c in cap]
            arr = lambda xs: (ctypes.c_void_p * max(n, 1))(*[x.ctypes.data for x in xs])
            st = _lib().vh_batch_fo_surfaces(self.b, n, _ptr(ids), _ptr(w), arr(counts), arr(el), arr(var),
                                             _ptr(cap), _ptr(tot))
            if st not in (VH_OK, VH_ERR_CAPACITY):
                Hydro._check(st, "vh_batch_fo_surfaces")
            return st, el, var

        if caps is None:
            call(np.zeros(n, dtype=np.int64))
            caps = tot.copy()
        cap = np.asarray(caps, dtype=np.int64)
        st, el, var = call(cap)
        out = []
        for k, e in enumerate(eids):
            t = int(tot[k])
            if t > cap[k]:
                out.append(None)
                continue
            dim = 4 if self.events[e][2] > 1 else 3
            out.append((counts[k][:nfs[k]].copy(), el[k][:t * 2 * dim].reshape(t, 2 * dim).copy(),
                        var[k][:t * 10].reshape(t, 10).copy()))
        return st, tot.copy(), out

    def info(self, eid):
        a, s, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        Hydro._check(_lib().vh_batch_event_info(self.b, eid, ctypes.byref(a), ctypes.byref(s), ctypes.byref(c)),
                     "vh_batch_event_info")
        return {"active": bool(a.value), "steps_done": s.value, "cells": c.value}

    def free_cells(self):
        f, big = ctypes.c_int64(), ctypes.c_int64()
        Hydro._check(_lib().vh_batch_free_cells(self.b, ctypes.byref(f), ctypes.byref(big)), "vh_batch_free_cells")
        return f.value, big.value

    def last_ms(self):
        return _lib().vh_batch_last_ms(self.b)

    def autotune(self, reps=3):
        t2, t3 = ctypes.c_int(), ctypes.c_int()
        Hydro._check(_lib().vh_batch_autotune(self.b, reps, ctypes.byref(t2), ctypes.byref(t3)), "vh_batch_autotune")
        return t2.value, t3.value

    def set_face_mode(self, mode):
        Hydro._check(_lib().vh_batch_set_face_mode(self.b, mode), "vh_batch_set_face_mode")

    def set_tiles(self, tile2d, tile3d):
        Hydro._check(_lib().vh_batch_set_tiles(self.b, tile2d, tile3d), "vh_batch_set_tiles")

    def eos_energy_density(self, T):
        return _lib().vh_batch_eos_energy_density(self.b, T)


SERVE_DEFAULTS = dict(rate_per_s=10.0, n_requests=50, seed=1, frac_3d=0.2, side2_min=101, side2_max=201,
                      side3_min=101, side3_max=161, nz3_min=41, nz3_max=81, dx=0.1, dt=0.02, t0=0.6, etabar=0.08,
                      etas_Tk_GeV=0.154, etas_slope_lo=0.5, etas_slope_hi=1.0, T_reg_GeV=0.1, T_F_GeV=0.150,
                      coarse=5, e0=10.0, e_floor=0.01, hbm_budget_bytes=64 << 30, precision=PREC_MIXED, bulk=1,
                      max_steps=3000, max_resident=256, collect_surface=1, device=0, digest=0)


def serve(eos_table=None, on_surface=None, **kw):
    """Runs the serving loop (vh_serve_run); returns (stats dict, list of
    request dicts).  on_surface(request, rows): with collect_surface=2, each
    resident request's surface elements at each check, rows an (n, 20) array
    in surface.dat's columns."""
    cfg = dict(SERVE_DEFAULTS)
    unknown = set(kw) - set(cfg)
    if unknown:
        raise TypeError(f"unknown serving options {sorted(unknown)}")
    cfg.update(kw)
    c = ServeConfig(**cfg)
    if on_surface is not None:
        def _cb(user, request, n, rows):
            on_surface(request, np.ctypeslib.as_array(rows, shape=(n, 20)).copy())
        c.on_surface = SURFACE_CALLBACK(_cb)
    st = ServeStats()
    reqs = (ServeRequest * cfg["n_requests"])()
    table = None if eos_table is None else np.ascontiguousarray(eos_table, dtype=np.float64)
    r = _lib().vh_serve_run(ctypes.byref(c), None if table is None else table.ctypes.data, ctypes.byref(st), reqs)
    if r != VH_OK:
        raise VhError(r, "vh_serve_run")
    stats = {f: getattr(st, f) for f, _ in ServeStats._fields_}
    out = [{f: getattr(q, f) for f, _ in ServeRequest._fields_} for q in reqs]
    return stats, out


def build_info():
    return _lib().vh_build_info().decode()


def state_digest(fields):
    """The serving loop's digest (vh_serve_request.digest) of a state given as
    the arrays of fields 0..nc-1, 15..20 in that order."""
    M = (1 << 64) - 1
    h = 14695981039346656037
    for a in fields:
        bits = np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)
        w = 2 * np.arange(bits.size, dtype=np.uint64) + np.uint64(1)
        d = int(np.sum(bits * w, dtype=np.uint64))
        h = (h * 1099511628211 + d) & M
    return h


def mc_glauber(nx, ny, nz, dx, dy, dz, seed, A=197, R=6.38, a=0.535, sigma_nn_fm2=4.2, b=-1.0, b_max=10.0,
               width=0.5, eta_flat=4.0, eta_variance=1.96, e0=10.0, e_floor=1e-3):
    """Seeded MC-Glauber energy density (vh_ic_mc_glauber); returns (e, info)."""
    gp = GlauberParams(A, R, a, sigma_nn_fm2, b, b_max, width, eta_flat, eta_variance, e0, e_floor, seed)
    info = GlauberInfo()
    e = np.empty(nx * ny * nz, dtype=np.float64)
    st = _lib().vh_ic_mc_glauber(ctypes.byref(gp), nx, ny, nz, dx, dy, dz, _ptr(e), ctypes.byref(info))
    if st != VH_OK:
        raise VhError(st, "vh_ic_mc_glauber")
    return e, {"b": info.b, "n_part": info.n_part, "n_coll": info.n_coll}


def mc_glauber_classic(nx, ny, nz, dx, dy, dz, A=197, b=0.0, sigma_nn_mb=42.0, e0=1.0, eta_flat=4.0,
                       eta_variance=1.96):
    """The reference's own IC type 4 (vh_ic_mc_glauber_classic); returns (e, wounded nucleons)."""
    e = np.empty(nx * ny * nz, dtype=np.float64)
    nw = ctypes.c_int(0)
    st = _lib().vh_ic_mc_glauber_classic(A, b, sigma_nn_mb, e0, eta_flat, eta_variance, nx, ny, nz, dx, dy, dz,
                                         _ptr(e), ctypes.byref(nw))
    if st != VH_OK:
        raise VhError(st, "vh_ic_mc_glauber_classic")
    return e, nw.value


def surface_elements(corners, E0, d):
    """Freezeout surface elements of one square (4 corner values), cube (8) or
    hypercube (16): an array of rows (centroid..., normal...)."""
    corners = np.ascontiguousarray(corners, dtype=np.float64).ravel()
    d = np.ascontiguousarray(d, dtype=np.float64)
    dim = {4: 2, 8: 3, 16: 4}[corners.size]
    out = np.empty(16 * 2 * dim)
    n = ctypes.c_int(0)
    name = {2: "vh_surface_square", 3: "vh_surface_cube", 4: "vh_surface_hypercube"}[dim]
    st = getattr(_lib(), name)(_ptr(corners), E0, _ptr(d), _ptr(out), 16, ctypes.byref(n))
    if st != VH_OK:
        raise VhError(st, name)
    return out[:2 * dim * n.value].reshape(n.value, 2 * dim)


def optical_glauber(nx, ny, nz, dx, dy, dz, A=197, b=0.0, sigma_nn_mb=42.0, alpha=1.0, e0=1.0, eta_flat=4.0,
                    eta_variance=1.96):
    """The reference's IC type 2 (vh_ic_optical_glauber)."""
    e = np.empty(nx * ny * nz, dtype=np.float64)
    st = _lib().vh_ic_optical_glauber(A, b, sigma_nn_mb, alpha, e0, eta_flat, eta_variance, nx, ny, nz, dx, dy, dz,
                                      _ptr(e))
    if st != VH_OK:
        raise VhError(st, "vh_ic_optical_glauber")
    return e


def analytic_ic(type_, nx, ny, nz, dx, dy, e0=1.0):
    """The reference's analytic initial states, types 3 and 5-9 (vh_ic_test_problem): dict of x-fastest arrays."""
    N = nx * ny * nz
    out = {k: np.empty(N) for k in ("e", "p", "ut", "ux", "uy", "un")}
    st = _lib().vh_ic_test_problem(type_, e0, nx, ny, nz, dx, dy,
                                   *[_ptr(out[k]) for k in ("e", "p", "ut", "ux", "uy", "un")])
    if st != VH_OK:
        raise VhError(st, "vh_ic_test_problem")
    return out


def read_fields(root, nx, ny, nz, t0):
    """IC type 10 reader (vh_ic_read_fields): root/input/{e,ux,uy,un,pitt..pinn}.dat."""
    N = nx * ny * nz
    out = {k: np.empty(N) for k in ("e", "ut", "ux", "uy", "un")}
    out["pi"] = np.empty(10 * N)
    st = _lib().vh_ic_read_fields(root.encode(), nx, ny, nz, t0,
                                  *[_ptr(out[k]) for k in ("e", "ut", "ux", "uy", "un", "pi")])
    if st != VH_OK:
        raise VhError(st, "vh_ic_read_fields")
    return out


def read_gubser(path, nx, ny, nz):
    """IC type 1 reader (vh_ic_read_gubser)."""
    N = nx * ny * nz
    out = {k: np.empty(N) for k in ("e", "p", "ut", "ux", "uy", "un")}
    out["pi"] = np.empty(10 * N)
    st = _lib().vh_ic_read_gubser(path.encode(), nx, ny, nz,
                                  *[_ptr(out[k]) for k in ("e", "p", "ut", "ux", "uy", "un", "pi")])
    if st != VH_OK:
        raise VhError(st, "vh_ic_read_gubser")
    return out


def read_tmunu(path, nx, ny, nz):
    """IC type 11 reader (vh_ic_read_tmunu): dict of x-fastest arrays."""
    N = nx * ny * nz
    out = {k: np.empty(N) for k in ("e", "p", "ut", "ux", "uy", "un", "Pi")}
    out["pi"] = np.empty(10 * N)
    st = _lib().vh_ic_read_tmunu(path.encode(), nx, ny, nz, *[_ptr(out[k]) for k in
                                                              ("e", "p", "ut", "ux", "uy", "un", "pi", "Pi")])
    if st != VH_OK:
        raise VhError(st, "vh_ic_read_tmunu")
    return out


def write_tmunu(path, nx, ny, nz, dx, dy, dz, ic):
    arrs = [np.ascontiguousarray(ic[k], dtype=np.float64) for k in ("e", "p", "ut", "ux", "uy", "un", "pi")]
    Pi = ic.get("Pi")
    Pi = None if Pi is None else np.ascontiguousarray(Pi, dtype=np.float64)
    st = _lib().vh_ic_write_tmunu(path.encode(), nx, ny, nz, dx, dy, dz, *[_ptr(a) for a in arrs], _ptr(Pi))
    if st != VH_OK:
        raise VhError(st, "vh_ic_write_tmunu")
