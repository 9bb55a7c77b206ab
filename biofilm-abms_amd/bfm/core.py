"""ctypes binding of include/bfm.h and the reference-named host API."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)

# Column indices, ABM-3d/ABM-3d-cpp/include/indices.hpp:17-47 (+ gamma, simulation.hpp:1182).
COL = dict(id=0, rx=1, ry=2, rz=3, nx=4, ny=5, nz=6, drx=7, dry=8, drz=9, dnx=10, dny=11,
           dnz=12, l=13, half_l=14, t0=15, growth=16, eta0=17, eta1=18, sigma0=19, group=20,
           gamma=21)

BFM_OK, BFM_ERR_ARG, BFM_ERR_CAPACITY, BFM_ERR_NAN, BFM_ERR_JKR_ROOT, BFM_ERR_HIP, \
    BFM_ERR_UNSUPPORTED = range(7)
_STATUS = {1: "argument", 2: "capacity", 3: "nan", 4: "jkr_root", 5: "hip", 6: "unsupported"}

ADHESION_NONE, ADHESION_JKR_ISOTROPIC, ADHESION_JKR_ANISOTROPIC = 0, 1, 2
JKR_EXACT, JKR_TABLE_1D, JKR_TABLE_2D, JKR_SURROGATE = 0, 1, 2, 3


class BFMError(RuntimeError):
    """Error status of a C-ABI call; .status is the bfm_status code."""

    def __init__(self, status, msg):
        super().__init__(f"bfm error {status} ({_STATUS.get(status, '?')}): {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [("R", C.c_double), ("Rcell", C.c_double), ("E0", C.c_double),
                ("repulsion_prefactors", C.c_double * 3), ("nz_threshold", C.c_double),
                ("cell_cell_coulomb_coeff", C.c_double),
                ("cell_surface_coulomb_coeff", C.c_double), ("no_surface", C.c_int),
                ("adhesion_mode", C.c_int), ("jkr_mode", C.c_int), ("colidx_gamma", C.c_int),
                ("friction_mode", C.c_int), ("ncols", C.c_int),
                ("colidx_eta_cell_cell", C.c_int)]


def repulsion_prefactors(R, Rcell, E0, Ecell):
    """simulation.hpp:1271-1274."""
    return ((4. / 3.) * E0 * np.sqrt(R / 2),
            (4. / 3.) * Ecell * np.sqrt(Rcell / 2),
            (4. / 3.) * E0 * np.sqrt(R / 2) * pow(2 * (R - Rcell), 1.5))


def make_params(R=0.8, Rcell=0.5, E0=3900.0, Ecell=3.9e6, nz_threshold=1e-8,
                cell_surface_coulomb_coeff=1.0, no_surface=False, adhesion=True,
                jkr_mode=JKR_EXACT, ncols=None, friction=False, cell_cell_coulomb_coeff=1.0,
                colidx_eta_cell_cell=None, anisotropic=False):
    """bfm_params. ncols defaults to 22 (the gamma column 21 is present) plus
    one friction-coefficient column when friction is on, which is then the
    last column (22 with adhesion, as simulation.hpp:1176-1188 places it).
    anisotropic: adhesion mode JKR_ANISOTROPIC (the table path; set the tables
    with Context.set_aniso_jkr_tables)."""
    if ncols is None:
        ncols = 22 + int(bool(friction))
    if colidx_eta_cell_cell is None:
        colidx_eta_cell_cell = ncols - 1 if friction else -1
    p = Params()
    p.R, p.Rcell, p.E0 = R, Rcell, E0
    for k, v in enumerate(repulsion_prefactors(R, Rcell, E0, Ecell)):
        p.repulsion_prefactors[k] = v
    p.nz_threshold = nz_threshold
    p.cell_cell_coulomb_coeff = cell_cell_coulomb_coeff
    p.cell_surface_coulomb_coeff = cell_surface_coulomb_coeff
    p.no_surface = int(bool(no_surface))
    p.adhesion_mode = (ADHESION_JKR_ANISOTROPIC if anisotropic else ADHESION_JKR_ISOTROPIC) \
        if adhesion else ADHESION_NONE
    p.jkr_mode = jkr_mode
    p.colidx_gamma = COL["gamma"] if adhesion else -1
    p.friction_mode = 1 if friction else 0
    p.ncols = ncols
    p.colidx_eta_cell_cell = colidx_eta_cell_cell
    return p


_dpt = C.POINTER(C.c_double)


class AnisoTablesC(C.Structure):
    """bfm_aniso_jkr_tables."""
    _fields_ = [("theta", _dpt), ("n_theta", C.c_int), ("half_l", _dpt), ("n_half_l", C.c_int),
                ("coords", _dpt), ("n_coords", C.c_int), ("curvature_rx", _dpt),
                ("rx", _dpt), ("n_rx", C.c_int), ("phi", _dpt), ("n_phi", C.c_int),
                ("delta", _dpt), ("n_delta", C.c_int), ("gamma", _dpt), ("n_gamma", C.c_int),
                ("force_radius", _dpt)]


class AnisoJKRTables:
    """The tables of AnisotropicJKRDataFixedGamma / VariableGamma
    (mechanics.hpp:199-327): grids theta, half_l, coords (centerline
    coordinate / half_l), rx, phi, delta and gamma (None: fixed gamma);
    curvature_rx[i_theta, i_half_l, i_coord] = Rx at the contact point;
    force_radius[i_rx1, i_rx2, i_phi, i_delta(, i_gamma), 0/1] = (force,
    contact radius). Combinations absent from the precomputed files are 0."""

    def __init__(self, theta, half_l, coords, curvature_rx, rx, phi, delta, force_radius,
                 gamma=None, fixed_gamma=None):
        a = lambda x: np.ascontiguousarray(x, dtype=np.float64)   # noqa: E731
        self.theta, self.half_l, self.coords = a(theta), a(half_l), a(coords)
        self.rx, self.phi, self.delta = a(rx), a(phi), a(delta)
        self.gamma = None if gamma is None else a(gamma)
        self.curvature_rx = a(curvature_rx).reshape(len(self.theta), len(self.half_l),
                                                    len(self.coords))
        shape = (len(self.rx), len(self.rx), len(self.phi), len(self.delta)) + \
            ((len(self.gamma),) if self.gamma is not None else ()) + (2,)
        self.force_radius = a(force_radius).reshape(shape)
        self.fixed_gamma = fixed_gamma

    def cstruct(self):
        t = AnisoTablesC()
        for name in ("theta", "half_l", "coords", "rx", "phi", "delta"):
            arr = getattr(self, name)
            setattr(t, name, _ptr(arr))
            setattr(t, "n_" + name, len(arr))
        t.gamma = _ptr(self.gamma) if self.gamma is not None else None
        t.n_gamma = 0 if self.gamma is None else len(self.gamma)
        t.curvature_rx = _ptr(self.curvature_rx)
        t.force_radius = _ptr(self.force_radius)
        return t


# Butcher tableaux, simulation.hpp:1207-1268 (name -> (A, b, bs, error_order)).
TABLEAUX = {
    "heun_euler": ([[0, 0], [1, 0]], [0.5, 0.5], [1, 0], 1),
    "bogacki_shampine": ([[0, 0, 0, 0], [1. / 2., 0, 0, 0], [0, 3. / 4., 0, 0],
                          [2. / 9., 1. / 3., 4. / 9., 0]],
                         [2. / 9., 1. / 3., 4. / 9., 0], [7. / 24., 1. / 4., 1. / 3., 1. / 8.], 2),
    "rkf45": ([[0, 0, 0, 0, 0, 0], [1. / 4., 0, 0, 0, 0, 0], [3. / 32., 9. / 32., 0, 0, 0, 0],
               [1932. / 2197., -7200. / 2197., 7296. / 2197., 0, 0, 0],
               [439. / 216., -8.0, 3680. / 513., -845. / 4104., 0, 0],
               [-8. / 27., 2.0, -3544. / 2565., 1859. / 4104., -11. / 40., 0]],
              [16. / 135., 0, 6656. / 12825., 28561. / 56430., -9. / 50., 2. / 55.],
              [25. / 216., 0, 1408. / 2565., 2197. / 4104., -1. / 5., 0], 4),
    "dormand_prince": ([[0, 0, 0, 0, 0, 0, 0], [1. / 5., 0, 0, 0, 0, 0, 0],
                        [3. / 40., 9. / 40., 0, 0, 0, 0, 0],
                        [44. / 45., -56. / 15., 32. / 9., 0, 0, 0, 0],
                        [19372. / 6561., -25360. / 2187., 64448. / 6561., -212. / 729, 0, 0, 0],
                        [9017. / 3168., -355. / 33., 46732. / 5247., 49. / 176., -5103. / 18656., 0, 0],
                        [35. / 384., 0, 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0]],
                       [35. / 384., 0, 500. / 1113., 125. / 192., -2187. / 6784., 11. / 84., 0],
                       [5179. / 57600., 0, 7571. / 16695., 393. / 640., -92097. / 339200.,
                        187. / 2100., 1. / 40.], 4),
}


def tableau(name):
    A, b, bs, order = TABLEAUX[name]
    return (np.ascontiguousarray(A, dtype=np.float64), np.asarray(b, dtype=np.float64),
            np.asarray(bs, dtype=np.float64), order)


def library_path():
    return os.path.join(_PKG, "libbfm.so")


_LIB = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)


def load_library():
    """Loads the in-tree libbfm.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    # torch ships its own HIP runtime under the same soname as /opt/rocm's; the
    # first one loaded serves the whole process. Load torch's first so torch and
    # libbfm share one runtime (device pointers and streams pass between them).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "bfm_create": [C.POINTER(vp), C.c_int, C.c_int, C.c_int64],
        "bfm_destroy": [vp],
        "bfm_set_params": [vp, C.POINTER(Params)],
        "bfm_set_stream": [vp, vp],
        "bfm_sync": [vp],
        "bfm_set_jkr_table": [vp, _dp, C.c_int, _dp, C.c_int, _dp],
        "bfm_set_surrogate": [vp, vp, vp, vp, vp, vp, vp, vp, vp, _dp],
        "bfm_set_aniso_jkr_tables": [vp, C.POINTER(AnisoTablesC)],
        "bfm_read_aniso_jkr_tables": [C.c_char_p, C.c_char_p, C.c_int,
                                      C.POINTER(C.POINTER(AnisoTablesC)), _dp],
        "bfm_get_cell_neighbors": [vp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                   _dp, C.c_int64, _i64p],
        "bfm_update_neighbor_distances": [vp, _dp, C.c_int, C.c_int, _dp, C.c_int64, C.c_int64],
        "bfm_get_conservative_forces": [vp, _dp, C.c_int, C.c_int, _dp, C.c_int64, C.c_int64, _ip,
                                        _dp],
        "bfm_get_velocities": [vp, _dp, C.c_int, C.c_int, _dp, C.c_int64, C.c_int64, _ip, _dp,
                               _dp],
        "bfm_step_rk": [vp, _dp, _dp, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_int, _dp,
                        C.c_int64, C.c_int64, _dp, _dp, _dp],
        "bfm_dev_build_neighbors": [vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    _i64p],
        "bfm_dev_export_neighbors": [vp, vp, C.c_int64],
        "bfm_dev_import_neighbors": [vp, vp, C.c_int64, C.c_int64, C.c_int],
        "bfm_dev_get_velocities": [vp, vp, C.c_int, C.c_int, vp, vp, vp],
        "bfm_dev_stage": [vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, vp],
        "bfm_dev_get_velocities_noalloc": [C.POINTER(Params), vp, C.c_int, C.c_int, vp, C.c_int64,
                                           C.c_int64, vp, vp, vp, vp],
        "bfm_dev_step_rk": [vp, _dp, _dp, _dp, C.c_int, C.c_double, vp, C.c_int, C.c_int, vp, vp,
                            vp, C.c_double, vp],
        "bfm_enable_timing": [vp, C.c_int],
        "bfm_kernel_times": [vp, C.c_char_p, C.c_int, _dp, _i64p, C.c_int, _ip],
        "bfm_reset_timing": [vp],
        "bfm_stats": [vp, _i64p, _i64p, _i64p, _dp, _ip],
        "bfm_set_tuning": [vp, C.c_double, C.c_int, C.c_int],
        "bfm_dev_update_neighbor_distances": [vp, vp, C.c_int, C.c_int],
        "bfm_simulate": [vp, vp, _dp, C.c_int, C.c_int, _ip, C.c_int, _dp, C.c_int, _ip, _ip, _ip,
                         _dp, C.c_int, vp],
        "bfm_divide_cells": [vp, _dp, _ip, C.c_int, C.c_int, C.c_double, C.c_uint64, _ip, _ip,
                             C.c_int, _ip],
        "bfm_rng_uniform": [C.c_uint64, C.c_int, _dp],
        "bfm_rng_draws_device": [C.c_int, C.c_uint64, C.c_int, _dp],
        "bfm_jkr_equilibrium_distance": [C.c_double] * 7 + [C.c_int, _dp, _dp],
        "bfm_jkr_optimal_surface_energy_density": [C.c_double] * 9 + [C.c_int, C.c_int, _dp, _dp],
        "bfm_ensemble_simulate": [vp, vp, C.c_int, _ip, _dp, C.c_int, vp, C.c_int, _dp, C.c_int, _ip,
                                  _dp, vp, _ip, vp],
        "bfm_autotune_neighbors": [vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                   _ip, _dp],
        "bfm_step_verlet": [vp, C.c_double, C.c_double, _dp, C.c_int, C.c_int, _dp, C.c_int64, C.c_int64,
                            _dp, _dp],
        "bfm_dev_step_verlet": [vp, C.c_double, C.c_double, vp, C.c_int, C.c_int, vp, vp, C.c_int],
        "bfm_write_cells": [C.c_char_p, _dp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_char_p),
                            C.POINTER(C.c_char_p), C.c_int, _ip, _ip, C.c_int],
        "bfm_read_cells": [C.c_char_p, _dp, C.c_int, C.c_int, C.c_int, _ip, _ip, C.c_char_p, C.c_int,
                           _ip],
        "bfm_write_lineage": [C.c_char_p, _ip, C.c_int],
        "bfm_read_lineage": [C.c_char_p, _ip, C.c_int, _ip],
        "bfm_load_config": [C.c_char_p, vp],
        "bfm_run": [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.bfm_last_error.argtypes = [vp]
    lib.bfm_last_error.restype = C.c_char_p
    lib.bfm_io_last_error.argtypes = []
    lib.bfm_io_last_error.restype = C.c_char_p
    lib.bfm_free_aniso_jkr_tables.argtypes = [C.POINTER(AnisoTablesC)]
    lib.bfm_free_aniso_jkr_tables.restype = None
    _LIB = lib
    return lib


def dev_get_velocities_noalloc(params, d_cells, n, ld, d_neighbors, m, ld_nb, d_vel, d_a2d=0,
                               d_noise=0, stream=0):
    """bfm_dev_get_velocities_noalloc: getVelocities on device pointers with no
    device allocation and no context (n <= 65536, m <= 262144). Raises
    BFMError with the returned status."""
    lib = load_library()
    rc = lib.bfm_dev_get_velocities_noalloc(C.byref(params), C.c_void_p(d_cells), int(n), int(ld),
                                            C.c_void_p(d_neighbors or None), int(m), int(ld_nb),
                                            C.c_void_p(d_a2d or None), C.c_void_p(d_noise or None),
                                            C.c_void_p(d_vel), C.c_void_p(stream or None))
    if rc != BFM_OK:
        raise BFMError(rc, "bfm_dev_get_velocities_noalloc")


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Context:
    """One bfm_ctx: device buffers sized for max_cells / max_pairs."""

    def __init__(self, max_cells, max_pairs=None, device=0, params=None):
        self.lib = load_library()
        self.h = C.c_void_p()
        max_pairs = max_pairs or max(16 * max_cells, 1024)
        rc = self.lib.bfm_create(C.byref(self.h), device, int(max_cells), int(max_pairs))
        self._check(rc)
        self.max_cells, self.max_pairs = max_cells, max_pairs
        self.params = None
        if params is not None:
            self.set_params(params)

    def _check(self, rc):
        if rc != BFM_OK:
            msg = self.lib.bfm_last_error(self.h).decode() if self.h else "create failed"
            raise BFMError(rc, msg)

    def close(self):
        if self.h:
            self.lib.bfm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        self.params = p
        self._check(self.lib.bfm_set_params(self.h, C.byref(p)))

    def set_stream(self, stream_handle):
        self._check(self.lib.bfm_set_stream(self.h, C.c_void_p(stream_handle)))

    def sync(self):
        self._check(self.lib.bfm_sync(self.h))

    def set_jkr_table(self, overlaps, radii, gammas=None):
        ov = np.ascontiguousarray(overlaps, dtype=np.float64)
        rad = np.ascontiguousarray(radii, dtype=np.float64)
        g = None if gammas is None else np.ascontiguousarray(gammas, dtype=np.float64)
        self._check(self.lib.bfm_set_jkr_table(self.h, _ptr(ov), len(ov), _ptr(g),
                                               0 if g is None else len(g), _ptr(rad)))

    def set_aniso_jkr_tables(self, tables):
        """bfm_set_aniso_jkr_tables from an AnisoJKRTables."""
        t = tables.cstruct()
        self._check(self.lib.bfm_set_aniso_jkr_tables(self.h, C.byref(t)))

    def set_tuning(self, bin_scale=0.0, block_threads=0, kernel_variant=-1):
        self._check(self.lib.bfm_set_tuning(self.h, float(bin_scale), int(block_threads),
                                            int(kernel_variant)))

    def get_cell_neighbors(self, cells, thr, R, Ldiv):
        cells = _f64(cells)
        n = cells.shape[0]
        cap = max(16 * n, 16)
        while True:
            out = np.zeros((cap, 7), dtype=np.float64, order="F")
            m = C.c_int64()
            rc = self.lib.bfm_get_cell_neighbors(self.h, _ptr(cells), n, n, thr, R, Ldiv,
                                                 _ptr(out), cap, C.byref(m))
            if rc == BFM_ERR_CAPACITY and m.value > cap:
                cap = m.value
                continue
            self._check(rc)
            return np.asfortranarray(out[:m.value])

    def update_neighbor_distances(self, cells, nb):
        cells = _f64(cells)
        assert nb.flags.f_contiguous and nb.dtype == np.float64
        self._check(self.lib.bfm_update_neighbor_distances(
            self.h, _ptr(cells), cells.shape[0], cells.shape[0], _ptr(nb), nb.shape[0],
            max(nb.shape[0], 1)))

    def get_conservative_forces(self, cells, nb, assume_2d):
        cells, nb = _f64(cells), _f64(nb)
        n = cells.shape[0]
        a2d = np.ascontiguousarray(assume_2d, dtype=np.int32)
        out = np.zeros((n, 6), order="F")
        self._check(self.lib.bfm_get_conservative_forces(
            self.h, _ptr(cells), n, n, _ptr(nb), nb.shape[0], max(nb.shape[0], 1),
            a2d.ctypes.data_as(_ip), _ptr(out)))
        return out

    def get_velocities(self, cells, nb, assume_2d, noise=None):
        cells, nb = _f64(cells), _f64(nb)
        n = cells.shape[0]
        a2d = np.ascontiguousarray(assume_2d, dtype=np.int32)
        nz = None if noise is None else _f64(noise)
        out = np.zeros((n, 6), order="F")
        self._check(self.lib.bfm_get_velocities(
            self.h, _ptr(cells), n, n, _ptr(nb), nb.shape[0], max(nb.shape[0], 1),
            a2d.ctypes.data_as(_ip), _ptr(nz), _ptr(out)))
        return out

    def step_rk(self, A, b, bs, dt, cells, nb, noise=None):
        cells, nb = _f64(cells), _f64(nb)
        n = cells.shape[0]
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        bs = np.ascontiguousarray(bs, dtype=np.float64)
        nz = None if noise is None else _f64(noise)
        out = np.zeros((n, cells.shape[1]), order="F")
        err = np.zeros((n, 6), order="F")
        self._check(self.lib.bfm_step_rk(
            self.h, _ptr(A), _ptr(b), _ptr(bs), len(b), dt, _ptr(cells), n, n, _ptr(nb),
            nb.shape[0], max(nb.shape[0], 1), _ptr(nz), _ptr(out), _ptr(err)))
        return out, err

    def step_verlet(self, dt, rho, cells, nb, noise=None):
        """stepVelocityVerlet (mechanics.hpp:2668-2996); returns the new cells."""
        cells, nb = _f64(cells), _f64(nb)
        n = cells.shape[0]
        nz = None if noise is None else _f64(noise)
        out = np.zeros((n, cells.shape[1]), order="F")
        self._check(self.lib.bfm_step_verlet(self.h, dt, rho, _ptr(cells), n, n, _ptr(nb), nb.shape[0],
                                             max(nb.shape[0], 1), _ptr(nz), _ptr(out)))
        return out

    def dev_step_verlet(self, dt, rho, d_cells, n, ld, d_new, ld_new, d_noise=None):
        self._check(self.lib.bfm_dev_step_verlet(self.h, dt, rho, C.c_void_p(d_cells), n, ld,
                                                 C.c_void_p(d_noise) if d_noise else None,
                                                 C.c_void_p(d_new), ld_new))

    # ---- device-resident entry points (device pointers as ints) ----------
    def dev_build_neighbors(self, d_cells, n, ld, thr, R, Ldiv, want_m=True):
        m = C.c_int64()
        self._check(self.lib.bfm_dev_build_neighbors(self.h, C.c_void_p(d_cells), n, ld, thr, R,
                                                     Ldiv, C.byref(m) if want_m else None))
        return m.value if want_m else None

    def dev_export_neighbors(self, d_nb, ld_out):
        self._check(self.lib.bfm_dev_export_neighbors(self.h, C.c_void_p(d_nb), int(ld_out)))

    def dev_import_neighbors(self, d_nb, m, ld, n):
        self._check(self.lib.bfm_dev_import_neighbors(self.h, C.c_void_p(d_nb), int(m), int(ld),
                                                      int(n)))

    def dev_get_velocities(self, d_cells, n, ld, d_vel, d_a2d=0, d_noise=0):
        self._check(self.lib.bfm_dev_get_velocities(self.h, C.c_void_p(d_cells), n, ld,
                                                    C.c_void_p(d_a2d or None),
                                                    C.c_void_p(d_noise or None), C.c_void_p(d_vel)))

    def dev_stage(self, d_cells, n, ld, thr, R, Ldiv, d_vel):
        self._check(self.lib.bfm_dev_stage(self.h, C.c_void_p(d_cells), n, ld, thr, R, Ldiv,
                                           C.c_void_p(d_vel)))

    def dev_step_rk(self, A, b, bs, dt, d_cells, n, ld, d_cells_new, d_errors, tol=0.0,
                    d_err_scaled=0, d_noise=0):
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        bs = np.ascontiguousarray(bs, dtype=np.float64)
        self._check(self.lib.bfm_dev_step_rk(
            self.h, _ptr(A), _ptr(b), _ptr(bs), len(b), dt, C.c_void_p(d_cells), n, ld,
            C.c_void_p(d_noise or None), C.c_void_p(d_cells_new), C.c_void_p(d_errors),
            float(tol), C.c_void_p(d_err_scaled or None)))

    def enable_timing(self, on=True):
        self._check(self.lib.bfm_enable_timing(self.h, int(on)))

    def reset_timing(self):
        self._check(self.lib.bfm_reset_timing(self.h))

    def kernel_times(self):
        cap = 64
        names = C.create_string_buffer(8192)
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        nk = C.c_int()
        self._check(self.lib.bfm_kernel_times(self.h, names, 8192, _ptr(ms),
                                              cnt.ctypes.data_as(_i64p), cap, C.byref(nk)))
        keys = names.value.decode().split("\n")[:nk.value]
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(keys)}

    def stats(self):
        m, scr, tests, edge, nb = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double(), C.c_int()
        self._check(self.lib.bfm_stats(self.h, C.byref(m), C.byref(scr), C.byref(tests),
                                       C.byref(edge), C.byref(nb)))
        return dict(m=m.value, screened=scr.value, segment_tests=tests.value, bin_edge=edge.value,
                    nbins=nb.value)


# ---- reference-named API (mechanics.hpp) -----------------------------------
_CTX = {}


def _ctx_for(n, params):
    key = id(params)
    c = _CTX.get(key)
    if c is None or c.max_cells < n:
        c = Context(max(n, 1024), params=params)
        _CTX.clear()
        _CTX[key] = c
    return c


def getCellNeighbors(cells, neighbor_threshold, R, Ldiv, params=None):
    """mechanics.hpp:520-593. Returns the m x 7 array (i, j, d, s_i, s_j)."""
    params = params or make_params(R=R, ncols=np.shape(cells)[1])
    return _ctx_for(np.shape(cells)[0], params).get_cell_neighbors(cells, neighbor_threshold, R,
                                                                   Ldiv)


def updateNeighborDistances(cells, neighbors, params=None):
    """mechanics.hpp:604-648 (in place on the Fortran-ordered neighbors)."""
    params = params or make_params(ncols=np.shape(cells)[1])
    _ctx_for(np.shape(cells)[0], params).update_neighbor_distances(cells, neighbors)


def getConservativeForces(cells, neighbors, assume_2d, params):
    """mechanics.hpp:1702-1769 with include_constraint = false."""
    return _ctx_for(np.shape(cells)[0], params).get_conservative_forces(cells, neighbors,
                                                                        assume_2d)


def getVelocities(cells, neighbors, assume_2d, noise, params):
    """mechanics.hpp:2011-2347."""
    return _ctx_for(np.shape(cells)[0], params).get_velocities(cells, neighbors, assume_2d, noise)


def stepVelocityVerlet(cells, neighbors, dt, rho, params, noise=None):
    """mechanics.hpp:2668-2996; returns cells_new. noise is drawn by the caller."""
    return _ctx_for(np.shape(cells)[0], params).step_verlet(dt, rho, cells, neighbors, noise)


def stepRungeKuttaAdaptive(A, b, bs, cells, neighbors, dt, params, noise=None):
    """mechanics.hpp:2406-2612; returns (cells_new, errors). noise is drawn by the caller."""
    return _ctx_for(np.shape(cells)[0], params).step_rk(A, b, bs, dt, cells, neighbors, noise)
