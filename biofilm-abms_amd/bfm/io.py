"""Frames, lineage files, the JSON configuration and the driver of the 3-D
model, through the C ABI of include/bfm.h (bfm_write_cells, bfm_read_cells,
bfm_write_lineage, bfm_read_lineage, bfm_load_config, bfm_run).

Names and argument meaning follow ABM-3d/ABM-3d-cpp/include/utils.hpp:95-258
(readCells, writeCells, readLineage, writeLineage) and the driver
ABM-3d/ABM-3d-cpp/src/abmConstAdhesion.cpp (main: config file, output prefix,
random seed).
"""
import ctypes as C

import numpy as np

from .core import BFMError, load_library, JKR_TABLE_1D
from .sim import SimStats

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class RunConfig(C.Structure):
    """bfm_run_config: the keys of abmConstAdhesion.cpp:38-330."""
    _fields_ = [(k, C.c_double) for k in (
        "R", "Rcell", "L0", "Ldiv", "E0", "Ecell", "sigma0", "eta_ambient", "eta_surface",
        "max_stepsize", "min_stepsize", "dt_write", "max_time", "max_error_allowed",
        "growth_mean", "growth_std", "daughter_length_std", "daughter_angle_xy_bound",
        "daughter_angle_z_bound", "nz_threshold", "max_rxy_noise", "max_rz_noise",
        "max_nxy_noise", "max_nz_noise", "basal_min_overlap", "adhesion_eqdist", "jkr_imag_tol",
        "jkr_aberth_tol", "eta_cell_cell", "cell_cell_coulomb_coeff",
        "cell_surface_coulomb_coeff", "M0", "neighbor_threshold")] + \
        [(k, C.c_int) for k in (
            "iter_update_stepsize", "iter_update_neighbors", "max_iter", "n_cells", "basal_only",
            "adhesion_mode", "friction_mode", "integration_mode", "no_surface", "n_mesh_overlap",
            "n_mesh_gamma")] + \
        [("init_filename", C.c_char * 1024), ("lineage_filename", C.c_char * 1024),
         ("precompute_jkr_forces", C.c_int), ("curvature_filename", C.c_char * 1024),
         ("jkr_forces_filename", C.c_char * 1024)] + \
        [(k, C.c_double) for k in (
            "calibrate_endpoint_radii", "min_aspect_ratio", "ellipsoid_project_tol", "brent_tol",
            "init_bracket_dx")] + \
        [(k, C.c_int) for k in (
            "n_mesh_theta", "n_mesh_half_l", "n_mesh_centerline_coords", "n_mesh_curvature_radii",
            "ellipsoid_project_max_iter", "brent_max_iter", "n_tries_bracket")]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = v.decode() if isinstance(v, bytes) else v
        return out


def _check(rc):
    if rc != 0:
        raise BFMError(rc, load_library().bfm_io_last_error().decode())


def writeCells(cells, params, filename, write_other_cols=None):
    """writeCells (utils.hpp:146-198): params is a dict of strings (written
    sorted by key), write_other_cols maps extra column -> 0 (int) / 1 (float)."""
    lib = load_library()
    cells = np.asfortranarray(cells, dtype=np.float64)
    n, ncols = cells.shape
    keys = [str(k).encode() for k in params]
    vals = [str(params[k]).encode() for k in params]
    ka = (C.c_char_p * max(len(keys), 1))(*keys)
    va = (C.c_char_p * max(len(vals), 1))(*vals)
    extra = write_other_cols or {}
    ec = np.array(list(extra.keys()), dtype=np.int32)
    et = np.array([extra[k] for k in extra], dtype=np.int32)
    _check(lib.bfm_write_cells(str(filename).encode(), cells.ctypes.data_as(_dp), n, max(n, 1),
                               ncols, ka, va, len(keys), ec.ctypes.data_as(_ip),
                               et.ctypes.data_as(_ip), len(ec)))


def readCells(filename):
    """readCells (utils.hpp:95-143): (cells n x ncols, Fortran order; params dict)."""
    lib = load_library()
    n, ncols, plen = C.c_int(), C.c_int(), C.c_int()
    fn = str(filename).encode()
    _check(lib.bfm_read_cells(fn, None, 0, 0, 0, C.byref(n), C.byref(ncols), None, 0,
                              C.byref(plen)))
    cells = np.zeros((n.value, ncols.value), order="F")
    buf = C.create_string_buffer(plen.value + 1)
    _check(lib.bfm_read_cells(fn, cells.ctypes.data_as(_dp), n.value, max(n.value, 1),
                              ncols.value, C.byref(n), C.byref(ncols), buf, plen.value + 1,
                              C.byref(plen)))
    params = {}
    for line in buf.value.decode().splitlines():
        k, v = line.split(" = ", 1)
        params[k] = v
    return cells, params


def writeLineage(parents, filename):
    """writeLineage (utils.hpp:243-258)."""
    p = np.ascontiguousarray(parents, dtype=np.int32)
    _check(load_library().bfm_write_lineage(str(filename).encode(), p.ctypes.data_as(_ip), len(p)))


def readLineage(filename):
    """readLineage (utils.hpp:207-233) as the parents vector of the driver
    (abmConstAdhesion.cpp:386-393): parents[child] = parent."""
    lib = load_library()
    n = C.c_int()
    fn = str(filename).encode()
    _check(lib.bfm_read_lineage(fn, None, 0, C.byref(n)))
    out = np.zeros(n.value, dtype=np.int32)
    _check(lib.bfm_read_lineage(fn, out.ctypes.data_as(_ip), n.value, C.byref(n)))
    return out


def readAnisotropicJKRTables(curvature_filename, forces_filename, variable_gamma=False):
    """parseCurvatureRadiiTable + parseReducedJKRForceTable (variable_gamma
    False: the force file's first line is the fixed gamma) or
    parseJKRForceTable (simulation.hpp:565-1008), as an AnisoJKRTables."""
    from .core import AnisoJKRTables, AnisoTablesC
    lib = load_library()
    h = C.POINTER(AnisoTablesC)()
    g = C.c_double()
    _check(lib.bfm_read_aniso_jkr_tables(str(curvature_filename).encode(),
                                         str(forces_filename).encode(), int(bool(variable_gamma)),
                                         C.byref(h), C.byref(g)))
    try:
        t = h.contents
        arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0)  # noqa: E731
        ng = t.n_gamma
        tabs = AnisoJKRTables(
            arr(t.theta, t.n_theta), arr(t.half_l, t.n_half_l), arr(t.coords, t.n_coords),
            arr(t.curvature_rx, t.n_theta * t.n_half_l * t.n_coords),
            arr(t.rx, t.n_rx), arr(t.phi, t.n_phi), arr(t.delta, t.n_delta),
            arr(t.force_radius, t.n_rx * t.n_rx * t.n_phi * t.n_delta * max(ng, 1) * 2),
            gamma=arr(t.gamma, ng) if ng else None,
            fixed_gamma=None if variable_gamma else g.value)
    finally:
        lib.bfm_free_aniso_jkr_tables(h)
    return tabs


def load_config(json_filename):
    """The driver's configuration (abmConstAdhesion.cpp:31-333) as a RunConfig."""
    cfg = RunConfig()
    _check(load_library().bfm_load_config(str(json_filename).encode(), C.byref(cfg)))
    return cfg


def run(json_filename, outprefix, rng_seed, device=0, jkr_mode=JKR_TABLE_1D, max_cells=0,
        verbose=False):
    """main() of abmConstAdhesion.cpp with runSimulation's frame output; returns
    the run statistics (iterations, RK attempts, divisions, final time, ...)."""
    st = SimStats()
    _check(load_library().bfm_run(str(json_filename).encode(), str(outprefix).encode(),
                                  int(rng_seed), int(device), int(jkr_mode), int(max_cells),
                                  int(bool(verbose)), C.byref(st)))
    return {k: getattr(st, k) for k, _ in SimStats._fields_}
