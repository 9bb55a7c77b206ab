"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker of the HIP path. Array layouts follow include/bfm.h
(column-major numpy arrays).
"""
import ctypes as C
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        L = C.CDLL(path)
        d, i64 = C.c_double, C.c_int64
        L.oracle_nearest_cell_body_coord.argtypes = [_dp, _dp, d, _dp]
        L.oracle_nearest_cell_body_coord.restype = d
        L.oracle_dist_between_cells.argtypes = [_dp, _dp, d, _dp, _dp, d, _dp, _dp, _dp]
        L.oracle_dist_between_cells.restype = None
        for nm in ("oracle_get_cell_neighbors",):
            getattr(L, nm).argtypes = [_dp, C.c_int, C.c_int, d, d, d, _dp, i64]
            getattr(L, nm).restype = i64
        L.oracle_get_cell_neighbors_binned.argtypes = [_dp, C.c_int, C.c_int, d, d, d, _dp, i64,
                                                       C.c_int]
        L.oracle_get_cell_neighbors_binned.restype = i64
        L.oracle_count_neighbor_rows.argtypes = [_dp, C.c_int, C.c_int, d, d, d, C.c_int,
                                                 C.c_int]
        L.oracle_count_neighbor_rows.restype = i64
        L.oracle_update_neighbor_distances.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64]
        L.oracle_update_neighbor_distances.restype = None
        L.oracle_jkr_contact_radius.argtypes = [d, d, d, d, _ip]
        L.oracle_jkr_contact_radius.restype = d
        L.oracle_jkr_force_magnitude.argtypes = [d, d, d, d]
        L.oracle_jkr_force_magnitude.restype = d
        L.oracle_jkr_equilibrium_distance.argtypes = [d] * 7 + [C.c_int, _dp]
        L.oracle_jkr_equilibrium_distance.restype = d
        L.oracle_jkr_optimal_surface_energy_density.argtypes = [d] * 9 + [C.c_int, C.c_int, _dp]
        L.oracle_jkr_optimal_surface_energy_density.restype = d
        L.oracle_linspace.argtypes = [d, d, C.c_int, _dp]
        L.oracle_linspace.restype = None
        L.oracle_nearest_value.argtypes = [_dp, C.c_int, d]
        L.oracle_nearest_value.restype = C.c_int
        L.oracle_jkr_table_radius_1d.argtypes = [_dp, _dp, C.c_int, d]
        L.oracle_jkr_table_radius_1d.restype = d
        L.oracle_jkr_table_radius_2d.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, d, d]
        L.oracle_jkr_table_radius_2d.restype = d
        L.oracle_set_jkr_table.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp]
        L.oracle_set_jkr_table.restype = None
        L.oracle_set_aniso_tables.argtypes = [_dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp,
                                              C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp]
        L.oracle_set_aniso_tables.restype = None
        L.oracle_aniso_jkr_force.argtypes = [_dp, d, _dp, d, _dp, d, d, d, d, d, _dp, _dp]
        L.oracle_aniso_jkr_force.restype = None
        for nm in ("oracle_integral1", "oracle_integral2", "oracle_integral3"):
            getattr(L, nm).argtypes = [d] * 6
            getattr(L, nm).restype = d
        for nm in ("oracle_integral4", "oracle_integral5", "oracle_integral6"):
            getattr(L, nm).argtypes = [d] * 3
            getattr(L, nm).restype = d
        L.oracle_area_integrals.argtypes = [d, d, d, d, d, _dp]
        L.oracle_area_integrals.restype = None
        vp = C.c_void_p
        L.oracle_cell_cell_forces.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, vp, _dp]
        L.oracle_friction_forces.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, vp, _dp]
        L.oracle_cell_cell_forces_ic.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, vp, _dp]
        L.oracle_conservative_forces_ic.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, _ip, vp, _dp]
        L.oracle_ambient_viscosity_forces.argtypes = [_dp, C.c_int, C.c_int, _dp]
        L.oracle_surface_friction_forces.argtypes = [_dp, C.c_int, C.c_int, _dp, d, _ip, _dp]
        L.oracle_step_verlet.argtypes = [d, d, _dp, C.c_int, C.c_int, _dp, i64, i64, _dp, vp, _dp]
        L.oracle_surface_repulsion_forces.argtypes = [_dp, C.c_int, C.c_int, _dp, d, d, _ip, _dp]
        L.oracle_surface_adhesion_forces.argtypes = [_dp, C.c_int, C.c_int, _dp, d, _ip, _dp]
        L.oracle_conservative_forces.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, _ip, vp,
                                                 _dp]
        L.oracle_get_velocities.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, _ip, _dp, vp,
                                            _dp]
        L.oracle_normalize_orientations.argtypes = [_dp, C.c_int, C.c_int]
        L.oracle_step_rk.argtypes = [_dp, _dp, _dp, C.c_int, d, _dp, C.c_int, C.c_int, _dp, i64,
                                     i64, _dp, vp, _dp, _dp]
        L.oracle_max_scaled_error.argtypes = [_dp, C.c_int, C.c_int, _dp, d, d]
        L.oracle_max_scaled_error.restype = d
        L.oracle_controller_trace.argtypes = [d, d, C.c_int, d, d, _dp, C.c_int, _dp]
        L.oracle_controller_trace.restype = C.c_int
        L.oracle_divide_cells.argtypes = [vp, _dp, _ip, C.c_int, C.c_int, d, C.c_uint64, _ip, _ip,
                                          C.c_int, _ip]
        L.oracle_divide_cells.restype = C.c_int
        L.oracle_simulate.argtypes = [vp, vp, _dp, C.c_int, _dp, C.c_int, _ip, _dp, C.c_int,
                                      C.POINTER(i64), C.POINTER(i64), C.c_int]
        L.oracle_simulate.restype = C.c_int
        _LIB = L
    return _LIB


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _v3(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def nearest_cell_body_coord(r, n, half_l, q):
    return lib().oracle_nearest_cell_body_coord(_p(_v3(r)), _p(_v3(n)), half_l, _p(_v3(q)))


def dist_between_cells(r1, n1, hl1, r2, n2, hl2):
    d = np.zeros(3)
    s, t = C.c_double(), C.c_double()
    lib().oracle_dist_between_cells(_p(_v3(r1)), _p(_v3(n1)), hl1, _p(_v3(r2)), _p(_v3(n2)), hl2,
                                    _p(d), C.byref(s), C.byref(t))
    return d, s.value, t.value


def get_cell_neighbors(cells, thr, R, Ldiv, binned=False, nthreads=0):
    cells = _f(cells)
    n = cells.shape[0]
    cap = max(16 * n, 16)
    while True:
        out = np.zeros((cap, 7), order="F")
        if binned:
            m = lib().oracle_get_cell_neighbors_binned(_p(cells), n, n, thr, R, Ldiv, _p(out), cap,
                                                       nthreads)
        else:
            m = lib().oracle_get_cell_neighbors(_p(cells), n, n, thr, R, Ldiv, _p(out), cap)
        if m <= cap:
            return np.asfortranarray(out[:m])
        cap = m


def count_neighbor_rows(cells, thr, R, Ldiv, row_lo, row_hi):
    cells = _f(cells)
    return lib().oracle_count_neighbor_rows(_p(cells), cells.shape[0], cells.shape[0], thr, R,
                                            Ldiv, row_lo, row_hi)


def update_neighbor_distances(cells, nb):
    cells = _f(cells)
    lib().oracle_update_neighbor_distances(_p(cells), cells.shape[0], cells.shape[0], _p(nb),
                                           nb.shape[0], max(nb.shape[0], 1))


def jkr_contact_radius(delta, Req, E, gamma):
    it = C.c_int()
    a = lib().oracle_jkr_contact_radius(delta, Req, E, gamma, C.byref(it))
    return a, it.value


def jkr_force_magnitude(radius, Req, E0, gamma):
    return lib().oracle_jkr_force_magnitude(radius, Req, E0, gamma)


def jkr_equilibrium_distance(R, E0, gamma, dinit, dmin, increment_overlap=1e-8,
                            newton_tol=1e-8, max_iter=1000):
    """oracle_jkr_equilibrium_distance; NaN when an iterate has no JKR root."""
    return lib().oracle_jkr_equilibrium_distance(R, E0, gamma, dinit, dmin, increment_overlap,
                                                 newton_tol, max_iter, None)


def jkr_optimal_surface_energy_density(R, Rcell, E0, deq_target, init_gamma=10.0,
                                       log_increment_gamma=1e-6, increment_overlap=1e-8,
                                       outer_tol=1e-8, inner_tol=1e-8, outer_max_iter=1000,
                                       inner_max_iter=1000):
    return lib().oracle_jkr_optimal_surface_energy_density(
        R, Rcell, E0, deq_target, init_gamma, log_increment_gamma, increment_overlap,
        outer_tol, inner_tol, outer_max_iter, inner_max_iter, None)


def linspace(lo, hi, n):
    out = np.zeros(n)
    lib().oracle_linspace(lo, hi, n, _p(out))
    return out


def nearest_value(values, x):
    v = np.ascontiguousarray(values, dtype=np.float64)
    return lib().oracle_nearest_value(_p(v), len(v), x)


def set_jkr_table(overlaps, radii, gammas=None):
    ov = np.ascontiguousarray(overlaps, dtype=np.float64)
    rad = np.ascontiguousarray(radii, dtype=np.float64)
    g = None if gammas is None else np.ascontiguousarray(gammas, dtype=np.float64)
    lib().oracle_set_jkr_table(_p(ov), len(ov), _p(g), 0 if g is None else len(g), _p(rad))


def set_aniso_tables(t):
    """Registers an AnisoJKRTables (bfm.core) with the oracle."""
    a = lambda x: np.ascontiguousarray(x, dtype=np.float64)   # noqa: E731
    keep = [a(t.theta), a(t.half_l), a(t.coords), a(t.curvature_rx), a(t.rx), a(t.phi),
            a(t.delta), None if t.gamma is None else a(t.gamma), a(t.force_radius)]
    th, hl, co, cv, rx, ph, de, ga, fr = keep
    lib().oracle_set_aniso_tables(_p(th), len(th), _p(hl), len(hl), _p(co), len(co), _p(cv),
                                  _p(rx), len(rx), _p(ph), len(ph), _p(de), len(de), _p(ga),
                                  0 if ga is None else len(ga), _p(fr))


def aniso_jkr_force(n1, half_l1, n2, half_l2, d12, R, gamma, s, t, max_overlap=-1.0):
    """(force, radius) of forcesAnisotropicJKRLagrange's table form."""
    a = lambda x: np.ascontiguousarray(x, dtype=np.float64)   # noqa: E731
    n1, n2, d12 = a(n1), a(n2), a(d12)
    f, r = C.c_double(), C.c_double()
    lib().oracle_aniso_jkr_force(_p(n1), half_l1, _p(n2), half_l2, _p(d12), R, gamma, s, t,
                                 max_overlap, C.byref(f), C.byref(r))
    return f.value, r.value


def table_radius_1d(overlaps, radii, delta):
    ov = np.ascontiguousarray(overlaps, dtype=np.float64)
    rad = np.ascontiguousarray(radii, dtype=np.float64)
    return lib().oracle_jkr_table_radius_1d(_p(ov), _p(rad), len(ov), delta)


def table_radius_2d(overlaps, gammas, radii, delta, gamma):
    ov = np.ascontiguousarray(overlaps, dtype=np.float64)
    g = np.ascontiguousarray(gammas, dtype=np.float64)
    rad = np.ascontiguousarray(radii, dtype=np.float64)
    return lib().oracle_jkr_table_radius_2d(_p(ov), len(ov), _p(g), len(g), _p(rad), delta, gamma)


def integral(k, *args):
    return getattr(lib(), f"oracle_integral{k}")(*args)


def area_integrals(rz, nz, R, hl, ss):
    out = np.zeros(3)
    lib().oracle_area_integrals(rz, nz, R, hl, ss, _p(out))
    return out


def cell_cell_forces(cells, nb, params):
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    out = np.zeros((n, 6), order="F")
    lib().oracle_cell_cell_forces(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                                  C.byref(params), _p(out))
    return out


def friction_forces(cells, nb, params):
    """cellCellFrictionForces (mechanics.hpp:1031-1175): dE/dq n x 6."""
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    out = np.zeros((n, 6), order="F")
    lib().oracle_friction_forces(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                                 C.byref(params), _p(out))
    return out


def cell_cell_forces_ic(cells, nb, params):
    """cellCellInteractionForces with include_constraint = true: dE/dq n x 6."""
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    out = np.zeros((n, 6), order="F")
    lib().oracle_cell_cell_forces_ic(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                                     C.byref(params), _p(out))
    return out


def conservative_forces_ic(cells, nb, assume_2d, params):
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    a2d = np.ascontiguousarray(assume_2d, dtype=np.int32)
    out = np.zeros((n, 6), order="F")
    lib().oracle_conservative_forces_ic(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                                        a2d.ctypes.data_as(_ip), C.byref(params), _p(out))
    return out


def step_verlet(dt, rho, cells, nb, params, noise=None):
    """stepVelocityVerlet (mechanics.hpp:2668-2996); returns the new cells."""
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    out = np.zeros((n, params.ncols), order="F")
    nz = _f(noise) if noise is not None else None
    lib().oracle_step_verlet(dt, rho, _p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                             _p(nz) if nz is not None else None, C.byref(params), _p(out))
    return out


def _ss(cells, R, a2d):
    rz, nz = cells[:, 3], cells[:, 6]
    with np.errstate(divide="ignore", invalid="ignore"):
        ss = (R - rz) / nz
    ss[np.asarray(a2d) != 0] = np.nan
    return np.ascontiguousarray(ss)


def surface_repulsion_forces(cells, R, E0, assume_2d):
    cells = _f(cells)
    n = cells.shape[0]
    a2d = np.ascontiguousarray(assume_2d, dtype=np.int32)
    out = np.zeros((n, 6), order="F")
    lib().oracle_surface_repulsion_forces(_p(cells), n, n, _p(_ss(cells, R, a2d)), R, E0,
                                          a2d.ctypes.data_as(_ip), _p(out))
    return out


def surface_adhesion_forces(cells, R, assume_2d):
    cells = _f(cells)
    n = cells.shape[0]
    a2d = np.ascontiguousarray(assume_2d, dtype=np.int32)
    out = np.zeros((n, 6), order="F")
    lib().oracle_surface_adhesion_forces(_p(cells), n, n, _p(_ss(cells, R, a2d)), R,
                                         a2d.ctypes.data_as(_ip), _p(out))
    return out


def conservative_forces(cells, nb, assume_2d, params):
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    a2d = np.ascontiguousarray(assume_2d, dtype=np.int32)
    out = np.zeros((n, 6), order="F")
    lib().oracle_conservative_forces(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                                     a2d.ctypes.data_as(_ip), C.byref(params), _p(out))
    return out


def get_velocities(cells, nb, assume_2d, noise, params):
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    a2d = np.ascontiguousarray(assume_2d, dtype=np.int32)
    nz = None if noise is None else _f(noise)
    out = np.zeros((n, 6), order="F")
    lib().oracle_get_velocities(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                                a2d.ctypes.data_as(_ip), _p(nz), C.byref(params), _p(out))
    return out


def normalize_orientations(cells):
    cells = _f(cells).copy(order="F")
    lib().oracle_normalize_orientations(_p(cells), cells.shape[0], cells.shape[0])
    return cells


def step_rk(A, b, bs, dt, cells, nb, params, noise=None):
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    bs = np.ascontiguousarray(bs, dtype=np.float64)
    nz = None if noise is None else _f(noise)
    out = np.zeros((n, cells.shape[1]), order="F")
    err = np.zeros((n, 6), order="F")
    lib().oracle_step_rk(_p(A), _p(b), _p(bs), len(b), dt, _p(cells), n, n, _p(nb), nb.shape[0],
                         max(nb.shape[0], 1), _p(nz), C.byref(params), _p(out), _p(err))
    return out, err


def max_scaled_error(cells, errors, tol, min_error=1e-30):
    cells, errors = _f(cells), _f(errors)
    return lib().oracle_max_scaled_error(_p(cells), cells.shape[0], cells.shape[0], _p(errors),
                                         tol, min_error)


def controller_trace(dt, error_order, max_tries, min_step, max_step, errs):
    e = np.ascontiguousarray(errs, dtype=np.float64)
    out = np.zeros(max_tries + 4)
    k = lib().oracle_controller_trace(dt, error_order, max_tries, min_step, max_step, _p(e), len(e),
                                      _p(out))
    return out[:k]


def divide_cells(cfg, cells, t, seed, parents=None):
    """oracle_divide_cells on a copy; returns (cells, parents, tries)."""
    cells = np.asarray(cells, dtype=np.float64)
    n, ncols = cells.shape
    cap = 2 * n + 1
    buf = np.zeros((cap, ncols), order="F")
    buf[:n] = cells
    if parents is None:
        parents = -np.ones(int(cells[:, 0].max()) + 1, dtype=np.int32)
    pcap = len(parents) + 2 * n + 2
    par = np.zeros(pcap, dtype=np.int32)
    par[:len(parents)] = parents
    n_io, n_par, tries = C.c_int(n), C.c_int(len(parents)), C.c_int()
    rc = lib().oracle_divide_cells(C.byref(cfg), _p(buf), C.byref(n_io), cap, ncols, t, seed,
                                   par.ctypes.data_as(_ip), C.byref(n_par), pcap, C.byref(tries))
    if rc != 0:
        raise RuntimeError(f"oracle_divide_cells: status {rc}")
    return np.asfortranarray(buf[:n_io.value]), par[:n_par.value].copy(), tries.value


def simulate(cfg, params, cells, cap, trace_cap=1_000_000, nthreads=0):
    """oracle_simulate; returns (cells, dt_trace, iterations, rk_steps)."""
    cells = _f(cells)
    n_in = cells.shape[0]
    out = np.zeros((cap, params.ncols), order="F")
    n_out = C.c_int()
    trace = np.zeros(trace_cap)
    iters, steps = C.c_int64(), C.c_int64()
    rc = lib().oracle_simulate(C.byref(cfg), C.byref(params), _p(cells), n_in, _p(out), cap,
                               C.byref(n_out), _p(trace), trace_cap, C.byref(iters),
                               C.byref(steps), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_simulate: status {rc}")
    it = min(iters.value, trace_cap)
    return np.asfortranarray(out[:n_out.value]), trace[:it].copy(), iters.value, steps.value


# ---- the 2-D model -----------------------------------------------------------
def _lib2d():
    L = lib()
    if not getattr(L, "_o2d", False):
        d, i64 = C.c_double, C.c_int64
        vp = C.c_void_p
        L.oracle2d_get_cell_neighbors.argtypes = [_dp, C.c_int, C.c_int, d, d, d, _dp, i64]
        L.oracle2d_get_cell_neighbors.restype = i64
        L.oracle2d_forces.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, _ip, vp, _dp, _dp]
        L.oracle2d_forces.restype = None
        L.oracle2d_get_velocities.argtypes = [_dp, C.c_int, C.c_int, _dp, i64, i64, _ip, _dp, vp, _dp]
        L.oracle2d_get_velocities.restype = None
        L.oracle2d_step_rk.argtypes = [_dp, _dp, _dp, C.c_int, d, _dp, C.c_int, C.c_int, C.c_int, _dp, i64,
                                       i64, _ip, _dp, vp, _dp, _dp]
        L.oracle2d_step_rk.restype = None
        L.oracle2d_divide_cells.argtypes = [vp, _dp, _ip, C.c_int, C.c_int, d, C.c_uint64, _ip, _ip, C.c_int, _ip]
        L.oracle2d_divide_cells.restype = C.c_int
        L.oracle2d_simulate.argtypes = [vp, vp, _ip, C.c_int, _dp, C.c_int, _dp, C.c_int, _ip, _ip, _ip, _dp,
                                        C.c_int, C.POINTER(i64), C.POINTER(i64)]
        L.oracle2d_simulate.restype = C.c_int
        L._o2d = True
    return L


def _adh(nb, to_adhere):
    if to_adhere is None:
        to_adhere = np.ones(nb.shape[0], dtype=np.int32)
    return np.ascontiguousarray(to_adhere, dtype=np.int32)


def get_cell_neighbors_2d(cells, thr, R, Ldiv):
    cells = _f(cells)
    n = cells.shape[0]
    cap = max(8 * n, 16)
    while True:
        out = np.zeros((cap, 6), order="F")
        m = _lib2d().oracle2d_get_cell_neighbors(_p(cells), n, n, thr, R, Ldiv, _p(out), cap)
        if m <= cap:
            return np.asfortranarray(out[:m])
        cap = m


def forces_2d(cells, nb, params, to_adhere=None):
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    rep, adh = np.zeros((n, 4), order="F"), np.zeros((n, 4), order="F")
    ta = _adh(nb, to_adhere)
    _lib2d().oracle2d_forces(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                             ta.ctypes.data_as(_ip), C.cast(C.byref(params), C.c_void_p), _p(rep), _p(adh))
    return rep, adh


def get_velocities_2d(cells, nb, params, to_adhere=None, noise=None):
    cells, nb = _f(cells), _f(nb)
    n = cells.shape[0]
    out = np.zeros((n, 4), order="F")
    ta = _adh(nb, to_adhere)
    nz = None if noise is None else _f(noise)
    _lib2d().oracle2d_get_velocities(_p(cells), n, n, _p(nb), nb.shape[0], max(nb.shape[0], 1),
                                     ta.ctypes.data_as(_ip), _p(nz), C.cast(C.byref(params), C.c_void_p),
                                     _p(out))
    return out


def step_rk_2d(A, b, bs, dt, cells, nb, params, to_adhere=None, noise=None):
    cells, nb = _f(cells), _f(nb)
    n, ncols = cells.shape
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    bs = np.ascontiguousarray(bs, dtype=np.float64)
    out = np.zeros((n, ncols), order="F")
    err = np.zeros((n, 4), order="F")
    ta = _adh(nb, to_adhere)
    nz = None if noise is None else _f(noise)
    _lib2d().oracle2d_step_rk(_p(A), _p(b), _p(bs), len(b), dt, _p(cells), n, n, ncols, _p(nb), nb.shape[0],
                              max(nb.shape[0], 1), ta.ctypes.data_as(_ip), _p(nz),
                              C.cast(C.byref(params), C.c_void_p), _p(out), _p(err))
    return out, err


def divide_cells_2d(cfg, cells, t, seed, parents=None):
    """oracle2d_divide_cells on a copy (2-D columns); returns (cells, parents, tries)."""
    cells = np.asarray(cells, dtype=np.float64)
    n, ncols = cells.shape
    cap = 2 * n + 1
    buf = np.zeros((cap, ncols), order="F")
    buf[:n] = cells
    if parents is None:
        parents = -np.ones(int(cells[:, 0].max()) + 1, dtype=np.int32)
    pcap = len(parents) + 2 * n + 2
    par = np.zeros(pcap, dtype=np.int32)
    par[:len(parents)] = parents
    n_io, n_par, tries = C.c_int(n), C.c_int(len(parents)), C.c_int()
    rc = _lib2d().oracle2d_divide_cells(C.cast(C.byref(cfg), C.c_void_p), _p(buf), C.byref(n_io), cap, ncols, t,
                                        seed, par.ctypes.data_as(_ip), C.byref(n_par), pcap, C.byref(tries))
    if rc != 0:
        raise RuntimeError(f"oracle2d_divide_cells: status {rc}")
    return np.asfortranarray(buf[:n_io.value]), par[:n_par.value].copy(), tries.value


def simulate_2d(cfg, params, adhesion_map, cells, cap, trace_cap=1_000_000):
    """oracle2d_simulate; returns (cells, parents, dt_trace, iterations, rk_steps)."""
    cells = _f(cells)
    n_in = cells.shape[0]
    out = np.zeros((cap, params.ncols), order="F")
    amap = None if adhesion_map is None else np.ascontiguousarray(adhesion_map, dtype=np.int32)
    ng = 0 if amap is None else amap.shape[0]
    n_out, n_par = C.c_int(), C.c_int()
    par = np.zeros(int(cells[:, 0].max()) + 1 + 2 * cap + 2, dtype=np.int32)
    trace = np.zeros(trace_cap)
    iters, steps = C.c_int64(), C.c_int64()
    rc = _lib2d().oracle2d_simulate(C.cast(C.byref(cfg), C.c_void_p), C.cast(C.byref(params), C.c_void_p),
                                    None if amap is None else amap.ctypes.data_as(_ip), ng, _p(cells), n_in,
                                    _p(out), cap, C.byref(n_out), par.ctypes.data_as(_ip), C.byref(n_par),
                                    _p(trace), trace_cap, C.byref(iters), C.byref(steps))
    if rc != 0:
        raise RuntimeError(f"oracle2d_simulate: status {rc}")
    it = min(iters.value, trace_cap)
    return (np.asfortranarray(out[:n_out.value]), par[:n_par.value].copy(), trace[:it].copy(), iters.value,
            steps.value)
