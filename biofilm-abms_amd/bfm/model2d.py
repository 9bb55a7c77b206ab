"""The 2-D model of ABM-2d/ABM-2d-cpp through the C ABI (bfm2d_set_params,
bfm2d_get_cell_neighbors, bfm2d_get_velocities, bfm2d_step_rk), with the
reference's function names: getCellNeighbors (mechanics.hpp:293-374 of
ABM-2d), getVelocities (:1503-1708, no confinement) and
stepRungeKuttaAdaptive (:1710-1817). Cells use the 2-D tree's column layout
(include/indices.hpp of ABM-2d); neighbor rows are (i, j, dx, dy, s, t)."""
import ctypes as C

import numpy as np

from .core import BFMError, Context, load_library  # noqa: F401

# Column indices, include/indices.hpp of ABM-2d.
COL2D = dict(id=0, rx=1, ry=2, nx=3, ny=4, drx=5, dry=6, dnx=7, dny=8, l=9, half_l=10, t0=11,
             growth=12, eta0=13, eta1=14, maxeta1=15, group=16)
NCOLS2D = 17
ADHESION2D_NONE, ADHESION2D_JKR, ADHESION2D_KIHARA, ADHESION2D_JKR_ISOTROPIC = 0, 1, 2, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params2D(C.Structure):
    """bfm2d_params."""
    _fields_ = [("R", C.c_double), ("Rcell", C.c_double), ("prefactors", C.c_double * 3),
                ("surface_contact_density", C.c_double), ("adhesion_mode", C.c_int),
                ("ncols", C.c_int), ("strength", C.c_double), ("mindist", C.c_double),
                ("distance_exp", C.c_double), ("E0", C.c_double), ("gamma", C.c_double)]


def cell_cell_prefactors(R, Rcell, E0, Ecell, cap_exponent=2.5):
    """The repulsion prefactors as the 2-D driver forms them
    (simulation.hpp:231-234 of ABM-2d: 2.5 E0 sqrt(R), 2.5 E0 sqrt(R)
    (2 (R - Rcell))^2.5, 2.5 Ecell sqrt(Rcell)). The force of the Hertz
    potential (potentialHertz, adhesion.hpp:70-92 of ABM-2d) at the cap is
    2.5 E0 sqrt(R) (2 (R - Rcell))^1.5, the exponent the reference's own
    repulsion tests use (cap_exponent=1.5)."""
    return (2.5 * E0 * np.sqrt(R), 2.5 * E0 * np.sqrt(R) * pow(2 * (R - Rcell), cap_exponent),
            2.5 * Ecell * np.sqrt(Rcell))


def make_params(R=0.8, Rcell=0.5, E0=3900.0, Ecell=3.9e6, sigma0=100.0, adhesion_mode=ADHESION2D_NONE,
                strength=1.0, mindist=1.2, distance_exp=6.0, ncols=NCOLS2D, cap_exponent=2.5, gamma=100.0):
    """bfm2d_params; gamma (Pa um) and E0 feed ADHESION2D_JKR_ISOTROPIC."""
    p = Params2D()
    p.R, p.Rcell = R, Rcell
    for k, v in enumerate(cell_cell_prefactors(R, Rcell, E0, Ecell, cap_exponent)):
        p.prefactors[k] = v
    p.surface_contact_density = pow(sigma0 * R * R / (4 * E0), 1. / 3.)
    p.adhesion_mode = adhesion_mode
    p.ncols = ncols
    p.strength, p.mindist, p.distance_exp = strength, mindist, distance_exp
    p.E0, p.gamma = E0, gamma
    return p


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _lib():
    lib = load_library()
    if not getattr(lib, "_bfm2d_sigs", False):
        vp = C.c_void_p
        sig = {
            "bfm2d_set_params": [vp, C.POINTER(Params2D)],
            "bfm2d_get_cell_neighbors": [vp, _dp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         _dp, C.c_int64, C.POINTER(C.c_int64)],
            "bfm2d_get_velocities": [vp, _dp, C.c_int, C.c_int, _dp, C.c_int64, C.c_int64, _ip, _dp, _dp],
            "bfm2d_step_rk": [vp, _dp, _dp, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_int, _dp,
                              C.c_int64, C.c_int64, _ip, _dp, _dp, C.c_int, _dp],
            "bfm2d_dev_build_neighbors": [vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _ip,
                                          C.c_int, C.POINTER(C.c_int64)],
            "bfm2d_dev_update_neighbor_distances": [vp, vp, C.c_int, C.c_int],
            "bfm2d_dev_step_rk": [vp, _dp, _dp, _dp, C.c_int, C.c_double, vp, C.c_int, C.c_int, vp, vp, C.c_int,
                                  vp],
            "bfm2d_simulate": [vp, vp, _ip, C.c_int, _dp, C.c_int, _ip, C.c_int, _dp, C.c_int, _ip, _ip, _ip, _dp,
                               C.c_int, vp],
            "bfm2d_divide_cells": [vp, _dp, _ip, C.c_int, C.c_int, C.c_double, C.c_uint64, _ip, _ip, C.c_int, _ip],
        }
        for name, args in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int
        lib._bfm2d_sigs = True
    return lib


class Model2D:
    """A context running the 2-D mechanics (its own bfm_ctx)."""

    def __init__(self, max_cells, params, max_pairs=None, device=0):
        self.ctx = Context(max_cells, max_pairs=max_pairs, device=device)
        self.lib = _lib()
        self.params = params
        self.ctx._check(self.lib.bfm2d_set_params(self.ctx.h, C.byref(params)))

    def getCellNeighbors(self, cells, neighbor_threshold, R, Ldiv):
        cells = _f64(cells)
        n = cells.shape[0]
        cap = max(8 * n, 16)
        while True:
            out = np.zeros((cap, 6), order="F")
            m = C.c_int64()
            rc = self.lib.bfm2d_get_cell_neighbors(self.ctx.h, _ptr(cells), n, n, neighbor_threshold, R, Ldiv,
                                                   _ptr(out), cap, C.byref(m))
            if rc == 2 and m.value > cap:
                cap = m.value
                continue
            self.ctx._check(rc)
            return np.asfortranarray(out[:m.value])

    def _adhere(self, nb, to_adhere):
        if to_adhere is None:
            to_adhere = np.ones(nb.shape[0], dtype=np.int32)
        return np.ascontiguousarray(to_adhere, dtype=np.int32)

    def getVelocities(self, cells, neighbors, to_adhere=None, noise=None):
        cells, nb = _f64(cells), _f64(neighbors)
        n = cells.shape[0]
        ta = self._adhere(nb, to_adhere)
        nz = None if noise is None else _f64(noise)
        out = np.zeros((n, 4), order="F")
        self.ctx._check(self.lib.bfm2d_get_velocities(
            self.ctx.h, _ptr(cells), n, n, _ptr(nb), nb.shape[0], max(nb.shape[0], 1),
            ta.ctypes.data_as(_ip), _ptr(nz), _ptr(out)))
        return out

    def stepRungeKuttaAdaptive(self, A, b, bs, cells, neighbors, dt, to_adhere=None, noise=None):
        cells, nb = _f64(cells), _f64(neighbors)
        n = cells.shape[0]
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        bs = np.ascontiguousarray(bs, dtype=np.float64)
        ta = self._adhere(nb, to_adhere)
        nz = None if noise is None else _f64(noise)
        out = np.zeros((n, cells.shape[1]), order="F")
        err = np.zeros((n, 4), order="F")
        self.ctx._check(self.lib.bfm2d_step_rk(
            self.ctx.h, _ptr(A), _ptr(b), _ptr(bs), len(b), dt, _ptr(cells), n, n, _ptr(nb), nb.shape[0],
            max(nb.shape[0], 1), ta.ctypes.data_as(_ip), _ptr(nz), _ptr(out), n, _ptr(err)))
        return out, err


def dev_build_neighbors(model, d_cells, n, ld, thr, R, Ldiv, adhesion_map=None, sync=True):
    """bfm2d_dev_build_neighbors on a device array (pointer); adhesion_map:
    an n_groups x n_groups 0/1 matrix or None. Returns the row count (sync)."""
    amap = None if adhesion_map is None else np.ascontiguousarray(adhesion_map, dtype=np.int32)
    m = C.c_int64()
    model.ctx._check(model.lib.bfm2d_dev_build_neighbors(
        model.ctx.h, C.c_void_p(d_cells), n, ld, thr, R, Ldiv,
        None if amap is None else amap.ctypes.data_as(_ip), 0 if amap is None else amap.shape[0],
        C.byref(m) if sync else None))
    return m.value if sync else None


def dev_update_neighbor_distances(model, d_cells, n, ld):
    model.ctx._check(model.lib.bfm2d_dev_update_neighbor_distances(model.ctx.h, C.c_void_p(d_cells), n, ld))


def dev_step_rk(model, A, b, bs, dt, d_cells, n, ld, d_out, ld_out, d_noise=None, d_errors=None):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    bs = np.ascontiguousarray(bs, dtype=np.float64)
    model.ctx._check(model.lib.bfm2d_dev_step_rk(
        model.ctx.h, _ptr(A), _ptr(b), _ptr(bs), len(b), dt, C.c_void_p(d_cells), n, ld,
        C.c_void_p(d_noise) if d_noise else None, C.c_void_p(d_out), ld_out,
        C.c_void_p(d_errors) if d_errors else None))


def to_adhere_rule(cells, nb, R, Rcell, adhesion_map):
    """The 2-D driver's to_adhere for host rows (simulation.hpp:242-262 of ABM-2d)."""
    amap = np.asarray(adhesion_map)
    g = cells[:, 16].astype(int)
    gi, gj = g[nb[:, 0].astype(int)], g[nb[:, 1].astype(int)]
    lo, hi = np.minimum(gi, gj), np.maximum(gi, gj)
    dist = np.sqrt(nb[:, 2] ** 2 + nb[:, 3] ** 2)
    ok = (lo >= 0) & (hi < amap.shape[0])
    flag = np.zeros(nb.shape[0], dtype=np.int32)
    flag[ok] = amap[lo[ok], hi[ok]] != 0
    return (flag.astype(bool) & (dist > R + Rcell) & (dist < 2 * R)).astype(np.int32)


def random_layer(n, seed, spacing=(3.0, 1.4), tilt=0.3, hl_lo=0.4, hl_hi=1.2, eta0=0.7, eta1=720.0):
    """n 2-D cells in rows (for tests and benchmarks): centers on a lattice
    of the given spacing with jitter (0.2 along a row, 0.1 across), nearly
    horizontal orientations (angle within +-tilt, either direction) and
    half-lengths in [hl_lo, hl_hi], so that no two centerlines cross (a zero
    distance has no direction) while neighbors along and across the rows are
    in contact. The 2-D tree's columns."""
    rng = np.random.default_rng(seed)
    sx, sy = spacing
    cols = max(1, int(np.ceil(np.sqrt(n * sy / sx))))
    k = np.arange(n)
    c = np.zeros((n, NCOLS2D), order="F")
    c[:, 0] = k
    c[:, 1] = (k % cols) * sx + rng.uniform(-0.2, 0.2, n)
    c[:, 2] = (k // cols) * sy + rng.uniform(-0.1, 0.1, n)
    th = rng.uniform(-tilt, tilt, n) + np.pi * (rng.uniform(0, 1, n) < 0.5)
    c[:, 3], c[:, 4] = np.cos(th), np.sin(th)
    hl = rng.uniform(hl_lo, hl_hi, n)
    c[:, 9], c[:, 10] = 2 * hl, hl
    c[:, 12], c[:, 13], c[:, 14], c[:, 15], c[:, 16] = 1.5, eta0, eta1, eta1, 1
    c[:, 5:9] = rng.normal(size=(n, 4))
    return c


# ---- the growth loop (runSimulationAdaptiveLagrangian of ABM-2d) -------------

def sim_config_2d(n_cells=0, max_iter=0, max_time=0.0, R=0.8, Rcell=0.5, L0=0.4, max_stepsize=1e-4,
                  min_stepsize=1e-9, max_error_allowed=1e-8, min_error=1e-30, growth_mean=1.5, growth_std=0.15,
                  daughter_length_std=0.125, daughter_angle_bound=0.05, iter_update_neighbors=10,
                  iter_update_stepsize=1, max_tries_update_stepsize=3, tableau="bogacki_shampine", seed=42,
                  report_seconds=0.0, max_wall_seconds=0.0):
    """bfm_sim_config for bfm2d_simulate, named as the 2-D driver's JSON keys
    (abmDualAdhesionSwitch.cpp:37-61 of ABM-2d): Ldiv = 2 L0 + 2 R,
    neighbor_threshold = 2 R + Ldiv, one growth mean/std for both groups,
    max_tries_update_stepsize 3 and min_error 1e-30 (the driver's constants,
    :25-29); the tableau defaults to the 2-D tree's Bogacki-Shampine."""
    from .sim import sim_config
    c = sim_config(n_cells=n_cells, max_iter=max_iter, max_time=max_time, R=R, Rcell=Rcell, L0=L0,
                   max_stepsize=max_stepsize, min_stepsize=min_stepsize, max_error_allowed=max_error_allowed,
                   min_error=min_error, growth_mean=(growth_mean, growth_mean), growth_std=(growth_std, growth_std),
                   daughter_length_std=daughter_length_std, daughter_angle_xy_bound=daughter_angle_bound,
                   daughter_angle_z_bound=0.0, iter_update_neighbors=iter_update_neighbors,
                   iter_update_stepsize=iter_update_stepsize, max_tries_update_stepsize=max_tries_update_stepsize,
                   tableau=tableau, seed=seed, report_seconds=report_seconds, max_wall_seconds=max_wall_seconds)
    c.neighbor_threshold = 2 * R + c.Ldiv
    return c


def founder_cell_2d(L0=0.4, growth_mean=1.5, eta_ambient=0.7, eta_surface=720.0, ncols=NCOLS2D):
    """The 2-D driver's founder (abmDualAdhesionSwitch.cpp:133-135 of ABM-2d):
    at the origin along x, length L0, mean growth rate, group 1."""
    c = np.zeros((1, ncols), order="F")
    c[0, COL2D["nx"]] = 1.0
    c[0, COL2D["l"]] = L0
    c[0, COL2D["half_l"]] = L0 / 2
    c[0, COL2D["growth"]] = growth_mean
    c[0, COL2D["eta0"]] = eta_ambient
    c[0, COL2D["eta1"]] = eta_surface
    c[0, COL2D["maxeta1"]] = eta_surface
    c[0, COL2D["group"]] = 1
    return c


def simulate_2d(model, cfg, cells, adhesion_map=((0, 0), (0, 1)), cap=None, trace_cap=2_000_000,
                allow_nan=False):
    """bfm2d_simulate on the model's context from `cells` (n x ncols, the
    2-D columns); adhesion_map: n_groups x n_groups 0/1 (groups are column
    16; the default lets group-1 cells adhere to each other, as the 2-D
    drivers do) or None. Returns (cells, parents, dt_trace, stats dict).
    allow_nan: a run stopped by a non-finite stage velocity returns its
    last state with stats["nan"] = True instead of raising."""
    from .sim import SimStats
    lib = _lib()
    cells = _f64(cells)
    n_in, ncols = cells.shape
    cap = int(cap or model.ctx.max_cells)
    amap = None if adhesion_map is None else np.ascontiguousarray(adhesion_map, dtype=np.int32)
    out = np.zeros((cap, ncols), order="F")
    parents = np.zeros(int(cells[:, 0].max()) + 1 + 2 * cap + 8, dtype=np.int32)
    n_out, n_par = C.c_int(), C.c_int()
    trace = np.zeros(trace_cap)
    st = SimStats()
    rc = lib.bfm2d_simulate(
        model.ctx.h, C.cast(C.byref(cfg), C.c_void_p), None if amap is None else amap.ctypes.data_as(_ip),
        0 if amap is None else amap.shape[0], _ptr(cells), n_in, None, 0, _ptr(out), cap, C.byref(n_out),
        parents.ctypes.data_as(_ip), C.byref(n_par), _ptr(trace), trace_cap, C.cast(C.byref(st), C.c_void_p))
    if not (allow_nan and rc == 3):
        model.ctx._check(rc)
    it = min(st.iterations, trace_cap)
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    stats["nan"] = rc == 3
    return np.asfortranarray(out[:n_out.value]), parents[:n_par.value].copy(), trace[:it].copy(), stats


def divide_cells_2d(cfg, cells, t, seed, parents=None):
    """bfm2d_divide_cells (host only) on a copy; returns (cells, parents, tries)."""
    lib = _lib()
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
    rc = lib.bfm2d_divide_cells(C.cast(C.byref(cfg), C.c_void_p), _ptr(buf), C.byref(n_io), cap, ncols, t, seed,
                                par.ctypes.data_as(_ip), C.byref(n_par), pcap, C.byref(tries))
    if rc != 0:
        raise BFMError(rc, "bfm2d_divide_cells failed")
    return np.asfortranarray(buf[:n_io.value]), par[:n_par.value].copy(), tries.value
