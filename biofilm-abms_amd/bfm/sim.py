"""Growth / division loop of runSimulation (ABM-3d/ABM-3d-cpp/include/simulation.hpp:1106-2081)
over the device mechanical step, through bfm_simulate of include/bfm.h.

The loop keeps the reference's order per iteration: divide the cells past the
division length (divideCells, growth.hpp:456-720, with the reference's random
draws), rebuild the neighbor list after a division, take an adaptive RK step
with the step-size controller (simulation.hpp:1677-1769), grow the cells
(growth.hpp:51-58), refresh the neighbor distances, advance t and iter, and
rebuild the list every iter_update_neighbors iterations.
"""
import ctypes as C

import numpy as np

from .core import COL, Context, load_library, BFMError

TABLEAU_IDS = {"heun_euler": 0, "bogacki_shampine": 1, "rkf45": 2, "dormand_prince": 3,
               "velocity_verlet": -1}


class SimConfig(C.Structure):
    _fields_ = [("R", C.c_double), ("Rcell", C.c_double), ("Ldiv", C.c_double),
                ("max_stepsize", C.c_double), ("min_stepsize", C.c_double),
                ("max_error_allowed", C.c_double), ("min_error", C.c_double),
                ("neighbor_threshold", C.c_double),
                ("growth_mean", C.c_double * 2), ("growth_std", C.c_double * 2),
                ("daughter_length_std", C.c_double), ("daughter_angle_xy_bound", C.c_double),
                ("daughter_angle_z_bound", C.c_double), ("max_time", C.c_double),
                ("iter_update_neighbors", C.c_int), ("iter_update_stepsize", C.c_int),
                ("max_tries_update_stepsize", C.c_int), ("n_cells", C.c_int),
                ("max_iter", C.c_int), ("tableau", C.c_int), ("seed", C.c_uint64),
                ("report_seconds", C.c_double), ("max_wall_seconds", C.c_double),
                ("rho", C.c_double)]


class SimStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("rk_steps", C.c_int64), ("divisions", C.c_int64),
                ("neighbor_builds", C.c_int64), ("t_end", C.c_double),
                ("wall_seconds", C.c_double), ("host_division_seconds", C.c_double)]


class EnsembleStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("attempts", C.c_int64), ("divisions", C.c_int64),
                ("neighbor_builds", C.c_int64), ("stage_cell_updates", C.c_int64),
                ("wall_seconds", C.c_double), ("host_division_seconds", C.c_double),
                ("autotune_ms", C.c_double), ("tuned_bin_scale", C.c_double),
                ("tuned_variant", C.c_int), ("n_final", C.c_int), ("failed", C.c_int),
                ("pad", C.c_int)]


def sim_config(n_cells=0, max_iter=0, max_time=0.0, R=0.8, Rcell=0.5, L0=0.4,
               max_stepsize=1e-4, min_stepsize=1e-9, max_error_allowed=1e-8, min_error=1e-30,
               growth_mean=(1.5, 1.5), growth_std=(0.15, 0.15), daughter_length_std=0.125,
               daughter_angle_xy_bound=0.05, daughter_angle_z_bound=0.0,
               iter_update_neighbors=10, iter_update_stepsize=1, max_tries_update_stepsize=3,
               tableau="heun_euler", seed=42, report_seconds=0.0, max_wall_seconds=0.0, M0=0.0):
    """Loop parameters named as the JSON keys of abmConstAdhesion.cpp:38-66:
    Ldiv = 2 L0 + 2 R and neighbor_threshold = (2 R + Ldiv) / 2; tableau
    "velocity_verlet" with M0 > 0 runs stepVelocityVerlet at dt = max_stepsize."""
    c = SimConfig()
    c.R, c.Rcell = R, Rcell
    c.Ldiv = 2 * L0 + 2 * R
    c.neighbor_threshold = 0.5 * (2 * R + c.Ldiv)
    c.max_stepsize, c.min_stepsize = max_stepsize, min_stepsize
    c.max_error_allowed, c.min_error = max_error_allowed, min_error
    c.growth_mean[0], c.growth_mean[1] = growth_mean
    c.growth_std[0], c.growth_std[1] = growth_std
    c.daughter_length_std = daughter_length_std
    c.daughter_angle_xy_bound = daughter_angle_xy_bound
    c.daughter_angle_z_bound = daughter_angle_z_bound
    c.max_time = max_time
    c.iter_update_neighbors = iter_update_neighbors
    c.iter_update_stepsize = iter_update_stepsize
    c.max_tries_update_stepsize = max_tries_update_stepsize
    c.n_cells, c.max_iter = n_cells, max_iter
    c.tableau = TABLEAU_IDS[tableau] if isinstance(tableau, str) else int(tableau)
    c.seed = seed
    c.report_seconds, c.max_wall_seconds = report_seconds, max_wall_seconds
    # velocity Verlet: density from the mass M0 of a cell of length L0 (simulation.hpp:1280-1290)
    c.rho = M0 / (np.pi * ((4. / 3.) * pow(R, 3) + R * R * L0)) if M0 > 0 else 0.0
    return c


def founder_cell(R=0.8, L0=0.4, E0=3900.0, sigma0=100.0, eta_ambient=0.7, eta_surface=720.0,
                 growth_mean=1.5, gamma=100.0, ncols=22):
    """The single starting cell of abmConstAdhesion.cpp:405-421: horizontal
    along x, resting at the height where surface repulsion balances adhesion."""
    cells = np.zeros((1, ncols), order="F")
    rz = R - pow(sigma0 * np.sqrt(R) / (4 * E0), 2. / 3.)
    cells[0, COL["rz"]] = rz
    cells[0, COL["nx"]] = 1.0
    cells[0, COL["l"]] = L0
    cells[0, COL["half_l"]] = L0 / 2
    cells[0, COL["growth"]] = growth_mean
    cells[0, COL["eta0"]] = eta_ambient
    cells[0, COL["eta1"]] = eta_surface
    cells[0, COL["sigma0"]] = sigma0
    cells[0, COL["group"]] = 1
    if ncols > 21:
        cells[0, COL["gamma"]] = gamma
    return cells


def simulate(ctx: Context, cfg: SimConfig, cells, cap=None, trace_cap=2_000_000):
    """Runs the loop on the GPU from `cells` (n x ncols) and returns
    (cells, parents, dt_trace, stats dict)."""
    lib = load_library()
    cells = np.asfortranarray(cells, dtype=np.float64)
    n_in, ncols = cells.shape
    cap = int(cap or ctx.max_cells)
    out = np.zeros((cap, ncols), order="F")
    parents = np.zeros(n_in + 2 * cap + 8, dtype=np.int32)
    n_out, n_par = C.c_int(), C.c_int()
    trace = np.zeros(trace_cap)
    st = SimStats()
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    rc = lib.bfm_simulate(ctx.h, C.byref(cfg), cells.ctypes.data_as(dp), n_in, ncols, None, 0,
                          out.ctypes.data_as(dp), cap, C.byref(n_out), parents.ctypes.data_as(ip),
                          C.byref(n_par), trace.ctypes.data_as(dp), trace_cap, C.byref(st))
    ctx._check(rc)
    stats = {k: getattr(st, k) for k, _ in SimStats._fields_}
    it = min(stats["iterations"], trace_cap)
    return (np.asfortranarray(out[:n_out.value]), parents[:n_par.value].copy(), trace[:it].copy(),
            stats)


def divide_cells(cfg: SimConfig, cells, t, seed, parents=None):
    """One divideMaxLength + divideCells call on host data (no device needed).
    Returns (cells, parents, tries)."""
    lib = load_library()
    cells = np.asarray(cells, dtype=np.float64)
    n, ncols = cells.shape
    cap = 2 * n + 1
    buf = np.zeros((cap, ncols), order="F")
    buf[:n] = cells
    if parents is None:
        parents = -np.ones(int(cells[:, COL["id"]].max()) + 1, dtype=np.int32)
    pcap = len(parents) + 2 * n + 2
    par = np.zeros(pcap, dtype=np.int32)
    par[:len(parents)] = parents
    n_io, n_par, tries = C.c_int(n), C.c_int(len(parents)), C.c_int()
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    rc = lib.bfm_divide_cells(C.byref(cfg), buf.ctypes.data_as(dp), C.byref(n_io), cap, ncols, t,
                              seed, par.ctypes.data_as(ip), C.byref(n_par), pcap, C.byref(tries))
    if rc != 0:
        raise BFMError(rc, "bfm_divide_cells failed")
    return np.asfortranarray(buf[:n_io.value]), par[:n_par.value].copy(), tries.value


def rng_draws_device(seed, count, device=0):
    """count uniform_01 draws, then count standard normals, of the division
    random stream as the device generator of the growth loop draws them."""
    lib = load_library()
    out = np.zeros(2 * count)
    rc = lib.bfm_rng_draws_device(device, seed, count, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise BFMError(rc, "bfm_rng_draws_device failed")
    return out[:count], out[count:]


def rng_uniform(seed, count):
    """The first `count` uniform_01 draws of the division random stream."""
    lib = load_library()
    out = np.zeros(count)
    rc = lib.bfm_rng_uniform(seed, count, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise BFMError(rc, "bfm_rng_uniform failed")
    return out


def jkr_equilibrium_distance(R, E0, gamma, dinit, dmin, increment_overlap=1e-8, newton_tol=1e-8,
                            max_iter=1000):
    """jkrEquilibriumDistance (jkr.hpp:1310-1340) with its default arguments:
    the cell-cell distance at which the JKR force vanishes."""
    lib = load_library()
    deq, res = C.c_double(), C.c_double()
    rc = lib.bfm_jkr_equilibrium_distance(R, E0, gamma, dinit, dmin, increment_overlap,
                                          newton_tol, max_iter, C.byref(deq), C.byref(res))
    if rc != 0:
        raise BFMError(rc, "bfm_jkr_equilibrium_distance failed")
    return deq.value


def jkr_optimal_surface_energy_density(R, Rcell, E0, deq_target, init_gamma=10.0,
                                       log_increment_gamma=1e-6, increment_overlap=1e-8,
                                       outer_newton_tol=1e-8, inner_newton_tol=1e-8,
                                       outer_max_iter=1000, inner_max_iter=1000):
    """jkrOptimalSurfaceEnergyDensity (jkr.hpp:1369-1408) with its default
    arguments: the surface energy density whose JKR equilibrium distance is
    deq_target."""
    lib = load_library()
    g, res = C.c_double(), C.c_double()
    rc = lib.bfm_jkr_optimal_surface_energy_density(
        R, Rcell, E0, deq_target, init_gamma, log_increment_gamma, increment_overlap,
        outer_newton_tol, inner_newton_tol, outer_max_iter, inner_max_iter, C.byref(g),
        C.byref(res))
    if rc != 0:
        raise BFMError(rc, "bfm_jkr_optimal_surface_energy_density failed")
    return g.value


def init_surface_energy(cells, eqdist, R=0.8, Rcell=0.5, E0=3900.0):
    """The adhesion setup of runSimulation (simulation.hpp:1352-1374): the
    surface energy density for the equilibrium distance `eqdist` (solved from
    init_gamma 100), written to the gamma column for group-1 cells and zero for
    the others. Returns the density."""
    gamma = jkr_optimal_surface_energy_density(R, Rcell, E0, eqdist, 100.0, 1e-6, 1e-8, 1e-8,
                                               1e-8, 1000, 1000)
    group1 = cells[:, COL["group"]] == 1
    cells[:, COL["gamma"]] = np.where(group1, gamma, 0.0)
    return gamma


def ensemble_simulate(ctx: Context, cfg: SimConfig, replicates, seeds, autotune=True, cap=None):
    """Runs every replicate's loop on the GPU in lockstep (bfm_ensemble_simulate).
    replicates: list of (n_r x ncols) cell arrays; seeds: one division stream
    seed per replicate. Returns (list of final cell arrays, t per replicate,
    iterations per replicate, stats dict with the failed replicates' indices)."""
    lib = load_library()
    n_rep = len(replicates)
    ncols = replicates[0].shape[1]
    sizes = np.array([r.shape[0] for r in replicates], dtype=np.int32)
    cells = np.asfortranarray(np.vstack(replicates), dtype=np.float64)
    cap = int(cap or ctx.max_cells)
    out = np.zeros((cap, ncols), order="F")
    sizes_out = np.zeros(n_rep, dtype=np.int32)
    t_out = np.zeros(n_rep)
    it_out = np.zeros(n_rep, dtype=np.int64)
    failed = np.zeros(n_rep, dtype=np.int32)
    seeds_a = np.ascontiguousarray(seeds, dtype=np.uint64)
    st = EnsembleStats()
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    rc = lib.bfm_ensemble_simulate(ctx.h, C.byref(cfg), n_rep, sizes.ctypes.data_as(ip),
                                   cells.ctypes.data_as(dp), ncols,
                                   seeds_a.ctypes.data_as(C.POINTER(C.c_uint64)), int(bool(autotune)),
                                   out.ctypes.data_as(dp), cap, sizes_out.ctypes.data_as(ip),
                                   t_out.ctypes.data_as(dp),
                                   it_out.ctypes.data_as(C.POINTER(C.c_int64)),
                                   failed.ctypes.data_as(ip), C.byref(st))
    ctx._check(rc)
    stats = {k: getattr(st, k) for k, _ in EnsembleStats._fields_}
    stats["failed_replicates"] = np.nonzero(failed)[0].tolist()
    res, o = [], 0
    for r in range(n_rep):
        res.append(np.asfortranarray(out[o:o + sizes_out[r]]))
        o += sizes_out[r]
    return res, t_out, it_out, stats
