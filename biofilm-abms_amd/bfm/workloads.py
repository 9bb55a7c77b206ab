"""Seeded synthetic inputs of the benchmark configurations (libbfm_workloads.so)."""
import ctypes as C
import os

import numpy as np

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "libbfm_workloads.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
        lib = C.CDLL(path)
        d = C.c_double
        lib.bfm_make_dome.argtypes = [C.c_int, C.c_int, d, d, d, d, d, d, d, d, d, d, C.c_uint64,
                                      C.POINTER(d), C.POINTER(d)]
        lib.bfm_make_dome.restype = C.c_int
        _LIB = lib
    return _LIB


def make_dome(n, seed=42, R=0.8, Rcell=0.5, phi=0.5, hl_lo=0.4, hl_hi=1.6, gamma=100.0,
              eta0=0.7, eta1=720.0, sigma0=100.0, growth=1.5, ncols=22):
    """Config-1 snapshot: n spherocylinders in a hemispherical dome at packing fraction phi.

    Returns (cells n x ncols Fortran-ordered, dome radius)."""
    cells = np.zeros((n, ncols), order="F")
    rd = C.c_double()
    got = _lib().bfm_make_dome(n, ncols, R, Rcell, phi, hl_lo, hl_hi, gamma, eta0, eta1, sigma0,
                               growth, seed, cells.ctypes.data_as(C.POINTER(C.c_double)),
                               C.byref(rd))
    if got != n:
        raise RuntimeError(f"dome generator placed only {got} of {n} cells")
    return cells, rd.value


def ensemble_sizes(n_rep=256, lo=500, hi=20000, seed=0):
    """Replicate cell counts round(exp(U[ln lo, ln hi])) from one seeded stream."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n_rep))).astype(np.int64)


def assign_phenotypes(cells, gammas=(500.0, 50.0), seed=1234):
    """Two phenotypes at a seeded 50/50 assignment per cell (half of the cells,
    rounded down, in group 1): the group column is 1 or 2 and the surface
    energy density column follows the group, gamma = gammas[group - 1]
    (the reference initialises gamma per cell from its group,
    simulation.hpp:1368-1374; a pair adheres with min(gamma_i, gamma_j),
    mechanics.hpp:1464). Daughters inherit both columns. In place; returns
    the group per cell."""
    n = cells.shape[0]
    group = np.random.default_rng(seed).permutation(np.arange(n) % 2) + 1
    cells[:, 20] = group
    cells[:, 21] = np.asarray(gammas, dtype=np.float64)[group - 1]
    return group


def make_ensemble(sizes, seeds=None, gammas=(500.0, 50.0), phenotype_seed=1234, **dome_kw):
    """One dome snapshot per replicate (replicate r: make_dome(sizes[r], seed=seeds[r]))
    whose cells carry one of two phenotypes (surface energy gammas[0] for group
    1, gammas[1] for group 2, Pa um) by a seeded 50/50 draw per cell
    (assign_phenotypes with seed phenotype_seed + r). Returns (list of cell
    arrays, list of group vectors)."""
    n_rep = len(sizes)
    seeds = list(range(n_rep)) if seeds is None else list(seeds)
    reps, groups = [], []
    for r in range(n_rep):
        cells, _ = make_dome(int(sizes[r]), seed=seeds[r], **dome_kw)
        groups.append(assign_phenotypes(cells, gammas, phenotype_seed + r))
        reps.append(cells)
    return reps, groups
