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


def make_ensemble(sizes, seeds=None, gammas=(50.0, 500.0), phenotype_seed=1234, **dome_kw):
    """One dome snapshot per replicate (replicate r: make_dome(sizes[r], seed=seeds[r])),
    each replicate assigned one of two phenotypes (surface energy gammas[0] or
    gammas[1], Pa um) by a seeded 50/50 draw. Returns (list of cell arrays,
    phenotype per replicate)."""
    n_rep = len(sizes)
    seeds = list(range(n_rep)) if seeds is None else list(seeds)
    pheno = np.random.default_rng(phenotype_seed).permutation(np.arange(n_rep) % 2)
    reps = []
    for r in range(n_rep):
        cells, _ = make_dome(int(sizes[r]), seed=seeds[r], gamma=gammas[int(pheno[r])], **dome_kw)
        reps.append(cells)
    return reps, pheno
