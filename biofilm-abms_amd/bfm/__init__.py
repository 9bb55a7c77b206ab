"""Python host side of the MI355X mechanical step (ctypes over libbfm.so).

The functions named like the reference (getCellNeighbors, updateNeighborDistances,
getConservativeForces, getVelocities, stepRungeKuttaAdaptive) keep the argument
meaning and array layouts of ABM-3d/ABM-3d-cpp/include/mechanics.hpp; arrays are
numpy float64 in column-major (Fortran) order like the reference's Eigen arrays.
Every call runs on the GPU through the C ABI of include/bfm.h: there is no CPU
fallback, and importing the library fails loudly when it is missing.
"""
from .core import (  # noqa: F401
    BFMError, Context, Params, load_library, library_path, tableau, TABLEAUX,
    getCellNeighbors, updateNeighborDistances, getConservativeForces, getVelocities,
    stepRungeKuttaAdaptive, stepVelocityVerlet, repulsion_prefactors, COL,
)
from .workloads import make_dome, make_ensemble, ensemble_sizes  # noqa: F401
