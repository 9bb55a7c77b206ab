"""Two-cell relaxation scenarios of test/src/testTwoCellSimulations.cpp
(ABM-3d/ABM-3d-cpp, lines 36-308): two cells of length 1 (half length 0.5)
lifted off the surface (z = 1.1 R, no surface forces, no adhesion, no
growth), started at a centerline distance of 1.0, 1.4 or 1.8 um in nine
configurations, run for 10,000 Heun-Euler iterations with dt <= 1e-7 h.

Hertz repulsion acts only at overlap (distance < 2R = 1.6), so the pairs
started at 1.0 and 1.4 are pushed apart to 2R and the pair at 1.8 never
moves. The reference test prints the equilibrium distance next to this
expectation (testTwoCellSimulations.cpp:114-133); the tests here assert it.
With JKR adhesion at the surface energy density of a target equilibrium
distance (jkrOptimalSurfaceEnergyDensity), the overlapping pairs end at that
distance instead.

Cell columns: the reference's comma initializer (testTwoCellSimulations.cpp:
89-96) lists 22 values per row for its 21-column array, an older layout with
eta_surface twice; the values meant by name are used here: eta0 = 1e-3,
eta1 = 720, sigma0 = 100, group 1."""
import numpy as np

R = 0.8
HALF_L = 0.5
Z = 1.1 * R
DISTS = (1.0, 1.4, 1.8)
_C6, _S6 = np.cos(np.pi / 6), np.sin(np.pi / 6)

# second cell (center, orientation) for initial distance d; the first cell
# sits at (0, 0, 1.1 R) along x (testTwoCellSimulations.cpp:139-308)
CONFIGS = {
    "head_to_head": lambda d: ((2 * HALF_L + d, 0.0, Z), (1.0, 0.0, 0.0)),
    "side_by_side": lambda d: ((0.0, d, Z), (1.0, 0.0, 0.0)),
    "shifted_side_by_side": lambda d: ((0.2, d, Z), (1.0, 0.0, 0.0)),
    "T": lambda d: ((HALF_L + d, 0.0, Z), (0.0, 1.0, 0.0)),
    "shifted_T": lambda d: ((HALF_L + d, 0.2, Z), (0.0, 1.0, 0.0)),
    "cross": lambda d: ((0.0, 0.0, Z + d), (0.0, 1.0, 0.0)),
    "skew_end_to_end": lambda d: ((HALF_L + d + HALF_L * _C6, HALF_L * _S6, Z), (_C6, _S6, 0.0)),
    "skew_side_by_side": lambda d: ((HALF_L * _C6, d + HALF_L * _S6, Z), (_C6, _S6, 0.0)),
    "skew_cross": lambda d: ((0.2, 0.0, Z + d), (_C6, _S6, 0.0)),
}

CASES = [(name, d) for name in CONFIGS for d in DISTS]


def cells(name, d, gamma=0.0, ncols=22):
    """The two cells of a case; gamma fills the surface energy column (21)."""
    r2, n2 = CONFIGS[name](d)
    c = np.zeros((2, ncols), order="F")
    for k, (r, n) in enumerate((((0.0, 0.0, Z), (1.0, 0.0, 0.0)), (r2, n2))):
        c[k, 0] = k
        c[k, 1:4] = r
        c[k, 4:7] = n
        c[k, 13], c[k, 14] = 2 * HALF_L, HALF_L
        c[k, 17], c[k, 18], c[k, 19], c[k, 20] = 1e-3, 720.0, 100.0, 1
        if ncols > 21:
            c[k, 21] = gamma
    return c


def config():
    """Loop settings of testTwoCellSimulations.cpp:42-66."""
    from bfm import sim
    c = sim.sim_config(max_iter=10000, R=R, max_stepsize=1e-7, min_stepsize=1e-9,
                       max_error_allowed=1e-10, min_error=1e-30, iter_update_neighbors=100,
                       iter_update_stepsize=1, max_tries_update_stepsize=3,
                       growth_mean=(0.0, 0.0), growth_std=(0.0, 0.0), daughter_length_std=0.0,
                       daughter_angle_xy_bound=0.0, daughter_angle_z_bound=0.0,
                       tableau="heun_euler", seed=42)
    c.Ldiv = 2 * (R + 2 * HALF_L)
    c.neighbor_threshold = 2 * (2 * R + 2 * HALF_L)
    return c


def params(adhesion=False):
    """No surface, Hertz only (the reference scenario), or with isotropic JKR
    adhesion between the cells (exact contact radius)."""
    from bfm.core import make_params
    return make_params(R=R, nz_threshold=1e-20, cell_surface_coulomb_coeff=0.0,
                       no_surface=True, adhesion=adhesion)


def distance(orc, c):
    dd, _, _ = orc.dist_between_cells(c[0, 1:4], c[0, 4:7], c[0, 14], c[1, 1:4], c[1, 4:7], c[1, 14])
    return float(np.linalg.norm(dd))
