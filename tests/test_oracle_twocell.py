"""The oracle's growth loop (runSimulationAdaptiveLagrangian restated) on the
two-cell relaxation scenarios of testTwoCellSimulations.cpp, against what
the mechanics fixes: overlapping pairs end at contact distance 2R, a pair
out of contact never moves, and with equal cells, no surface and no noise
the forces are equal and opposite under equal isotropic drag, so the
midpoint of the two centers stays where it started."""
import numpy as np
import pytest

import twocell_cases as tc

TOL_CONTACT = 5e-7   # um; the Hertz force ~ overlap^1.5 leaves a ~1e-7 tail at t = 1e-3 h


@pytest.mark.parametrize("name,d", tc.CASES)
def test_two_cell_equilibrium(orc, name, d):
    c0 = tc.cells(name, d)
    assert abs(tc.distance(orc, c0) - d) < 1e-12
    out, trace, iters, steps = orc.simulate(tc.config(), tc.params(), c0, cap=8)
    assert iters == 10000 and steps >= 2 * iters
    assert np.all(trace[:iters] <= 1e-7 * (1 + 1e-12))
    final = tc.distance(orc, out[:2])
    if d < 2 * tc.R:
        assert abs(final - 2 * tc.R) < TOL_CONTACT, final
    else:
        assert np.array_equal(out[:2, :22], c0)
    mid0 = 0.5 * (c0[0, 1:4] + c0[1, 1:4])
    mid1 = 0.5 * (out[0, 1:4] + out[1, 1:4])
    assert np.max(np.abs(mid1 - mid0)) < 1e-12
    assert np.allclose(np.linalg.norm(out[:2, 4:7], axis=1), 1.0, atol=1e-12)
    # the cells stay in their initial planes when both lie in z = 1.1 R
    if c0[1, 3] == c0[0, 3]:
        assert np.all(out[:2, 3] == c0[0, 3]) and np.all(out[:2, 6] == 0.0)


EQDIST = 1.5


@pytest.mark.parametrize("name,d", tc.CASES)
def test_two_cell_adhesion_equilibrium(orc, name, d):
    """With JKR adhesion at the surface energy density of a 1.5 um equilibrium
    distance (jkrOptimalSurfaceEnergyDensity), overlapping pairs end at 1.5 um
    (the density's own Newton residual is ~1e-10); the pair at 1.8 um has no
    overlap and no force."""
    gamma = orc.jkr_optimal_surface_energy_density(tc.R, 0.5, 3900.0, EQDIST, 100.0)
    deq = orc.jkr_equilibrium_distance(tc.R, 3900.0, gamma, 1.5 * tc.R, 1.0)
    assert abs(deq - EQDIST) < 1e-8
    c0 = tc.cells(name, d, gamma=gamma)
    out, trace, iters, steps = orc.simulate(tc.config(), tc.params(adhesion=True), c0, cap=8)
    assert iters == 10000
    final = tc.distance(orc, out[:2])
    if d < 2 * tc.R:
        assert abs(final - deq) < 1e-8, final
    else:
        assert np.array_equal(out[:2, :22], c0)
    mid0 = 0.5 * (c0[0, 1:4] + c0[1, 1:4])
    mid1 = 0.5 * (out[0, 1:4] + out[1, 1:4])
    assert np.max(np.abs(mid1 - mid0)) < 1e-12
